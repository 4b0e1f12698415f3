"""QRunes front end (qpanda_amd/cpp/QRunes.cpp), host code, no device needed.

Pins:
  * QPanda's own parser tests (testMain/testQRunesParser.cpp:9-35): an empty
    path and a missing file fail, shor_15.txt parses, testfile.txt does not
    ('X' is not a token of the lexer, so "X 0" leaves a bare integer);
  * the SDK's loadFile codes (testMain/testQPandaAPI.cpp:10-31) and run before
    loading (:67-85), through the handle-based facade;
  * the grammar's semantics, each case derived from the grammar actions
    (QRunesParser/y.tab.cpp:1290-1650) and the list controller
    (LinkedListController.cpp:51-339) by hand: DAGGER reversal and daggered
    flags by nesting parity, CONTROL lists, dropped measurements, the angle
    arithmetic including "a - b" = a + b and left-associative '^'.
tests/test_qrunes_tables.py compares the same parser, accept / reject and
instruction list, with a walk of the reference lexer's and parser's own
tables."""
import math
import os

import pytest

from qpanda_amd import sdk

G = os.path.join(os.path.dirname(__file__), "golden", "qrunes")


def P(*lines):
    return sdk.parse_text("\n".join(lines))


def brief(nodes):
    return [(n["kind"], tuple(n["qubits"]), n["dagger"], tuple(n["controls"])) for n in nodes]


def test_reference_parser_expectations():
    assert sdk.parse_file("")[0] == -1
    assert sdk.parse_file(os.path.join(G, "hahah"))[0] == -1
    q, nodes = sdk.parse_file(os.path.join(G, "shor_15.txt"))
    assert q == 7
    assert [(n["kind"], n["qubits"]) for n in nodes] == [("H", [0]), ("RX", [2]), ("MEASURE", [0])]
    assert nodes[1]["angle"] == 3.14159265358979323846 and nodes[2]["creg"] == 0
    assert sdk.parse_file(os.path.join(G, "testfile.txt"))[0] == -1


def test_nested_dagger_and_control_example():
    q, nodes = sdk.parse_file(os.path.join(G, "parser_test.txt"))
    assert q == 15
    # CNOT 0,2 | CONTROL 4 | CONTROL 5 | H 1 | DAGGER | H 2 | DAGGER | H 3 | ENDDAGGER | H 4 |
    # ENDCONTROL 4 | ENDCONTROL 5 | H 5 | ENDDAGGER | H 6:
    # the outer block [H2, H3, H4, H5] comes out reversed and daggered; the inner
    # block (dagger of dagger) is not daggered; controls are those in effect
    assert brief(nodes) == [
        ("CNOT", (0, 2), False, ()),
        ("H", (1,), False, (4, 5)),
        ("H", (5,), True, ()),
        ("H", (4,), True, (4, 5)),
        ("H", (3,), False, (4, 5)),
        ("H", (2,), True, (4, 5)),
        ("H", (6,), False, ()),
    ]


def test_pmeasure_example():
    q, nodes = sdk.parse_file(os.path.join(G, "pmeasure_test.txt"))
    assert q == 7
    assert [(n["kind"], n["qubits"]) for n in nodes] == [("H", [0]), ("H", [1]), ("NOT", [2]),
                                                          ("PMEASURE", [0, 1, 2, 3])]


@pytest.mark.parametrize("expr,value", [
    ("pi", math.pi), ("PI", math.pi), ("Pi/2", math.pi / 2), ("-pi", -math.pi), ("+3", 3.0),
    ("1-2", 3.0),                     # the '-' action adds
    ("2^3^2", 64.0),                  # left-associative power
    ("-2^2", 4.0),                    # unary minus binds tighter than '^'
    ("(1+2)*3", 9.0), ("3/2/4", 0.375), ("2*3^2", 18.0), ("1.5*pi", 1.5 * math.pi), ("-(-1)", 1.0),
])
def test_angle_expressions(expr, value):
    q, nodes = P("QINIT 1", "CREG 1", f'RX 0,"{expr}"')
    assert q == 1 and nodes[0]["angle"] == value


@pytest.mark.parametrize("expr", ["0/1", "pi/0", "0^2", '1e3', "", "pi pi", "1+"])
def test_bad_angles(expr):
    # quotients and powers that are not normal doubles, and syntax errors
    assert P("QINIT 1", "CREG 1", f'RX 0,"{expr}"')[0] == -1


def test_lexer_rules():
    # unknown characters are skipped; keywords are case sensitive; '%' comments
    q, nodes = P("% a comment with H 0", "QINIT 2", "CREG 1", 'RXq 0,"pi"', "HADAMARD 1", "CPHASE 0,1,\"pi\"")
    assert q == 2 and brief(nodes) == [("RX", (0,), False, ()), ("H", (1,), False, ()), ("CR", (0, 1), False, ())]
    assert P("QINIT 2", "CREG 1", "h 0")[0] == -1              # 'h' is skipped, leaving a bare integer
    assert P("QINIT 2\r\nCREG 1\r\nH 0\r\n")[0] == 2


def test_structure_errors():
    assert P("QINIT 2", "CREG 1")[0] == -1                     # at least one statement
    assert P("CREG 1", "QINIT 2", "H 0")[0] == -1              # QINIT first
    assert P("QINIT 2", "CREG 1", "H 2")[0] == -1              # qubit out of bounds
    assert P("QINIT 2", "CREG 1", "MEASURE 0,$1")[0] == -1     # register out of bounds
    assert P("QINIT 2", "CREG 1", "MEASURE 0,0")[0] == -1      # '$' required
    assert P("QINIT 2", "CREG 1", "RX 0,pi")[0] == -1          # quotes required
    assert P("QINIT 2", "CREG 1", "H 0", "PMEASURE 0", "H 1")[0] == -1   # PMEASURE only at the end
    assert P("QINIT 2", "CREG 1", "ISWAP 0,1")[0] == -1        # ISWAP takes an angle in this grammar
    assert P("QINIT 2", "CREG 1", "PMEASURE 0")[0] == -1       # a statement before PMEASURE


def test_controller_rules():
    # measurement inside DAGGER or CONTROL is dropped; an unclosed DAGGER drops
    # its statements; stray closers and repeated CONTROL are ignored
    q, nodes = P("QINIT 3", "CREG 2", "DAGGER", "MEASURE 0,$0", "H 0", "ENDDAGGER",
                 "CONTROL 1", "MEASURE 1,$1", "CONTROL 1", "NOT 2", "ENDCONTROL 1", "ENDCONTROL 1",
                 "ENDDAGGER", "ENDCONTROL 0", "H 1", "DAGGER", "H 2")
    assert brief(nodes) == [("H", (0,), True, ()), ("NOT", (2,), False, (1,)), ("H", (1,), False, ())]


def test_triple_nested_dagger_keeps_the_reference_order():
    # levels 2, 3, 4: [H2, H3] (level 4, even: not reversed on closing) joins
    # level 3 as [H1, H2, H3], which is reversed (odd) into level 2 as
    # [H0, H3, H2, H1], reversed again at the root: the innermost statements
    # are flagged daggered but keep program order
    q, nodes = P("QINIT 4", "CREG 1", "DAGGER", "H 0", "DAGGER", "H 1", "DAGGER", "H 2", "H 3",
                 "ENDDAGGER", "ENDDAGGER", "ENDDAGGER")
    assert brief(nodes) == [("H", (1,), False, ()), ("H", (2,), True, ()), ("H", (3,), True, ()),
                            ("H", (0,), True, ())]


def test_sdk_codes_without_device():
    with sdk.SDK() as s:
        assert s.load_file("") == sdk.Q_PARAMETER_ERROR
        assert s.load_file(os.path.join(G, "hahah")) == sdk.QBIT_ERROR
        assert s.run(1) == sdk.NOT_READ_FILE_ERROR
        assert s.load_file(os.path.join(G, "shor_15.txt")) == sdk.Q_ERROR_NONE
        assert s.run(-1) == sdk.Q_PARAMETER_ERROR
        assert s.run(1) == sdk.INIT_STATE_ERROR               # no compute unit
        assert s.set_compute_unit(0) == sdk.Q_PARAMETER_ERROR
        assert s.set_compute_unit(3) == sdk.Q_PARAMETER_ERROR
        assert s.set_compute_unit(sdk.CPU) == sdk.SET_COMPUTE_UNIT_ERROR
        assert s.result() == "" and s.state_text() == ""
