"""The QRunes parser (qpanda_amd/cpp/QRunes.cpp) against QPanda's own generated
parser.  tests/golden/qrunes_parse.json holds seeded programs and the result of
a walk of the reference's flex and bison tables, with its grammar actions and
its list controller (oracle/qrunes_tables.py; tests/golden/
gen_qrunes_parse_golden.py wrote the file).

Bar: accept / reject identical; on accept the qubit count and every
instruction identical: kind, qubits, angle bit for bit, daggered flag, control
list, classical bit.

The fixture test needs nothing but the file.  The other tests read the tables
from the reference tree and are skipped where it is absent: the walk itself
against the reference's own parser tests and example files, mistakes planted
in the walk that the fixture must notice, fresh programs, and the inputs of
the executor fixtures (tests/golden/qrunes_exec.json), so that those are
QPanda's own parses."""
import importlib.util
import json
import math
import os
import re

import pytest

import qrunes_tables as Q
from qpanda_amd import sdk

G = os.path.join(os.path.dirname(__file__), "golden")
need_ref = pytest.mark.skipif(not Q.available(), reason="needs the QPanda source tree for its parser tables")


def fixture():
    with open(os.path.join(G, "qrunes_parse.json")) as f:
        return json.load(f)["cases"]


def as_rows(nodes):
    return [[n["kind"], n["qubits"], n["angle"], int(n["dagger"]), n["controls"], n["creg"]] for n in nodes]


def same(q, rows, case):
    if q != case["qubits"] or len(rows) != len(case["nodes"]):
        return False
    for a, b in zip(rows, case["nodes"]):
        if math.isnan(a[2]) and math.isnan(b[2]):
            a, b = a[:2] + a[3:], b[:2] + b[3:]
        if a != b:
            return False
    return True


def dagger_depth(text):
    depth = top = 0
    for m in re.finditer(r"ENDDAGGER|DAGGER", text):
        depth = max(0, depth - 1) if m.group(0) == "ENDDAGGER" else depth + 1
        top = max(top, depth)
    return top


def generator():
    spec = importlib.util.spec_from_file_location("gen_qrunes_parse_golden",
                                                  os.path.join(G, "gen_qrunes_parse_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_engine_matches_reference_parser_fixture():
    bad = []
    for i, c in enumerate(fixture()):
        q, nodes = sdk.parse_text(c["text"])
        if not same(q, as_rows(nodes), c):
            bad.append(i)
    assert not bad, "cases %s differ, first: %r" % (bad[:10], fixture()[bad[0]]["text"])


def test_fixture_coverage():
    cs = fixture()
    acc = [c for c in cs if c["qubits"] >= 0]
    rows = [r for c in acc for r in c["nodes"]]
    assert 0.3 * len(cs) < len(acc) < 0.7 * len(cs)
    assert {r[0] for r in rows} == {"H", "NOT", "RESET", "RX", "RY", "RZ", "CNOT", "SQISWAP", "CR", "ISWAP",
                                    "TOFFOLI", "MEASURE", "PMEASURE"}
    assert sum(r[3] for r in rows) > 200                      # daggered
    assert sum(1 for r in rows if len(r[4]) >= 2) > 50      # two or more controls
    texts = [c["text"] for c in acc]
    assert sum(dagger_depth(t) >= 2 for t in texts) > 20 and sum(dagger_depth(t) >= 3 for t in texts) > 3
    for token in ("pi", "PI", "Pi", "pI", "%", "\t", "\r\n", "^", "/", "*", "(", "-"):
        assert sum(token in t for t in texts) > 10, token
    assert any(ord(ch) > 127 for t in texts for ch in t)


@need_ref
def test_walk_reproduces_the_fixture():
    bad = [i for i, c in enumerate(fixture()) if not same(*(lambda r: (r[0], as_rows(r[1])))(Q.parse(c["text"])), c)]
    assert not bad


@need_ref
def test_walk_matches_reference_parser_tests():
    # testMain/testQRunesParser.cpp:9-35 and the reference's example programs
    def walk(name):
        with open(os.path.join(G, "qrunes", name), "rb") as f:
            return Q.parse(f.read())
    q, nodes = walk("shor_15.txt")
    assert q == 7 and [(n["kind"], n["qubits"]) for n in nodes] == [("H", [0]), ("RX", [2]), ("MEASURE", [0])]
    assert nodes[1]["angle"] == 3.14159265358979323846
    assert walk("testfile.txt") == (-1, [])
    q, nodes = walk("parser_test.txt")
    assert q == 15
    assert [(n["kind"], tuple(n["qubits"]), n["dagger"], tuple(n["controls"])) for n in nodes] == [
        ("CNOT", (0, 2), False, ()), ("H", (1,), False, (4, 5)), ("H", (5,), True, ()),
        ("H", (4,), True, (4, 5)), ("H", (3,), False, (4, 5)), ("H", (2,), True, (4, 5)), ("H", (6,), False, ())]
    # the lexer's own classes: "pi" in any case, no exponent in numbers, '.' alone is no token
    toks = Q.scan(b'PI pI 1.5e3 .5 007 X', Q.tables())
    F, I = Q.tables().tokens["FLOAT"], Q.tables().tokens["INT"]
    assert toks == [(F, math.pi), (F, math.pi), (F, 1.5), (I, 3), (I, 5), (I, 7)]


@need_ref
@pytest.mark.parametrize("plant", ["sub", "no_isnormal", "dagger_parity", "no_reverse", "keep_block_measure",
                                   "dup_control"])
def test_fixture_notices_planted_mistakes(plant):
    differ = 0
    for c in fixture():
        q, nodes = Q.parse(c["text"], plant=(plant,))
        differ += not same(q, as_rows(nodes), c)
    assert differ >= 20, (plant, differ)


@need_ref
def test_fresh_programs_match_reference_parser():
    bad = []
    for text in generator().programs(777, 6000):
        q, nodes = Q.parse(text.encode("latin-1"))
        want = {"qubits": q, "nodes": as_rows(nodes)}
        gq, gnodes = sdk.parse_text(text)
        if not same(gq, as_rows(gnodes), want):
            bad.append(text)
    assert not bad, bad[:3]


@need_ref
def test_executor_fixture_inputs_are_reference_parses():
    with open(os.path.join(G, "qrunes_exec.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        q, nodes = Q.parse(c["text"].encode("latin-1"))
        assert q >= 0, c["name"]
        assert sdk.parse_text(c["text"]) == (q, nodes), c["name"]


WRAP_CASES = [
    # integers go through atoi, (int) strtol: the low 32 bits, as a signed value
    ("QINIT 4 CREG 1 H 4294967296", 4, [("H", [0], 0.0)]),
    ("QINIT 4294967300 CREG 1 H 1", 4, [("H", [1], 0.0)]),
    ("QINIT 4294967300 CREG 1 H 5", -1, []),
    ('QINIT 4 CREG 1 RX 0,"4294967296+2147483648"', 4, [("RX", [0], -2147483648.0)]),
]


def test_integers_wrap_as_atoi_reads_them():
    for text, q, nodes in WRAP_CASES:
        got = sdk.parse_text(text)
        assert (got[0], [(n["kind"], n["qubits"], n["angle"]) for n in got[1]]) == (q, nodes), text
    # deliberate: an index that wraps negative is refused (the reference would
    # use it as an index)
    assert sdk.parse_text("QINIT 4 CREG 1 H 2147483648") == (-1, [])
    assert sdk.parse_text("QINIT 4 CREG 1 H 99999999999999999999999") == (-1, [])


@need_ref
def test_walk_wraps_integers_the_same_way():
    for text, q, nodes in WRAP_CASES:
        got = Q.parse(text)
        assert (got[0], [(n["kind"], n["qubits"], n["angle"]) for n in got[1]]) == (q, nodes), text
    assert Q.parse("QINIT 4 CREG 1 H 2147483648")[1][0]["qubits"] == [-2 ** 31]
