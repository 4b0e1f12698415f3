"""Writes tests/golden/qrunes_parse.json: seeded QRunes programs and what
QPanda's own parser makes of them (qubit count or -1, and the instruction
list), from a walk of the reference's generated scanner and parser tables
(oracle/qrunes_tables.py reads them from the reference tree).

The programs cover the lexer (every keyword, "pi" in any case, numbers with
and without a fraction, characters no rule names, comments, missing blanks),
the grammar (headers, gate lines of every arity, angle expressions with
unary signs, parentheses, '^' chains and divisions, PMEASURE lists) and the
list controller (nested DAGGER / CONTROL blocks, unbalanced closers,
measurements inside blocks, out-of-range qubits and classical bits); a part of
them is then mutated byte by byte.  Integers stay below 2^31 (the reference
reads them with atoi).

Run from the repository root:  python tests/golden/gen_qrunes_parse_golden.py
"""
import json
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "qrunes_parse.json")

SINGLE = ["H", "HADAMARD", "NOT", "RESET"]
SINGLE_ANGLE = ["RX", "RY", "RZ"]
DOUBLE = ["CNOT", "SQISWAP"]
DOUBLE_ANGLE = ["CR", "CPHASE", "ISWAP"]
TRIPLE = ["TOFFOLI"]
WORDS = (SINGLE + SINGLE_ANGLE + DOUBLE + DOUBLE_ANGLE + TRIPLE +
         ["QINIT", "CREG", "DAGGER", "ENDDAGGER", "CONTROL", "ENDCONTROL", "MEASURE", "PMEASURE", "pi", "PI"])
NOISE = list("+-*/^()\"$,.%%0123456789 \t\n\rXxYyZzABCDEFGHIJKLMNOPQRSTUVWabcdeipq@#!?;:_") + ["\xe9", "\xff", "\r\n"]


def _ws(r):
    return r.choice([" ", " ", " ", "  ", "\t", "", " \t "])


def _sep(r):
    return r.choice(["\n", "\n", "\n", "\r\n", " ", "\n\n", "  % note CNOT 0,1\n", "\t"])


def _num(r):
    k = r.random()
    if k < 0.45:
        return str(r.randrange(1, 10))
    if k < 0.48:
        return "0"
    if k < 0.6:
        return str(r.randrange(1, 1000))
    if k < 0.85:
        return "%d.%s" % (r.randrange(0, 100), "".join(r.choice("0123456789") for _ in range(r.randrange(1, 6))))
    return r.choice(["pi", "PI", "Pi", "pI", "007", "0.0", "1.0", "3.5"])


def _angle(r, depth=0):
    """A primary / exp / multiplicative / additive expression."""
    def prim(d):
        k = r.random()
        if d < 3 and k < 0.15:
            return "(" + _ws(r) + _angle(r, d + 1) + _ws(r) + ")"
        if k < 0.25:
            return r.choice("-+") + _ws(r) + prim(d + 1)
        return _num(r)

    def exp(d):
        e = prim(d)
        for _ in range(r.choice([0, 0, 0, 0, 1, 2])):
            e += _ws(r) + "^" + _ws(r) + (prim(d + 1) if r.random() < 0.2 else
                                          r.choice(["2", "0.5", "1.5", "3", "1", "pi", "-1", "(1/3)"]))
        return e

    def mul(d):
        e = exp(d)
        for _ in range(r.choice([0, 0, 0, 1, 2])):
            e += _ws(r) + r.choice("*/") + _ws(r) + exp(d + 1)
        return e

    if depth == 0 and r.random() < 0.5:
        return r.choice([_num(r), "pi/" + str(r.randrange(1, 9)), "-" + _num(r), _num(r) + "*pi"])
    e = mul(depth)
    for _ in range(r.choice([0, 0, 1, 1, 2])):
        e += _ws(r) + r.choice("+-") + _ws(r) + mul(depth + 1)
    return e


def _q(r, nq):
    return r.randrange(0, nq + 1) if r.random() < 0.01 else r.randrange(0, max(1, nq))


def _gate(r, nq):
    k = r.random()
    c = ","
    if k < 0.3:
        return r.choice(SINGLE) + _ws(r) + str(_q(r, nq))
    if k < 0.5:
        return r.choice(SINGLE_ANGLE) + _ws(r) + str(_q(r, nq)) + _ws(r) + c + _ws(r) + '"' + _angle(r) + '"'
    if k < 0.65:
        a, b = _q(r, nq), _q(r, nq)
        return r.choice(DOUBLE) + _ws(r) + "%d%s%s%s%d" % (a, _ws(r), c, _ws(r), b)
    if k < 0.85:
        a, b = _q(r, nq), _q(r, nq)
        return (r.choice(DOUBLE_ANGLE) + _ws(r) + "%d%s,%s%d" % (a, _ws(r), _ws(r), b) + _ws(r) + "," + _ws(r) +
                '"' + _angle(r) + '"')
    a, b, t = _q(r, nq), _q(r, nq), _q(r, nq)
    return r.choice(TRIPLE) + _ws(r) + "%d,%d%s,%s%d" % (a, b, _ws(r), _ws(r), t)


def _body(r, nq, nc, lines, depth=0):
    out = []
    for _ in range(lines):
        k = r.random()
        if k < 0.55 or depth > 3:
            out.append(_gate(r, nq))
        elif k < 0.68:
            inner = _body(r, nq, nc, r.randrange(0, 5), depth + 1)
            closed = r.random() < 0.93
            out += ["DAGGER"] + inner + (["ENDDAGGER"] if closed else [])
        elif k < 0.8:
            q = _q(r, nq)
            inner = _body(r, nq, nc, r.randrange(0, 5), depth + 1)
            closer = q if r.random() < 0.9 else _q(r, nq)
            out += ["CONTROL" + _ws(r) + " " + str(q)] + inner + (["ENDCONTROL %d" % closer] if r.random() < 0.93 else [])
        elif k < 0.9:
            cr = r.randrange(0, nc + 1) if r.random() < 0.03 else r.randrange(0, max(1, nc))
            out.append("MEASURE" + _ws(r) + " %d%s,%s$%d" % (_q(r, nq), _ws(r), _ws(r), cr))
        elif k < 0.94:
            out.append(r.choice(["ENDDAGGER", "ENDCONTROL %d" % _q(r, nq), "CONTROL %d" % _q(r, nq)]))
        else:
            out.append("% " + r.choice(WORDS) + " " + str(r.randrange(9)))
    return out


def program(r):
    nq = r.choice([0, 1, 2, 3, 4, 4, 5, 5, 6, 6, 8, 10, 16, 30])
    nc = r.choice([0, 1, 2, 4, 4, 8])
    lines = []
    if r.random() < 0.97:
        lines.append("QINIT" + _ws(r) + " " + str(nq))
    if r.random() < 0.95:
        lines.append("CREG" + _ws(r) + " " + str(nc))
    lines += _body(r, max(nq, 1), nc, r.randrange(0, 12))
    if r.random() < 0.3:
        lst = [_q(r, nq) for _ in range(r.randrange(1, 5))]
        lines.append("PMEASURE " + (_ws(r) + "," + _ws(r)).join(str(q) for q in lst))
        if r.random() < 0.1:
            lines.append(_gate(r, nq))
    text = ""
    for ln in lines:
        text += ln + _sep(r)
    if r.random() < 0.25:                                 # byte mutations
        b = list(text)
        for _ in range(r.randrange(1, 4)):
            k = r.random()
            p = r.randrange(0, len(b) + 1)
            if k < 0.35 and b:
                del b[min(p, len(b) - 1)]
            elif k < 0.75:
                b.insert(p, r.choice(NOISE))
            else:
                b.insert(p, r.choice(WORDS))
        text = "".join(b)
    if r.random() < 0.05:
        text = text.rstrip("\n")
    return text


def programs(seed, count):
    r = random.Random(seed)
    return [program(r) for _ in range(count)]


def main():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import qrunes_tables as Q
    if not Q.available():
        sys.exit("the reference tree is needed to write this fixture")
    cases = []
    for text in programs(20261204, 1200):
        q, nodes = Q.parse(text.encode("latin-1"))
        cases.append({"text": text, "qubits": q,
                      "nodes": [[n["kind"], n["qubits"], n["angle"], int(n["dagger"]), n["controls"], n["creg"]]
                                for n in nodes]})
    with open(OUT, "w") as f:
        json.dump({"source": "QPandaSDK/QRunesParser lex.yy.cpp + y.tab.cpp tables, grammar actions, "
                             "LinkedListController.cpp (oracle/qrunes_tables.py)",
                   "cases": cases}, f, separators=(",", ":"))
        f.write("\n")
    acc = sum(c["qubits"] >= 0 for c in cases)
    print("%d programs, %d accepted, %d nodes -> %s (%d bytes)" %
          (len(cases), acc, sum(len(c["nodes"]) for c in cases), OUT, os.path.getsize(OUT)))


if __name__ == "__main__":
    main()
