"""Writes tests/golden/qrunes_exec.json and qrunes_exec_states.npz: QRunes
programs run through QPanda's own executor and CPU simulator
(oracle/_ref/qpanda_ref_exec, built by `make -C oracle ref` from the reference
tree), with the result text of every run (as QPandaSDK::getResult prints it)
and the final state.

The programs: the reference's example files (tests/golden/qrunes/*.txt that
parse), seeded programs with nested DAGGER / CONTROL blocks over every gate
kind, and seeded programs with mid-circuit MEASURE and RESET.  No gate is
controlled on its own target (the reference's controlled loops then address
the wrong amplitudes).  The instruction lists come from the engine's QRunes
parser (qpanda_amd/libqpanda_sdk.so), which its own tests pin.

Run from the repository root:  python tests/golden/gen_qrunes_golden.py
"""
import json
import os
import random
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from qpanda_amd import sdk  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "qpanda_ref_exec")
OUT = os.path.join(ROOT, "tests", "golden")
ANGLES = ['"pi/3"', '"-pi/7"', '"0.25"', '"2*pi/5"', '"1-0.5"', '"pi^0.5"', '"(pi+1)/3"']
ARITY = {"CNOT": 2, "CR": 2, "ISWAP": 2, "SQISWAP": 2, "TOFFOLI": 3}


def statement(rng, g, qs):
    args = ",".join(str(q) for q in qs)
    if g in ("RX", "RY", "RZ", "CR", "ISWAP"):
        args += "," + rng.choice(ANGLES)
    return f"{g} {args}"


def nested_program(n, seed, depth):
    rng = random.Random(seed)
    lines = [f"QINIT {n}", f"CREG {n}"]
    ctl, dag = [], 0
    for _ in range(depth):
        r = rng.random()
        if r < 0.06 and len(ctl) < 2:
            c = rng.choice([q for q in range(n) if q not in ctl])
            ctl.append(c)
            lines.append(f"CONTROL {c}")
        elif r < 0.10 and ctl:
            lines.append(f"ENDCONTROL {ctl.pop(rng.randrange(len(ctl)))}")
        elif r < 0.14 and dag < 3:
            dag += 1
            lines.append("DAGGER")
        elif r < 0.18 and dag:
            dag -= 1
            lines.append("ENDDAGGER")
        else:
            g = rng.choice(["H", "NOT", "RX", "RY", "RZ", "CNOT", "CR", "ISWAP", "SQISWAP", "TOFFOLI", "HADAMARD"])
            free = [q for q in range(n) if q not in ctl]
            lines.append(statement(rng, g, rng.sample(free, ARITY.get(g, 1))))
    lines += ["ENDDAGGER"] * dag + [f"ENDCONTROL {c}" for c in ctl]
    order = rng.sample(range(n), n)
    regs = rng.sample(range(n), n)
    lines += [f"MEASURE {q},${c}" for q, c in zip(order, regs)]
    return "\n".join(lines) + "\n"


def midcircuit_program(n, seed, depth):
    rng = random.Random(seed)
    lines = [f"QINIT {n}", f"CREG {n}"]
    for _ in range(depth):
        r = rng.random()
        if r < 0.10:
            lines.append(f"MEASURE {rng.randrange(n)},${rng.randrange(n)}")
        elif r < 0.16:
            lines.append(f"RESET {rng.randrange(n)}")
        else:
            g = rng.choice(["H", "NOT", "RX", "RY", "RZ", "CNOT", "CR", "ISWAP", "SQISWAP", "TOFFOLI"])
            lines.append(statement(rng, g, rng.sample(range(n), ARITY.get(g, 1))))
    lines += [f"MEASURE {q},${q}" for q in range(n)]
    return "\n".join(lines) + "\n"


def instruction_list(q, nodes):
    out = [f"QINIT {q}"]
    for nd in nodes:
        out.append(f"{nd['kind']} {' '.join(map(str, nd['qubits']))} | angle={nd['angle']!r} "
                   f"dagger={int(nd['dagger'])} ctrl={','.join(map(str, nd['controls']))} creg={nd['creg']}")
    return "\n".join(out) + "\n"


def run_reference(text, seed, repeat, runs):
    q, nodes = sdk.parse_text(text)
    assert q > 0
    with tempfile.TemporaryDirectory() as td:
        lst = os.path.join(td, "p.list")
        with open(lst, "w") as f:
            f.write(instruction_list(q, nodes))
        codes = subprocess.check_output([EXE, lst, str(seed), str(repeat), str(runs), os.path.join(td, "o")])
        codes = [int(c) for c in codes.decode().split()]
        results = []
        for i in range(runs):
            with open(os.path.join(td, f"o.{i}.txt")) as f:
                results.append(f.read())
        state = np.fromfile(os.path.join(td, "o.state"), dtype=np.complex128)
    assert len(state) == 1 << q
    return codes, results, state


def main():
    if not os.path.exists(EXE):
        raise SystemExit(f"{EXE} is not built (make -C oracle ref)")
    cases = []
    for f in ("shor_15.txt", "shor_15_sdk.txt", "shor.txt", "pmeasure_test.txt", "parser_test.txt"):
        with open(os.path.join(OUT, "qrunes", f)) as fh:
            cases.append({"name": f, "text": fh.read(), "seed": 4001, "repeat": 64, "runs": 2})
    for s in range(1, 13):
        cases.append({"name": f"nested_{s}", "text": nested_program(4 + s % 5, s, 45), "seed": 5000 + s,
                      "repeat": 48, "runs": 2})
    for s in range(1, 9):
        cases.append({"name": f"midcircuit_{s}", "text": midcircuit_program(4 + s % 4, 100 + s, 35),
                      "seed": 6000 + s, "repeat": 40, "runs": 2})
    states = {}
    for c in cases:
        codes, results, state = run_reference(c["text"], c["seed"], c["repeat"], c["runs"])
        c["codes"] = codes
        c["results"] = results
        states[c["name"]] = state
    with open(os.path.join(OUT, "qrunes_exec.json"), "w") as f:
        json.dump({"generator": "tests/golden/gen_qrunes_golden.py",
                   "producer": "QPanda Instruction.cpp / QuantumInstructionGet.cpp / CPUQuantumGates.cpp "
                               "via oracle/ref_exec_driver.cpp", "cases": cases}, f, indent=1)
    np.savez_compressed(os.path.join(OUT, "qrunes_exec_states.npz"), **states)
    print(f"{len(cases)} cases written")


if __name__ == "__main__":
    main()
