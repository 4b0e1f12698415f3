"""Generate golden vectors by running QPanda's own CPU simulator.

Runs oracle/_ref/qpanda_ref (the reference's QuantumGates.cpp and
CPUQuantumGates.cpp compiled from the reference tree by oracle/Makefile, driven
by oracle/ref_driver.cpp) on seeded circuits and stores, per case:
  <case>.qsvc    the circuit records (input)
  <case>.state   the state before the first measurement, 2^n x {f64 re, f64 im}
  <case>.shots   per-shot measurement words (bit c = classical register c)
and manifest.json with the seeds and sizes.

Run once in a container that holds the reference tree:
    python tests/golden/gen_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from qpanda_amd import circuits as C  # noqa: E402
import oracle as O  # noqa: E402


def mid_circuit_measure(n: int, seed: int) -> C.Circuit:
    """Layers with measurements and resets between them (draw-order test)."""
    base = C.random_layers(n, 6, seed)
    b = C._Builder(n)
    per_layer = (n + n // 2)
    for i, g in enumerate(base.gates):
        mat = None
        if g["kind"] == C.U2:
            mat = base.mats[g["mat"]:g["mat"] + 16]
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), int(g["q2"]), float(g["theta"]),
              int(g["dagger"]), mat)
        if (i + 1) % per_layer == 0:
            layer = (i + 1) // per_layer
            b.add(C.MEASURE, layer % n, layer % n)
            b.add(C.RESET, (layer + 3) % n)
    for q in range(n):
        b.add(C.MEASURE, q, q)
    return b.build(f"midmeasure_n{n}_s{seed}")


def prep_then_qft(n: int, seed: int) -> C.Circuit:
    prep = C.random_layers(n, 3, seed)
    q = C.qft(n)
    return C.Circuit(n, np.concatenate([prep.gates, q.gates]), prep.mats, f"prep_qft_n{n}_s{seed}")


CASES = [
    # name, circuit, Park-Miller seed, shots
    ("cfg0_random_n16_d20", C.random_layers(16, 20, 20240601, measure_all=True), 1234567, 1024),
    ("allkinds_n10", C.all_gate_kinds(10, 77, reps=4), 4242, 0),
    ("ghz_n12", C.ghz(12), 99, 0),
    ("prep_qft_n10", prep_then_qft(10, 31337), 5, 0),
    ("midmeasure_n8", mid_circuit_measure(8, 2718), 16807, 256),
    ("hadamard_measure_n4", None, 7, 4096),
]


def hadamard_measure(n: int = 4) -> C.Circuit:
    """testMain/testCPUQuantumGates.cpp:335-358 shape: H(0) then measure(0)."""
    b = C._Builder(n)
    b.add(C.H, 0)
    b.add(C.MEASURE, 0, 0)
    return b.build("hadamard_measure_n4")


def main():
    assert O.build_ref(), "reference tree not found"
    manifest = {}
    for name, circ, seed, shots in CASES:
        if circ is None:
            circ = hadamard_measure()
        path = os.path.join(HERE, name)
        circ.save(path + ".qsvc")
        O.ref_run(path + ".qsvc", path, seed, shots)
        manifest[name] = {"n": circ.n, "ngates": circ.ngates, "seed": seed, "shots": shots,
                          "counts": circ.counts()}
        print(name, manifest[name])
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
