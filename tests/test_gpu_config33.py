"""Configuration 4 (33 qubits) and the 30-qubit headline path under test.

* 33 qubits, depth 60, 10% Haar pairs, complex64 (64 GiB): the circuit runs
  as k = 5 FP16x3 windows and under the learned scheduler; both FP64 norms
  stay at 1, the two states agree (fidelity), 10^5 shots from the copy-free
  sampler give a linear cross-entropy of 1 (ideal sampling of a random
  circuit's Porter-Thomas output) and their first shots equal the per-shot
  copy path's; the same generator and seed at 28 qubits run through both
  fused paths match the plain FP64 path at fidelity >= 1 - 1e-6.
* 30 qubits, depth 20: the benchmark path (learned schedule, FP16x3 windows)
  against the plain per-gate FP64 path at fidelity >= 1 - 1e-6.
The reference cannot run these sizes (it caps programs at 30 qubits,
QPandaAPI.cpp:87, and keeps a 30-qubit state in mapped host memory,
GPUGates.cu:1211-1233), so the checks are the FP64 path and invariants."""
import os
import sys

import numpy as np
import pytest

from qpanda_amd import circuits as C
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu

SEED = 2033


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def costmodel():
    assert os.path.exists(E.DEFAULT_COSTMODEL), "trained cost model missing"
    return E.CostModel.load()


def test_config4_33_qubits_k5_windows_learned_shots_and_28q_restriction():
    n, depth = 33, 60
    circ = C.random_layers(n, depth, SEED, haar_frac=0.1)
    say("config 4: n", n, "gates", circ.ngates)
    cm = costmodel()
    with E.StateVector(n, E.C64) as a:
        st = a.run(circ, fuse_k=5, split=3, scheduler=3)
        say("  k=5 windows: blocks", st.blocks, "launches", st.launches, "device ms", round(st.device_ms, 1))
        assert st.planner == 1 and st.fuse_k == 5
        norm_a = a.norm2()
        say("  norm", norm_a)
        assert abs(norm_a - 1) < 1e-4
        qs = list(range(n))
        words, _, ties = a.sample_shots(qs, circ.ngates, 12345, 100000)
        p = np.abs(a.gather(words)) ** 2 / norm_a
        xeb = float((1 << n) * p.mean() - 1)
        say("  1e5 shots: linear XEB", xeb, "tie re-runs", ties)
        assert abs(xeb - 1) <= 0.05
        first, _, _ = a.sample_shots(qs, circ.ngates, 12345, 3, method="copy")
        assert np.array_equal(first, words[:3])
        with E.StateVector(n, E.C64) as b:
            b.attach_costmodel(cm)
            sl = b.run(circ, fuse_k=4, split=3, scheduler=1)
            say("  learned: planner", sl.planner, "k", sl.fuse_k, "blocks", sl.blocks, "device ms",
                round(sl.device_ms, 1))
            assert abs(b.norm2() - 1) < 1e-4
            fab = a.fidelity(b)
            say("  fidelity k=5 windows vs learned", fab)
            assert fab >= 1 - 1e-6
    r = 28
    circ28 = C.random_layers(r, depth, SEED, haar_frac=0.1)
    with E.StateVector(r, E.C128) as ref:
        ref.run(circ28)
        for sched, k in ((3, 5), (1, 4)):
            with E.StateVector(r, E.C64) as f:
                if sched == 1:
                    f.attach_costmodel(cm)
                f.run(circ28, fuse_k=k, split=3, scheduler=sched)
                fid = ref.fidelity(f)
            say("  28-qubit restriction, scheduler", sched, "fidelity vs FP64", fid)
            assert fid >= 1 - 1e-6


def test_headline_30_qubits_learned_fp16x3_vs_plain_fp64():
    n, depth = 30, 20
    circ = C.random_layers(n, depth, 2025, haar_frac=0.1)
    with E.StateVector(n, E.C128) as ref, E.StateVector(n, E.C64) as f:
        ref.run(circ)
        f.attach_costmodel(costmodel())
        st = f.run(circ, fuse_k=4, split=3, scheduler=1)
        fid = ref.fidelity(f)
        say("30q depth", depth, "learned planner", st.planner, "k", st.fuse_k, "blocks", st.blocks,
            "fidelity", fid)
        assert st.planner == 1                       # window plans, as in the benchmark
        assert fid >= 1 - 1e-6
