"""Fused FP32 MFMA blocks (complex64 state) against the FP64 oracle, for the
four products: FP32 (split 0), FP32 + BF16 lo correction (split 1), BF16x6
(split 2, the product on the BF16 matrix cores) and FP16x3 (split 3, two fp16
pieces of matrix and scaled state, three piece products on the FP16 matrix
cores).

Bars: one block application within 2e-6 max amplitude error (FP32 storage and
accumulation, ~1e-7 relative); whole fused circuits at fidelity >= 1 - 1e-6
against the FP64 oracle state.  Block qubit sets cover the low run of qubits
(inside the contiguous 7-bit tile run), high qubits and mixtures, so every
tile-addressing case of the kernel is exercised.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
from qpanda_amd import circuits as C
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


QSETS = {
    4: [[0, 1, 2, 3], [3, 2, 1, 0], [12, 13, 14, 15], [0, 7, 11, 15], [5, 9, 6, 14], [15, 0, 8, 3]],
    5: [[0, 1, 2, 3, 4], [11, 12, 13, 14, 15], [0, 6, 7, 13, 2], [15, 9, 1, 4, 10]],
    2: [[0, 1], [14, 3]],
    3: [[0, 5, 15], [8, 9, 10]],
}


@pytest.mark.parametrize("k", [2, 3, 4, 5])
@pytest.mark.parametrize("split", [0, 1, 2, 3])
def test_block_mfma_vs_oracle(k, split):
    n = 16
    for idx, qs in enumerate(QSETS[k]):
        u = C.haar_unitary(C.Xoshiro256(31 * k + idx), 1 << k)
        psi = C.random_gaussian_state(n, 7 * k + idx)
        st = O.State(n)
        st.set(psi)
        st.block(qs, u)
        say("block", k, qs, "split", split)
        with E.StateVector(n, E.C64) as sv:
            sv.set(psi)
            sv.block(qs, u, split=split)
            got = sv.get()
        err = np.abs(got - st.psi).max()
        assert err < 2e-6 * np.abs(st.psi).max() * 64, (qs, err)
        fid = abs(np.vdot(st.psi, got)) ** 2 / (np.vdot(got, got).real * np.vdot(st.psi, st.psi).real)
        assert fid > 1 - 1e-10, (qs, fid)


def test_split_is_not_less_accurate():
    n = 16
    qs = [2, 9, 13, 4, 6]
    u = C.haar_unitary(C.Xoshiro256(99), 32)
    psi = C.random_gaussian_state(n, 3)
    st = O.State(n)
    st.set(psi)
    for _ in range(20):
        st.block(qs, u)
    errs = {}
    for split in (0, 1, 2, 3):
        with E.StateVector(n, E.C64) as sv:
            sv.set(psi)
            for _ in range(20):
                sv.block(qs, u, split=split)
            errs[split] = np.abs(sv.get() - st.psi).max()
    say("20 blocks, max error FP32 / split / BF16x6 / FP16x3", errs[0], errs[1], errs[2], errs[3])
    assert errs[1] <= errs[0] * 1.1
    assert errs[2] <= errs[0] * 1.1
    # FP16x3 keeps 22 of the 24 operand bits: within 8x of the FP32 product
    assert errs[3] <= errs[0] * 8


@pytest.mark.parametrize("scale", [1e-7, 1.0, 300.0])
def test_fp16x3_scale_follows_the_norm(scale):
    """FP16x3 scales the state by 2^(15 - e), 2^e >= sqrt(norm): unnormalised
    states (norm 1e-14 .. 9e4) keep the relative accuracy of a normalised one."""
    n = 16
    qs = [0, 5, 11, 15]
    u = C.haar_unitary(C.Xoshiro256(5), 16)
    psi = C.random_gaussian_state(n, 11) * scale
    psi[12345] = 0.9 * scale          # one large amplitude next to ~2^-8 ones
    st = O.State(n)
    st.set(psi)
    st.block(qs, u)
    with E.StateVector(n, E.C64) as sv:
        sv.set(psi)
        sv.block(qs, u, split=3)
        got = sv.get()
    rel = np.abs(got - st.psi).max() / np.abs(st.psi).max()
    say("fp16x3 scale", scale, "relative max error", rel)
    assert np.all(np.isfinite(got))
    assert rel < 2e-6


@pytest.mark.parametrize("k", [2, 3, 4, 5])
@pytest.mark.parametrize("split", [1, 2, 3])
def test_fused_circuit_fidelity(k, split):
    n = 16
    circ = C.random_layers(n, 20, 500 + k, haar_frac=0.1)
    ref, _ = O.run_circuit(circ)
    say("fused circuit k", k, "split", split)
    with E.StateVector(n, E.C64) as sv:
        st = sv.run(circ, fuse_k=k, split=split)
        got = sv.get()
    assert st.blocks > 0 and st.blocks < circ.ngates
    fid = abs(np.vdot(ref.psi, got)) ** 2 / np.vdot(got, got).real
    say("k", k, "blocks", st.blocks, "fidelity", fid)
    assert fid >= 1 - 1e-6
    assert np.abs(got - ref.psi).max() < 1e-5


def test_fused_matches_plain_c64_and_device_fidelity():
    n = 18
    circ = C.random_layers(n, 12, 77, haar_frac=0.1)
    with E.StateVector(n, E.C128) as p, E.StateVector(n, E.C64) as f:
        p.run(circ)
        f.run(circ, fuse_k=4, split=True)
        fid = p.fidelity(f)
        ref = p.get()
        got = f.get()
    host = abs(np.vdot(ref, got)) ** 2 / (np.vdot(got, got).real * np.vdot(ref, ref).real)
    assert abs(fid - host) < 1e-12
    assert fid >= 1 - 1e-6


def test_fused_circuit_with_measurements_and_rng_accounting():
    n = 14
    b = C._Builder(n)
    base = C.random_layers(n, 6, 9)
    for g in base.gates:
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
    b.add(C.MEASURE, 3, 0)
    b.add(C.RESET, 5)
    for q in range(n):
        b.add(C.H, q)
    circ = b.build()
    out = np.zeros(64, np.int32)
    with E.StateVector(n, E.C64) as f, E.StateVector(n, E.C128) as p:
        f.run(circ, fuse_k=4, rng_state=4321, outcomes=out)
        rng_fused = f.last_rng
        out_p = np.zeros(64, np.int32)
        p.run(circ, rng_state=4321, outcomes=out_p)
        assert rng_fused == p.last_rng          # one draw per logical gate, as the reference
        assert abs(f.norm2() - 1) < 1e-5
        if out[0] == out_p[0]:
            assert f.fidelity(p) >= 1 - 1e-6


def test_fused_large_state_sweep_30q():
    n = 30
    circ = C.random_layers(n, 2, 1, haar_frac=0.1)
    say("fused 30q depth 2")
    with E.StateVector(n, E.C64) as sv:
        st = sv.run(circ, fuse_k=4, split=True, profile=True)
        assert abs(sv.norm2() - (2 * 0.707106781186548 ** 2) ** circ.counts().get("H", 0)) < 1e-4
    c = st.classes()["fused"]
    say("fused launches", c["launches"], "ms/launch", c["ms"] / c["launches"], "GB/s", c["gbs"])


@pytest.mark.parametrize("group,pass_tb", [(2, 0), (3, 0), (8, 0), (4, 13)])
@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("split", [1, 2, 3])
def test_multi_block_passes_vs_oracle(group, pass_tb, k, split):
    n = 18
    circ = C.random_layers(n, 16, 900 + group + k, haar_frac=0.1)
    ref, _ = O.run_circuit(circ)
    say("passes group", group, "tb", pass_tb, "k", k, "split", split)
    with E.StateVector(n, E.C64) as sv:
        st = sv.run(circ, fuse_k=k, split=split, scheduler=2, group=group, pass_tb=pass_tb, profile=True)
        got = sv.get()
        single = E.StateVector(n, E.C64)
        st1 = single.run(circ, fuse_k=k, split=split, profile=True)
        single.close()
    cl = st.classes()
    sweeps = cl.get("fused", {}).get("launches", 0) + cl.get("pass", {}).get("launches", 0)
    assert cl.get("pass", {}).get("launches", 0) > 0
    assert sweeps < st1.classes()["fused"]["launches"]     # blocks really shared sweeps
    fid = abs(np.vdot(ref.psi, got)) ** 2 / np.vdot(got, got).real
    say("sweeps", sweeps, "vs", st1.classes()["fused"]["launches"], "fidelity", fid)
    assert fid >= 1 - 1e-6
    assert np.abs(got - ref.psi).max() < 1e-5


@pytest.mark.parametrize("sched", [0, 2, 3])
def test_passes_with_measurements_keep_rng_accounting(sched):
    """One draw per record on every fused planner, no-op records (CNOT(q, q),
    which the reference loop skips but still draws for) included, also where
    a no-op is alone between two measurements."""
    n = 14
    circ = C.random_layers(n, 6, 5)
    b = C._Builder(n)
    for g in circ.gates:
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
    b.add(C.CNOT, 4, 4)
    b.add(C.MEASURE, 2, 0)
    b.add(C.CNOT, 9, 9)
    b.add(C.MEASURE, 6, 1)
    for g in C.random_layers(n, 3, 6).gates:
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
    b.add(C.CNOT, 1, 1)
    c2 = b.build()
    out_f = np.zeros(64, np.int32)
    out_p = np.zeros(64, np.int32)
    with E.StateVector(n, E.C64) as f, E.StateVector(n, E.C128) as p:
        f.run(c2, fuse_k=4, split=3, scheduler=sched, group=8, rng_state=99, outcomes=out_f)
        p.run(c2, rng_state=99, outcomes=out_p)
        assert f.last_rng == p.last_rng
        if np.array_equal(out_f[:2], out_p[:2]):
            assert f.fidelity(p) >= 1 - 1e-6


@pytest.mark.parametrize("pass_tb", [0, 12])
@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("split", [2, 3])
def test_window_scheduler_vs_oracle(pass_tb, k, split):
    """Window planner (scheduler 3): one sweep per window, every block inside
    it; the state matches the FP64 oracle at the fused-path bar."""
    n = 18
    circ = C.random_layers(n, 16, 700 + k + pass_tb, haar_frac=0.1)
    ref, _ = O.run_circuit(circ)
    say("window scheduler tb", pass_tb, "k", k, "split", split)
    with E.StateVector(n, E.C64) as sv:
        st = sv.run(circ, fuse_k=k, split=split, scheduler=3, pass_tb=pass_tb, profile=True)
        got = sv.get()
    assert st.planner == 1 and st.fuse_k == k
    # (18 qubits: qubits 0..3 stay in every window by default)
    windows = len({w for *_, w in E.window_plan(circ, k, pass_tb or 13)})
    cl = st.classes()
    sweeps = cl.get("fused", {}).get("launches", 0) + cl.get("pass", {}).get("launches", 0)
    assert windows <= sweeps <= windows + st.blocks // 16
    fid = abs(np.vdot(ref.psi, got)) ** 2 / np.vdot(got, got).real
    say("sweeps", sweeps, "blocks", st.blocks, "fidelity", fid)
    assert fid >= 1 - 1e-6
    assert np.abs(got - ref.psi).max() < 1e-5


@pytest.mark.parametrize("fixed_low", [0, -1])
def test_window_scheduler_with_measurements_keeps_rng_accounting(fixed_low):
    """fixed_low = -1 relabels qubits between sweeps at 14 qubits: the order
    is restored before the MEASURE and RESET records."""
    n = 14
    b = C._Builder(n)
    for g in C.random_layers(n, 6, 15).gates:
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
    b.add(C.CNOT, 6, 6)
    b.add(C.MEASURE, 9, 0)
    b.add(C.RESET, 2)
    for g in C.random_layers(n, 3, 16).gates:
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
    circ = b.build()
    out_f = np.zeros(8, np.int32)
    out_p = np.zeros(8, np.int32)
    with E.StateVector(n, E.C64) as f, E.StateVector(n, E.C128) as p:
        f.run(circ, fuse_k=4, split=3, scheduler=3, rng_state=321, outcomes=out_f, fixed_low=fixed_low)
        p.run(circ, rng_state=321, outcomes=out_p)
        assert f.last_rng == p.last_rng          # one draw per record, as the reference
        if out_f[0] == out_p[0]:
            assert f.fidelity(p) >= 1 - 1e-6


@pytest.mark.parametrize("n", [8, 10, 11, 12, 13, 14])
def test_window_and_learned_schedulers_at_small_sizes(n):
    """Edges of the window planner: below 11 qubits scheduler 3 falls back to
    one sweep per block; at 11..13 one window covers the whole state and its
    blocks are split into launches of <= 16; the learned scheduler picks among
    its candidates at every size.  States against the FP64 oracle, draws
    counted as the reference does."""
    circ = C.random_layers(n, 24, 6100 + n, haar_frac=0.1)
    ref, _ = O.run_circuit(circ)
    m = E.CostModel.load() if os.path.exists(E.DEFAULT_COSTMODEL) else E.CostModel(5)
    for sched in (3, 1):
        with E.StateVector(n, E.C64) as sv:
            if sched == 1:
                sv.attach_costmodel(m)
            st = sv.run(circ, fuse_k=4, split=3, scheduler=sched, rng_state=77)
            got = sv.get()
            fid = abs(np.vdot(ref.psi, got)) ** 2 / np.vdot(got, got).real
        say("n", n, "scheduler", sched, "planner", st.planner, "launches", st.launches, "fidelity", fid)
        assert (st.planner == 1) == (sched == 3 and n >= 11) or sched == 1
        assert fid >= 1 - 1e-6
        assert np.abs(got - ref.psi).max() < 1e-5


@pytest.mark.parametrize("n", [14, 16, 21])
@pytest.mark.parametrize("sched", [3, 1])
def test_relabelled_schedules_vs_oracle(n, sched):
    """Window schedules with relabelling (each sweep leaves four of its qubits
    at bits 0..3 for the next window; the order is restored at the end),
    forced below its default 28 qubits, against the FP64 oracle, and against
    the same schedule with qubits 0..3 fixed: fewer sweeps, the same state to
    the fused-path bar."""
    circ = C.random_layers(n, 30, 7300 + n, haar_frac=0.1)
    ref, _ = O.run_circuit(circ)
    m = E.CostModel.load() if os.path.exists(E.DEFAULT_COSTMODEL) else E.CostModel(5)
    sweeps, states = {}, {}
    for fixed in (False, True):
        with E.StateVector(n, E.C64) as sv:
            if sched == 1:
                sv.attach_costmodel(m)
            st = sv.run(circ, fuse_k=4, split=3, scheduler=sched, profile=True, fixed_low=1 if fixed else -1,
                        rng_state=5)
            got = sv.get()
        cl = st.classes()
        sweeps[fixed] = cl.get("fused", {}).get("launches", 0) + cl.get("pass", {}).get("launches", 0)
        states[fixed] = got
        fid = abs(np.vdot(ref.psi, got)) ** 2 / np.vdot(got, got).real
        say("n", n, "scheduler", sched, "fixed_low", fixed, "sweeps", sweeps[fixed], "blocks", st.blocks,
            "fidelity", fid)
        assert fid >= 1 - 1e-6
        assert np.abs(got - ref.psi).max() < 1e-5
    # at 14 qubits a window holds all but one qubit: no fewer windows, and the
    # last window's store puts the qubits back
    assert sweeps[False] <= sweeps[True] if n < 18 else sweeps[False] < sweeps[True]
    assert np.abs(states[False] - states[True]).max() < 1e-5
