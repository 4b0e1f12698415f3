"""The fusion planner (host C++ in libqsv.so, no device needed) against the
oracle: applying the planned FP64 blocks with the oracle's dense-block loop must
reproduce the oracle's gate-by-gate state to 1e-12, for every block size and
gate kind, with measurement records left in place."""
import numpy as np
import pytest

import oracle as O
from qpanda_amd import circuits as C
from qpanda_amd import engine as E


def apply_plan(plan, circ, psi, exact=False):
    st = O.State(circ.n, exact=exact)
    st.set(psi)
    for qs, u, ng in plan:
        assert qs is not None
        st.block(qs, u)
    return st


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_plan_reproduces_random_circuit(k):
    circ = C.random_layers(9, 8, 100 + k, haar_frac=0.2)
    psi = C.random_gaussian_state(9, k)
    ref = O.State(9)
    ref.set(psi)
    ref.run(circ)
    plan = E.fuse_plan(circ, k)
    assert all(len(qs) <= max(k, 2) for qs, _, _ in plan)
    assert sum(ng for _, _, ng in plan) == circ.ngates
    got = apply_plan(plan, circ, psi)
    assert np.abs(got.psi - ref.psi).max() < 1e-12


@pytest.mark.parametrize("k", [3, 4, 5])
@pytest.mark.parametrize("exact", [False, True])
def test_plan_every_gate_kind(k, exact):
    circ = C.all_gate_kinds(8, 5 + k, reps=4, u1=True)
    psi = C.random_gaussian_state(8, 50 + k)
    ref = O.State(8, exact=exact)
    ref.set(psi)
    ref.run(circ)
    plan = E.fuse_plan(circ, k, exact=exact)
    got = apply_plan(plan, circ, psi, exact=exact)
    assert np.abs(got.psi - ref.psi).max() < 1e-12


def test_plan_block_count_shrinks_with_k():
    circ = C.random_layers(16, 20, 7, haar_frac=0.1)
    counts = [len(E.fuse_plan(circ, k)) for k in (1, 2, 3, 4, 5)]
    assert counts == sorted(counts, reverse=True)
    assert counts[3] < circ.ngates / 4


def test_plan_keeps_measurements_in_order():
    b = C._Builder(6)
    for q in range(6):
        b.add(C.H, q)
    b.add(C.MEASURE, 2, 0)
    b.add(C.CNOT, 2, 3)
    b.add(C.RESET, 4)
    b.add(C.RX, 4, theta=0.4)
    circ = b.build()
    plan = E.fuse_plan(circ, 4)
    barriers = [ng for qs, _, ng in plan if qs is None]
    assert barriers == [6, 8]
    # every gate before record 6 is planned before the first barrier
    first = [i for i, (qs, _, _) in enumerate(plan) if qs is None][0]
    assert sum(ng for qs, _, ng in plan[:first]) == 6


def test_plan_blocks_are_unitary():
    circ = C.random_layers(10, 6, 3, haar_frac=0.3)
    for qs, u, _ in E.fuse_plan(circ, 5):
        d = u.shape[0]
        assert np.abs(u.conj().T @ u - np.eye(d)).max() < 1e-12


# ---- window planner (qsv_window.cpp) ------------------------------------------

def apply_window_plan(plan, n, psi, exact=False):
    st = O.State(n, exact=exact)
    st.set(psi)
    for qs, u, _, _ in plan:
        assert qs is not None
        st.block(qs, u)
    return st


@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("tb", [7, 8, 13])
def test_window_plan_reproduces_random_circuit(k, tb):
    n = 11
    circ = C.random_layers(n, 10, 300 + k + tb, haar_frac=0.2)
    psi = C.random_gaussian_state(n, k + tb)
    ref = O.State(n)
    ref.set(psi)
    ref.run(circ)
    plan = E.window_plan(circ, k, tb)
    assert sum(ng for _, _, ng, _ in plan) == circ.ngates
    assert all(len(qs) <= k for qs, _, _, _ in plan)
    # every window's blocks fit in tb qubits together with qubits 0..3
    wins = {}
    for qs, _, _, w in plan:
        wins.setdefault(w, set(range(4))).update(qs)
    assert all(len(s) <= max(tb, 4) for s in wins.values())
    got = apply_window_plan(plan, n, psi)
    assert np.abs(got.psi - ref.psi).max() < 1e-12


@pytest.mark.parametrize("exact", [False, True])
def test_window_plan_every_gate_kind(exact):
    n = 9
    circ = C.all_gate_kinds(n, 21, reps=4, u1=True)
    psi = C.random_gaussian_state(n, 77)
    ref = O.State(n, exact=exact)
    ref.set(psi)
    ref.run(circ)
    plan = E.window_plan(circ, 4, 7, exact=exact)
    assert sum(ng for _, _, ng, _ in plan) == circ.ngates
    got = apply_window_plan(plan, n, psi, exact=exact)
    assert np.abs(got.psi - ref.psi).max() < 1e-12


def test_window_plan_counts_noop_records():
    b = C._Builder(12)
    b.add(C.H, 5)
    b.add(C.CNOT, 7, 7)          # a no-op in the reference loop, but one draw
    b.add(C.CNOT, 5, 9)
    circ = b.build()
    plan = E.window_plan(circ, 4, 13)
    assert sum(ng for _, _, ng, _ in plan) == 3


def test_window_plan_keeps_measurements_in_order():
    b = C._Builder(12)
    for q in range(12):
        b.add(C.H, q)
    b.add(C.MEASURE, 7, 0)
    b.add(C.CNOT, 7, 11)
    b.add(C.RESET, 4)
    b.add(C.RX, 4, theta=0.4)
    circ = b.build()
    plan = E.window_plan(circ, 4, 13)
    assert [ng for qs, _, ng, _ in plan if qs is None] == [12, 14]
    first = [i for i, p in enumerate(plan) if p[0] is None][0]
    assert sum(p[2] for p in plan[:first]) == 12
    # barriers sit in windows of their own
    wb = [p[3] for p in plan if p[0] is None]
    assert all(sum(1 for p in plan if p[3] == w) == 1 for w in wb)


def test_window_plan_needs_fewer_sweeps_than_block_first_grouping():
    """30 qubits, depth 200 (the benchmark circuit): the window planner needs
    at most 250 sweeps of 13 tile bits for its 9,000 gates (288 with greedy
    growth, no lookahead); block-first
    grouping of greedy k = 4 blocks (every run of consecutive blocks whose
    qubits above 3 number <= 9 shares a sweep) needs more than 400."""
    circ = C.random_layers(30, 200, 2025, haar_frac=0.1)
    plan = E.window_plan(circ, 4, 13)
    windows = len({w for *_, w in plan})
    blocks = E.fuse_plan(circ, 4)
    sweeps, cur = 0, None
    for qs, _, _ in blocks:
        hi = {q for q in qs if q > 3}
        if cur is not None and len(cur | hi) <= 9:
            cur |= hi
        else:
            sweeps += 1
            cur = set(hi)
    assert windows <= 250 < 400 < sweeps, (windows, sweeps)


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_plans_leave_no_mergeable_neighbours(k):
    """Neighbouring blocks commute into one another's place trivially, so the
    planners merge every neighbouring pair whose union has <= k qubits: none
    may be left (block-first plans, and within each window of a window
    plan)."""
    circ = C.random_layers(16, 30, 77 + k, haar_frac=0.1)
    plan = E.fuse_plan(circ, k)
    for (qa, _, _), (qb, _, _) in zip(plan, plan[1:]):
        assert len(set(qa) | set(qb)) > k, (qa, qb)
    if k >= 4:
        wplan = E.window_plan(circ, k, 10)
        for (qa, _, _, wa), (qb, _, _, wb) in zip(wplan, wplan[1:]):
            if wa == wb:
                assert len(set(qa) | set(qb)) > k, (qa, qb)


def test_window_blocks_are_mostly_full_width():
    """The benchmark circuit (30 qubits, depth 200, k = 4): merging narrow
    blocks into their neighbours leaves no single-qubit block and at least
    60% of the blocks at the full 4 qubits (greedy fusion alone: 41%)."""
    circ = C.random_layers(30, 200, 2025, haar_frac=0.1)
    widths = [len(qs) for qs, *_ in E.window_plan(circ, 4, 13) if qs is not None]
    assert min(widths) >= 2
    assert sum(w == 4 for w in widths) >= 0.6 * len(widths), np.bincount(widths)


def test_window_plan_tile_bits_range():
    """A window holds qubits 0..3 and grows from there: below 7 tile bits a
    3-qubit record may not fit beside them, so tb < 7 is rejected unless the
    window holds the whole state; from 7 up every record fits and every gate
    is planned (a Toffoli on qubits 5, 6, 7 included)."""
    n = 11
    b = C._Builder(n)
    for g in C.random_layers(n, 6, 31).gates:
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
    b.add(C.TOFFOLI, 5, 6, 7)
    b.add(C.CSWAP, 10, 4, 8)
    circ = b.build()
    for tb in (4, 5, 6):
        with pytest.raises(E.QsvError):
            E.window_plan(circ, 4, tb)
    for tb in (7, 8, 11, 13):
        plan = E.window_plan(circ, 4, tb)
        assert sum(ng for _, _, ng, _ in plan) == circ.ngates
    small = C.random_layers(5, 4, 3)
    assert sum(ng for _, _, ng, _ in E.window_plan(small, 4, 5)) == small.ngates


# ---- beam search over windows (WindowStream, qsv_window.cpp) -------------------

BEAM = (4, 3, 6, 4)         # the learned schedule's width, branch, first, delay at 29 qubits


def test_beam_width_one_is_the_window_plan():
    """With one state the beam keeps the lookahead growth's windows: the same
    blocks, matrices and windows as qsv_window_plan."""
    circ = C.random_layers(16, 30, 4040, haar_frac=0.1)
    a = E.window_plan(circ, 4, 13)
    b = E.window_plan(circ, 4, 13, beam=(1, 3, 6, 4))
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x[0] == y[0] and x[2] == y[2] and x[3] == y[3]
        assert np.array_equal(x[1], y[1])


@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("tb", [7, 9, 13])
@pytest.mark.parametrize("tries", [0, 8])
def test_beam_plan_reproduces_random_circuit(k, tb, tries):
    """Beam windows, fused in the closure's, program or (tries > 0) random
    dependency orders: the planned FP64 blocks reproduce the gate-by-gate
    state."""
    n = 12
    circ = C.random_layers(n, 12, 700 + k + tb, haar_frac=0.2)
    psi = C.random_gaussian_state(n, 3 * k + tb)
    ref = O.State(n)
    ref.set(psi)
    ref.run(circ)
    plan = E.window_plan(circ, k, tb, beam=BEAM + (tries,))
    assert sum(ng for _, _, ng, _ in plan) == circ.ngates
    assert all(len(qs) <= k for qs, _, _, _ in plan)
    wins = {}
    for qs, _, _, w in plan:
        wins.setdefault(w, set(range(4))).update(qs)
    assert all(len(s) <= max(tb, 4) for s in wins.values())
    got = apply_window_plan(plan, n, psi)
    assert np.abs(got.psi - ref.psi).max() < 1e-12


def test_beam_plan_every_gate_kind_and_measurements():
    n = 10
    b = C._Builder(n)
    for seed, rec in ((61, (C.MEASURE, 3, 0)), (62, (C.RESET, 8)), (63, None)):
        part = C.all_gate_kinds(n, seed, reps=2, u1=True)
        for g in part.gates:
            k, mi = int(g["kind"]), int(g["mat"])
            mat = None
            if mi >= 0:
                size = 4 if k == C.U1 else 16
                mat = part.mats[mi: mi + size]
            b.add(k, int(g["q0"]), int(g["q1"]), int(g["q2"]), theta=float(g["theta"]), dagger=int(g["dagger"]),
                  mat=mat)
        if rec:
            b.add(*rec)
    circ = b.build()
    plan = E.window_plan(circ, 4, 8, beam=BEAM + (8,))
    assert sum(p[2] for p in plan if p[0] is not None) == circ.ngates - 2
    # the barriers in order, each in a window of its own, after exactly the
    # records that precede them
    bars = [i for i, p in enumerate(plan) if p[0] is None]
    assert [plan[i][2] for i in bars] == [int(np.nonzero(circ.gates["kind"] == C.MEASURE)[0][0]),
                                          int(np.nonzero(circ.gates["kind"] == C.RESET)[0][0])]
    assert sum(p[2] for p in plan[:bars[0]]) == plan[bars[0]][2]
    assert all(sum(1 for p in plan if p[3] == plan[i][3]) == 1 for i in bars)


def test_beam_needs_fewer_windows_on_the_benchmark_circuit():
    """30 qubits, depth 200: 234 windows with the lookahead growth alone, at
    most 225 with the 29-qubit beam (221 measured) and at most 220 with the
    beam the learned schedule uses from 30 qubits (216 measured)."""
    circ = C.random_layers(30, 200, 2025, haar_frac=0.1)
    w1 = len({w for *_, w in E.window_plan(circ, 4, 13)})
    wb = len({w for *_, w in E.window_plan(circ, 4, 13, beam=BEAM)})
    w12 = len({w for *_, w in E.window_plan(circ, 4, 13, beam=(12, 4, 8, 6))})
    assert w12 <= 220 and wb <= 225 < w1, (w12, wb, w1)


def test_random_fusion_orders_need_fewer_blocks():
    """The learned schedule fuses each window in 24 random dependency orders
    too: on the benchmark circuit's beam windows 1,092 blocks become <= 1,050
    (1,027 measured), the same windows and records."""
    circ = C.random_layers(30, 200, 2025, haar_frac=0.1)
    a = E.window_plan(circ, 4, 13, beam=(12, 4, 8, 6, 0))
    b = E.window_plan(circ, 4, 13, beam=(12, 4, 8, 6, 24))
    assert len({w for *_, w in a}) == len({w for *_, w in b})
    assert sum(p[2] for p in b) == circ.ngates
    assert len(b) <= 1050 < len(a), (len(b), len(a))


@pytest.mark.parametrize("width", [1, 4])
@pytest.mark.parametrize("k", [4, 5])
def test_relabelled_plan_reproduces_random_circuit(width, k):
    """Relabelled windows (each grows from four qubits of the previous one
    instead of from qubits 0..3): the blocks keep their logical qubits, so the
    planned FP64 blocks reproduce the gate-by-gate state; consecutive windows
    share at least four qubits and every window holds at most 13."""
    n = 18
    circ = C.random_layers(n, 14, 900 + k + width, haar_frac=0.2)
    psi = C.random_gaussian_state(n, 11 * k + width)
    ref = O.State(n)
    ref.set(psi)
    ref.run(circ)
    plan = E.window_plan(circ, k, 13, beam=(width, 3, 6, 4, 0, 1))
    assert sum(ng for _, _, ng, _ in plan) == circ.ngates
    wins = {}
    for qs, _, _, w in plan:
        wins.setdefault(w, set()).update(qs)
    assert all(len(s) <= 13 for s in wins.values())
    got = apply_window_plan(plan, n, psi)
    assert np.abs(got.psi - ref.psi).max() < 1e-12


def test_relabelling_needs_fewer_windows_on_the_benchmark_circuit():
    """30 qubits, depth 200: windows that grow from the four qubits the
    previous sweep left at bits 0..3 need <= 195 sweeps with the lookahead
    growth alone (188 measured, against 234 from qubits 0..3) and <= 182 with
    the 30-qubit beam (176 measured, against 216)."""
    circ = C.random_layers(30, 200, 2025, haar_frac=0.1)
    w1 = len({w for *_, w in E.window_plan(circ, 4, 13, beam=(1, 3, 6, 4, 0, 1))})
    w12 = len({w for *_, w in E.window_plan(circ, 4, 13, beam=(12, 4, 8, 6, 0, 1))})
    assert w12 <= 182 and w1 <= 195, (w12, w1)
