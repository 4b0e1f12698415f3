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
    circ = C.all_gate_kinds(8, 5 + k, reps=4)
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
