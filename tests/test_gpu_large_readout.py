"""Readouts of large states against the oracle (CPUQuantumGates.cpp order):

* measurement decisions inside the tie band at 26 qubits: the state is
  shaped so that P(0) lies within rounding distance of a float, the draw is
  that float, and the outcome must follow the CPU simulator's serial
  hypot^2 sum (CPUQuantumGates.cpp:1074-1088), which the engine recomputes on
  the host;
* pMeasure over m > 16 qubits (up to all 26) with the device top 10, the
  full distribution for m = 25, and equal probabilities ordered by index;
* a QRunes PMEASURE of all 26 qubits through the QPandaSDK facade."""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
from qpanda_amd import circuits as C
from qpanda_amd import engine as E
from qpanda_amd import sdk

pytestmark = pytest.mark.gpu

N = 26


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def band(n, fr):
    return math.ldexp(np.finfo(np.float64).eps, n + 1) * max(1.0, fr)


@pytest.mark.parametrize("q", [0, 11, 25])
def test_measure_tie_band_26_qubits_matches_oracle(q):
    say("tie band n", N, "q", q)
    psi = C.random_gaussian_state(N, 300 + q)
    low = ((np.arange(1 << N, dtype=np.int64) >> q) & 1) == 0
    p0 = float(np.sum(np.abs(psi[low]) ** 2))
    f = float(np.float32(p0))
    psi[low] *= math.sqrt(f / p0)
    psi[~low] *= math.sqrt((1 - f) / (1 - p0))
    st = O.State(N)
    st.set(psi)
    p0s = st.prob0(q)
    assert abs(f - p0s) <= band(N, f)                 # the draw lies inside the band
    with E.StateVector(N) as sv:
        sv.set(psi)
        o = sv.measure(q, f)
        o_ref = st.measure(q, f)
        say("  p0 serial - f", p0s - f, "outcome", o, "oracle", o_ref)
        assert o == o_ref
        assert np.abs(sv.get() - st.psi).max() <= 1e-12


def _check_top(top, full):
    probs = [p for _, p in top]
    assert probs == sorted(probs, reverse=True)
    for i, p in top:
        assert abs(full[i] - p) <= 1e-12
    want = np.sort(full)[::-1][:len(top)]
    assert np.abs(np.array(probs) - want).max() <= 1e-12


def test_pmeasure_over_all_26_qubits_and_25_vs_oracle():
    psi = C.random_gaussian_state(N, 4242)
    st = O.State(N)
    st.set(psi)
    g = np.random.default_rng(5)
    with E.StateVector(N) as sv:
        sv.set(psi)
        qs = g.permutation(N).tolist()                 # m = n, caller order permuted
        say("pmeasure m", len(qs))
        full = st.marginals(qs)
        top = sv.pmeasure(qs)
        assert len(top) == 10
        _check_top(top, full)
        del full
        qs = g.permutation(N)[:25].tolist()            # m = 25: full distribution too
        say("marginals m", len(qs))
        full = st.marginals(qs)
        got = sv.marginals(qs)
        assert np.abs(got - full).max() <= 1e-12
        _check_top(sv.pmeasure(qs), full)
        qs = g.permutation(N)[:16].tolist()            # the LDS-tile path at its limit
        full = st.marginals(qs)
        assert np.abs(sv.marginals(qs) - full).max() <= 1e-12
        _check_top(sv.pmeasure(qs), full)


def test_pmeasure_equal_probabilities_ordered_by_index():
    n = 14
    with E.StateVector(n) as sv:
        for q in range(n):
            sv.h_gate(q, exact=True)
        for m in (3, 12, 14):
            top = sv.pmeasure(list(range(m)))
            assert [i for i, _ in top] == list(range(min(10, 1 << m)))
            assert all(abs(p - 2.0 ** -m) < 1e-15 for _, p in top)


def test_qrunes_pmeasure_of_26_qubits(tmp_path):
    """QINIT 26 / PMEASURE of all qubits through the facade; the ten printed
    outcomes are the oracle's ten most probable."""
    lines = [f"QINIT {N}", "CREG 1"]
    for q in range(N):
        lines.append(f'RY {q},"{0.3 + 0.11 * q}"')
    for q in range(0, N - 1, 3):
        lines.append(f"CNOT {q},{q + 1}")
    lines.append("PMEASURE " + ",".join(str(q) for q in range(N)))
    path = str(tmp_path / "pm26.txt")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    with sdk.SDK() as s:
        assert s.load_file(path) == sdk.Q_ERROR_NONE
        assert s.set_compute_unit(sdk.GPU) == sdk.Q_ERROR_NONE
        s.set_seed(7)
        assert s.run(1) == sdk.Q_ERROR_NONE
        res = s.result()
    got = [ln.split() for ln in res.strip().splitlines()]
    assert len(got) == 10 and all(len(k) == N for k, _ in got)
    nq, nodes = sdk.parse_file(path)
    st = O.State(nq)
    regs, pm, rng = {}, [], 7
    for nd in nodes:
        ok, rng = O._qrunes_node(st, nd, rng, regs, pm)
        assert ok
    full = pm[1]
    want = np.sort(full)[::-1][:10]
    # lines come in key order (the result map is keyed by the bit string)
    vals = sorted((float(v) for _, v in got), reverse=True)
    for k, v in got:
        assert abs(full[int(k, 2)] - float(v)) <= 5e-6 * float(v)
    assert np.abs(np.array(vals) - want).max() <= 5e-6 * want.max()
