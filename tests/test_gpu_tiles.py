"""Gate tiling (qsv_set_gate_tiling): queued gates applied in one sweep over
64 KiB LDS tiles (2^12 complex128 / 2^13 complex64 amplitudes).

Bars: the tiled state is bit-identical (np.array_equal) to the
one-sweep-per-gate state from the same input, at both precisions, through the
gate entry points and through qsv_run_circuit; complex128 results are within
1e-12 of the oracle; measurement outcomes and the Park-Miller stream are
unchanged.  Every entry point that is not a gate applies the queue first, and
qsv_reset_zero drops it.
"""
import sys

import numpy as np
import pytest

import oracle as O
from qpanda_amd import circuits as C
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu

TOL = 1e-12
TILE_BITS = {E.C128: 12, E.C64: 13}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _lib():
    E.lib()
    assert E.device_count() >= 1, "no HIP device visible"


def run_pair(circ, prec, psi, rng=None, measure=False):
    """(per-gate state, tiled state, per-gate stats, tiled stats, outcomes, rng)"""
    outs, rngs, states, stats = [], [], [], []
    for tiled in (False, True):
        with E.StateVector(circ.n, prec) as sv:
            sv.set(psi)
            if tiled:
                sv.set_gate_tiling(True)
            o = np.zeros(64, np.int32)
            st = sv.run(circ, rng_state=rng, outcomes=o if measure else None)
            states.append(sv.get())
            stats.append(st)
            outs.append(o)
            rngs.append(sv.last_rng)
    return states[0], states[1], stats[0], stats[1], outs, rngs


@pytest.mark.parametrize("prec", [E.C128, E.C64])
@pytest.mark.parametrize("n", [13, 14, 17])
def test_tiled_random_circuit_bit_identical(prec, n):
    circ = C.random_layers(n, 12, 900 + n, haar_frac=0.2)
    psi = C.random_gaussian_state(n, 11 + n)
    say("tiles random", "prec", prec, "n", n, "gates", circ.ngates)
    a, b, sa, sb, _, _ = run_pair(circ, prec, psi)
    assert np.array_equal(a, b)
    assert sb.launches < sa.launches / 4, (sa.launches, sb.launches)


@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_tiled_every_gate_kind_bit_identical(prec):
    n = 15
    circ = C.all_gate_kinds(n, 41, reps=5, u1=True)
    psi = C.random_gaussian_state(n, 5)
    say("tiles allkinds", "prec", prec)
    a, b, _, _, _, _ = run_pair(circ, prec, psi)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("exact", [False, True])
def test_tiled_matches_oracle(exact):
    n = 14
    circ = C.all_gate_kinds(n, 7 + exact, reps=4, u1=True)
    psi = C.random_gaussian_state(n, 3)
    ref = O.State(n, exact=exact)
    ref.set(psi)
    ref.run(circ)
    say("tiles oracle", "exact", exact)
    with E.StateVector(n) as sv:
        sv.set(psi)
        sv.set_gate_tiling(True)
        st = sv.run(circ, exact=exact)
        got = sv.get()
    assert np.abs(got - ref.psi).max() <= TOL
    assert st.launches < circ.ngates / 4


def test_tiled_measurements_keep_outcomes_and_draws():
    n = 14
    b = C._Builder(n)
    layers = C.random_layers(n, 10, 77, haar_frac=0.1)
    for i, g in enumerate(layers.gates):
        mat = layers.mats[int(g["mat"]):int(g["mat"]) + 16].reshape(4, 4) if int(g["kind"]) == C.U2 else None
        b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), int(g["q2"]), theta=float(g["theta"]), mat=mat)
        if i % 37 == 36:
            b.add(C.MEASURE, i % n, i % n)
        if i % 53 == 52:
            b.add(C.RESET, (3 * i) % n)
    circ = b.build()
    psi = C.random_gaussian_state(n, 9)
    say("tiles measure", "gates", circ.ngates)
    a, t, _, _, outs, rngs = run_pair(circ, E.C128, psi, rng=12345, measure=True)
    assert np.array_equal(a, t)
    assert np.array_equal(outs[0], outs[1])
    assert rngs[0] == rngs[1]


def _gate_sequence(sv, n, rng):
    """Every gate entry point, with controls, on random qubits."""
    def pick(k):
        qs = list(range(n))
        return [qs.pop(int(rng.integers(len(qs)))) for _ in range(k)]
    for _ in range(6):
        a, c, d = pick(3)
        th = float(rng.uniform(0, 6.28))
        sv.h_gate(a)
        sv.h_gate(c, ctrl=1 << d, exact=True)
        sv.x(a, ctrl=(1 << c) | (1 << d))
        sv.rx(c, th, ctrl=1 << a)
        sv.ry(d, th, dagger=1)
        sv.rz(a, th, ctrl=1 << d)
        u = C.haar_unitary(C.Xoshiro256(int(rng.integers(1 << 30))), 2)
        sv.u1(c, u, ctrl=1 << a)
        sv.diag(d, np.exp(0.3j), np.exp(-1.1j), ctrl=1 << c)
        sv.cphase(a, c, np.exp(0.7j), ctrl=1 << d)
        sv.cr(d, a, th, dagger=1)
        sv.iswap(a, d)
        sv.sqiswap(c, a, dagger=1)
        sv.cswap(a, c, d)
        sv.u2(d, c, C.haar_unitary(C.Xoshiro256(int(rng.integers(1 << 30))), 4))


@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_tiled_gate_entry_points_bit_identical(prec):
    n = 16
    psi = C.random_gaussian_state(n, 21)
    states = []
    for tiled in (False, True):
        with E.StateVector(n, prec) as sv:
            sv.set(psi)
            sv.set_gate_tiling(tiled)
            _gate_sequence(sv, n, np.random.default_rng(5))
            p = sv.prob0(3)                       # a readout between gates flushes
            _gate_sequence(sv, n, np.random.default_rng(6))
            states.append((sv.get(), p))
    assert np.array_equal(states[0][0], states[1][0])
    assert states[0][1] == states[1][1]


def test_queue_longer_than_one_sweep():
    """More than 512 queued gates on qubits 0..7 split over several sweeps."""
    n = 14
    b = C._Builder(n)
    g = np.random.default_rng(1)
    for i in range(1300):
        q = int(g.integers(8))
        b.add((C.H, C.RZ, C.RX)[i % 3], q, theta=float(g.uniform(0, 6.28)))
        if i % 5 == 4:
            b.add(C.CZ, q, (q + 1) % 8)
    circ = b.build()
    psi = C.random_gaussian_state(n, 2)
    a, t, sa, sb, _, _ = run_pair(circ, E.C128, psi)
    assert np.array_equal(a, t)
    assert 3 <= sb.launches <= 4, sb.launches


def test_flush_points():
    """copy_state, inner, gather, norm2, marginals, block k=3 and
    set_gate_tiling(False) all see the queued gates; reset_zero drops them."""
    n = 14
    psi = C.random_gaussian_state(n, 4)
    with E.StateVector(n) as ref, E.StateVector(n) as sv, E.StateVector(n) as cp:
        for s in (ref, sv):
            s.set(psi)
        sv.set_gate_tiling(True)
        for s in (ref, sv):
            s.h_gate(5)
            s.cphase(0, 9, np.exp(0.4j))
        cp.copy_from(sv)
        assert np.array_equal(cp.get(), ref.get())
        for s in (ref, sv):
            s.rx(2, 0.3)
        assert sv.inner(ref) == ref.inner(ref)
        for s in (ref, sv):
            s.x(11, ctrl=1)
        idx = np.arange(0, 1 << n, 97, dtype=np.uint64)
        assert np.array_equal(sv.gather(idx), ref.gather(idx))
        for s in (ref, sv):
            s.ry(13, 1.1)
        assert sv.norm2() == ref.norm2()
        for s in (ref, sv):
            s.iswap(1, 12)
        assert np.array_equal(sv.marginals([12, 1]), ref.marginals([12, 1]))
        u = C.haar_unitary(C.Xoshiro256(3), 8)
        for s in (ref, sv):
            s.h_gate(7)
            s.block([2, 7, 10], u)
        assert np.array_equal(sv.get(), ref.get())
        for s in (ref, sv):
            s.rz(4, 0.9)
        sv.set_gate_tiling(False)
        assert np.array_equal(sv.get(), ref.get())
        sv.set_gate_tiling(True)
        sv.h_gate(3)
        sv.reset()
        z = sv.get()
        assert z[0] == 1 and np.count_nonzero(z) == 1


def test_state_smaller_than_a_tile():
    """Below 12 qubits (complex128) every gate sweeps on its own."""
    n = 10
    circ = C.random_layers(n, 8, 3, haar_frac=0.2)
    psi = C.random_gaussian_state(n, 8)
    a, t, sa, sb, _, _ = run_pair(circ, E.C128, psi)
    assert np.array_equal(a, t)
    assert sa.launches == sb.launches


def test_profiled_and_fused_runs_ignore_tiling():
    """A profiled run times each record's own sweep; the fused path (complex64)
    launches its own block sweeps with the tiling setting left as it was."""
    n = 14
    circ = C.random_layers(n, 6, 8)
    psi = C.random_gaussian_state(n, 1)
    with E.StateVector(n) as sv:
        sv.set(psi)
        sv.set_gate_tiling(True)
        st = sv.run(circ, profile=True)
        assert st.launches == circ.ngates
        st = sv.run(circ)
        assert st.launches < circ.ngates / 4
        got = sv.get()
    ref = O.State(n)
    ref.set(psi)
    ref.run(circ)
    ref.run(circ)
    assert np.abs(got - ref.psi).max() <= TOL
    fused = []
    for tiled in (False, True):
        with E.StateVector(n, E.C64) as sv:
            sv.set(psi)
            sv.set_gate_tiling(tiled)
            st = sv.run(circ, fuse_k=4)
            assert st.blocks > 0
            sv.h_gate(2)                          # queued after the fused run
            fused.append(sv.get())
    assert np.array_equal(fused[0], fused[1])
