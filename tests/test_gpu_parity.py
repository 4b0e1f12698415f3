"""Parity of the HIP engine (through the C ABI) with the CPU oracle and the
reference's golden vectors.

Bars: complex128 per-gate path within 1e-12 max amplitude error of the oracle
(which is itself bit-identical to QPanda's CPU simulator on the golden cases);
measurement outcomes bit-exact for the same Park-Miller stream; complex64
states within 2e-6 (FP32 storage).  Sizes here keep the oracle to seconds; the
full configuration-1 sizes are checked through closed forms (GHZ) and numpy's
FFT (QFT).
"""
import json
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
from qpanda_amd import circuits as C
from qpanda_amd import engine as E

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _lib():
    E.lib()
    assert E.device_count() >= 1, "no HIP device visible"


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def upto_first_measure(circ):
    k = np.nonzero(circ.gates["kind"] == C.MEASURE)[0]
    return circ.prefix(int(k[0])) if k.size else circ


@pytest.mark.parametrize("case", ["cfg0_random_n16_d20", "allkinds_n10", "ghz_n12", "prep_qft_n10",
                                  "midmeasure_n8", "hadamard_measure_n4"])
def test_golden_state_c128(case):
    circ = C.Circuit.load(os.path.join(GOLDEN, case + ".qsvc"))
    ref = np.fromfile(os.path.join(GOLDEN, case + ".state"), np.complex128)
    say("golden", case, "n", circ.n)
    with E.StateVector(circ.n) as sv:
        sv.run(upto_first_measure(circ), rng_state=manifest()[case]["seed"])
        got = sv.get()
    assert np.abs(got - ref).max() <= TOL


@pytest.mark.parametrize("method", ["tree", "copy"])
@pytest.mark.parametrize("case", ["cfg0_random_n16_d20", "hadamard_measure_n4"])
def test_shots_bit_exact(case, method):
    meta = manifest()[case]
    circ = C.Circuit.load(os.path.join(GOLDEN, case + ".qsvc"))
    ref = np.fromfile(os.path.join(GOLDEN, case + ".shots"), np.uint64)
    pre = upto_first_measure(circ)
    mq = [int(g["q0"]) for g in circ.gates if g["kind"] == C.MEASURE]
    assert [int(g["q1"]) for g in circ.gates if g["kind"] == C.MEASURE] == list(range(len(mq)))
    say("shots", case, "n", circ.n, "shots", meta["shots"])
    with E.StateVector(circ.n) as sv:
        sv.run(pre)
        words, rng, ties = sv.sample_shots(mq, pre.ngates, meta["seed"], meta["shots"], method=method)
    say("tie resolutions", ties)
    assert np.array_equal(words, ref)


def test_midcircuit_measure_reset_bit_exact():
    case = "midmeasure_n8"
    meta = manifest()[case]
    circ = C.Circuit.load(os.path.join(GOLDEN, case + ".qsvc"))
    ref = np.fromfile(os.path.join(GOLDEN, case + ".shots"), np.uint64)
    rng = meta["seed"]
    out = np.zeros(64, np.int32)
    got = []
    say("midmeasure shots", meta["shots"])
    with E.StateVector(circ.n) as sv:
        for _ in range(meta["shots"]):
            sv.reset()
            out[:] = 0
            sv.run(circ, rng_state=rng, outcomes=out)
            rng = sv.last_rng
            got.append(sum(int(out[c] & 1) << c for c in range(circ.n)))
    assert np.array_equal(np.array(got, np.uint64), ref)


@pytest.mark.parametrize("prec,tol", [(E.C128, TOL), (E.C64, 2e-6)])
def test_every_gate_kind_vs_oracle(prec, tol):
    """Every record kind, U1 (general 2x2) included, in plain and exact mode;
    the oracle's exact-mode gates are pinned by closed forms on the CPU."""
    n = 9
    circ = C.all_gate_kinds(n, 1234, reps=5, u1=True)
    assert (circ.gates["kind"] == C.U1).sum() == 10
    psi = C.random_gaussian_state(n, 8)
    for exact in (False, True):
        st = O.State(n, exact=exact)
        st.set(psi)
        st.run(circ)
        with E.StateVector(n, prec) as sv:
            sv.set(psi)
            sv.run(circ, exact=exact)
            got = sv.get()
        say("gate kinds prec", prec, "exact", exact, "max err", np.abs(got - st.psi).max())
        assert np.abs(got - st.psi).max() <= tol


def ref_single(op, th, dag):
    c, s = math.cos(th / 2), math.sin(th / 2)
    if op == "RX":
        m = np.array([[c, -1j * s], [-1j * s, c]])
    elif op == "RY":
        m = np.array([[c, -s], [s, c]], complex)
    else:
        m = np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)])
    return m.conj().T if dag else m


@pytest.mark.parametrize("prec,tol", [(E.C128, TOL), (E.C64, 2e-6)])
def test_controlled_forms_vs_oracle(prec, tol):
    n = 10
    g = np.random.default_rng(17)
    for trial in range(48):
        m = int(g.integers(1, 5))
        qs = g.permutation(n)[:m + 1].tolist()          # controls..., target last
        op = ["H", "X", "RX", "RY", "RZ", "U1"][trial % 6]
        th = float(g.uniform(0, 2 * math.pi))
        dag = int(g.integers(0, 2))
        psi = C.random_gaussian_state(n, trial)
        u = C.haar_unitary(C.Xoshiro256(900 + trial), 2)
        st = O.State(n)
        st.set(psi)
        st.controlled(op, qs, th, dag, u=u if op == "U1" else None)
        ctrl = E.mask(qs[:-1])
        t = qs[-1]
        with E.StateVector(n, prec) as sv:
            sv.set(psi)
            if op == "H":
                sv.h_gate(t, ctrl)
            elif op == "X":
                sv.x(t, ctrl)
            elif op == "RX":
                sv.rx(t, th, dag, ctrl)
            elif op == "RY":
                sv.ry(t, th, dag, ctrl)
            elif op == "U1":
                sv.u1(t, u, ctrl)
            else:
                sv.rz(t, th, dag, ctrl)
            got = sv.get()
        assert np.abs(got - st.psi).max() <= tol, (op, qs)


@pytest.mark.parametrize("prec,tol", [(E.C128, TOL), (E.C64, 2e-6)])
def test_two_three_qubit_gates_vs_oracle(prec, tol):
    n = 10
    g = np.random.default_rng(3)
    for trial in range(30):
        a, b, c = g.permutation(n)[:3].tolist()
        psi = C.random_gaussian_state(n, 100 + trial)
        st = O.State(n)
        st.set(psi)
        dag = trial % 2
        th = float(g.uniform(0, 2 * math.pi))
        u = C.haar_unitary(C.Xoshiro256(trial), 4)
        with E.StateVector(n, prec) as sv:
            sv.set(psi)
            kind = trial % 7
            if kind == 0:
                st.cnot(a, b); sv.x(b, 1 << a)
            elif kind == 1:
                st.cr(a, b, th, dag); sv.cr(a, b, th, dag)
            elif kind == 2:
                st.iswap(a, b, dag); sv.iswap(a, b, dag)
            elif kind == 3:
                st.sqiswap(a, b, dag); sv.sqiswap(a, b, dag)
            elif kind == 4:
                st.toffoli(a, b, c); sv.x(c, (1 << a) | (1 << b))
            elif kind == 5:
                st.cswap(a, b, c); sv.cswap(a, b, c)
            else:
                st.u2(a, b, u); sv.u2(a, b, u)
            got = sv.get()
        assert np.abs(got - st.psi).max() <= tol, trial


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_block_c128_vs_oracle(k):
    n = 11
    g = np.random.default_rng(40 + k)
    qs = g.permutation(n)[:k].tolist()
    u = C.haar_unitary(C.Xoshiro256(k), 1 << k)
    psi = C.random_gaussian_state(n, k)
    st = O.State(n)
    st.set(psi)
    st.block(qs, u)
    with E.StateVector(n) as sv:
        sv.set(psi)
        sv.block(qs, u)
        got = sv.get()
    assert np.abs(got - st.psi).max() <= TOL


@pytest.mark.parametrize("n", [3, 9, 14])
def test_readouts_vs_oracle(n):
    psi = C.random_gaussian_state(n, n)
    st = O.State(n)
    st.set(psi)
    g = np.random.default_rng(n)
    with E.StateVector(n) as sv:
        sv.set(psi)
        assert abs(sv.norm2() - st.norm2()) <= TOL
        for q in range(n):
            assert abs(sv.prob0(q) - st.prob0(q)) <= TOL
        assert abs(sv.expect_zz_chain() - st.expect_zz_chain()) <= TOL
        sets = [[0], [n - 1], list(range(n)), g.permutation(n)[:min(n, 5)].tolist(),
                list(range(n - 1, -1, -1))[: min(n, 7)]]
        for qs in sets:
            assert np.abs(sv.marginals(qs) - st.marginals(qs)).max() <= TOL, qs
        # pMeasure view: top 10, probability descending (CPUQuantumGates.cpp:178-187)
        qs = [2, 1, 0] if n >= 3 else [0]
        full = st.marginals(qs)
        top = sv.pmeasure(qs)
        assert len(top) == min(10, 1 << len(qs))
        probs = [p for _, p in top]
        assert probs == sorted(probs, reverse=True)
        for i, p in top:
            assert abs(full[i] - p) <= TOL


def test_measure_matches_oracle_including_tie_band():
    n = 10
    psi = C.random_gaussian_state(n, 77)
    with E.StateVector(n) as sv:
        for trial in range(40):
            st = O.State(n)
            st.set(psi)
            sv.set(psi)
            q = trial % n
            p0 = st.prob0(q)
            # draws straddling p0 at float resolution, plus generic draws
            r = [p0, float(np.float32(p0)), float(np.nextafter(np.float32(p0), np.float32(2))), 0.5,
                 0.999][trial % 5]
            o_ref = st.measure(q, r)
            o = sv.measure(q, r)
            assert o == o_ref
            assert np.abs(sv.get() - st.psi).max() <= TOL


def test_runs_are_bit_identical():
    circ = C.random_layers(14, 12, 5, haar_frac=0.1)
    outs = []
    for _ in range(2):
        with E.StateVector(14) as sv:
            sv.run(circ)
            outs.append(sv.get().copy())
            outs.append(np.array([sv.norm2(), sv.prob0(3), sv.expect_zz_chain()]))
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[3])


def test_haar_random_circuit_vs_oracle():
    circ = C.random_layers(12, 10, 11, haar_frac=0.1)
    assert circ.counts().get("U2", 0) > 0
    st, _ = O.run_circuit(circ)
    with E.StateVector(12) as sv:
        sv.run(circ)
        assert np.abs(sv.get() - st.psi).max() <= TOL


def test_invalid_arguments_are_rejected():
    with E.StateVector(5) as sv:
        for bad in (lambda: sv.x(5), lambda: sv.x(-1), lambda: sv.x(1, 1 << 1), lambda: sv.x(0, 1 << 7),
                    lambda: sv.cswap(1, 1, 2), lambda: sv.u2(2, 2, np.eye(4)),
                    lambda: sv.marginals([0, 0]), lambda: sv.block([0, 1, 1], np.eye(8))):
            with pytest.raises(E.QsvError):
                bad()
        assert abs(sv.norm2() - 1) < 1e-15
    with pytest.raises(E.QsvError):
        E.StateVector(0)


# ----------------------------------------------------- configuration 1 ----
def test_cfg1_ghz_26q():
    n = 26
    say("ghz", n)
    with E.StateVector(n) as sv:
        sv.run(C.ghz(n))
        psi = sv.get()
    sq2 = 0.707106781186548
    assert abs(psi[0] - sq2) <= TOL and abs(psi[-1] - sq2) <= TOL
    psi[0] = psi[-1] = 0
    assert np.abs(psi).max() == 0.0


def test_cfg1_qft_26q_vs_numpy_fft():
    n = 26
    say("qft", n)
    psi = C.random_gaussian_state(n, 2026)
    with E.StateVector(n) as sv:
        sv.set(psi)
        st = sv.run(C.qft(n), exact=True)
        out = sv.get()
    say("qft gates", st.gates, "device ms", st.device_ms)
    N = 1 << n
    ref = np.fft.ifft(psi) * math.sqrt(N)
    k = np.arange(N, dtype=np.uint64)
    rev = np.zeros(N, np.uint64)
    for b in range(n):
        rev |= ((k >> np.uint64(b)) & np.uint64(1)) << np.uint64(n - 1 - b)
    err = np.abs(out[rev] - ref).max()
    say("qft max err", err)
    assert err <= TOL


def test_30q_state_fits_and_normalises():
    n = 30
    say("30q c128")
    with E.StateVector(n) as sv:
        for q in range(n):
            sv.h_gate(q)
        sv.x(3, 1 << 29)
        assert abs(sv.norm2() - (2 * 0.707106781186548 ** 2) ** n) < 1e-12
        assert abs(sv.prob0(29) - 0.5 * (2 * 0.707106781186548 ** 2) ** n) < 1e-12
