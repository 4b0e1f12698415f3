"""Pins the CPU oracle (oracle/qsv_oracle.c) before anything is compared to it.

Checks, in order of strength:
  1. golden vectors written by QPanda's own CPU simulator (tests/golden/,
     generated by tests/golden/gen_golden.py from the compiled reference);
  2. the reference's own known-answer tests, testMain/testCPUQuantumGates.cpp,
     re-expressed (tolerance ZERO = 1e-12, :6);
  3. closed forms and independent routines: the Park-Miller KAT (CACM 31(10),
     1988), QFT = inverse DFT (numpy.fft) with bit reversal, GHZ, Kronecker-product
     brute force for dense blocks, <ZZ> from the probability vector.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from qpanda_amd import circuits as C

from conftest import GOLDEN

ZERO = 1e-12
PI = 3.141592653589793
SQ2 = 0.707106781186548


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


# ------------------------------------------------------------------ RNG ----
def test_park_miller_kat():
    s = 1
    for _ in range(10000):
        s = O.pm_next(s)
    assert s == 1043618065
    assert O.pm_jump(1, 10000) == 1043618065


@pytest.mark.parametrize("seed", [1, 12345, 2147483646, 987654321])
def test_park_miller_jump_matches_stepping(seed):
    s = seed
    for k in range(1, 300):
        s = O.pm_next(s)
        assert O.pm_jump(seed, k) == s


def test_seed_expression():
    # QuantumGates.cpp:68-72 with tm_year = S - 70 and zeros elsewhere gives S.
    assert O.pm_seed_from_tm(1234567 - 70, 0, 0, 0, 0, 0) == 1234567
    assert O.pm_seed_from_tm(118, 5, 3, 10, 20, 30) == 118 + 70 * (6 + 12 * (3 + 31 * (10 + 23 * (20 + 59 * 30))))


# --------------------------------------------------------- golden vectors --
@pytest.mark.parametrize("case", ["cfg0_random_n16_d20", "allkinds_n10", "ghz_n12", "prep_qft_n10",
                                  "midmeasure_n8", "hadamard_measure_n4"])
def test_oracle_state_matches_reference_golden(case):
    meta = manifest()[case]
    circ = C.Circuit.load(os.path.join(GOLDEN, case + ".qsvc"))
    ref = np.fromfile(os.path.join(GOLDEN, case + ".state"), np.complex128)
    # state before the first measurement
    kinds = circ.gates["kind"]
    first_meas = np.nonzero(kinds == C.MEASURE)[0]
    upto = int(first_meas[0]) if first_meas.size else circ.ngates
    st = O.State(circ.n)
    st.run(circ.prefix(upto), rng_state=meta["seed"])
    assert np.abs(st.psi - ref).max() <= ZERO


@pytest.mark.parametrize("case", ["cfg0_random_n16_d20", "midmeasure_n8", "hadamard_measure_n4"])
def test_oracle_shots_bit_exact_vs_reference(case):
    meta = manifest()[case]
    circ = C.Circuit.load(os.path.join(GOLDEN, case + ".qsvc"))
    ref = np.fromfile(os.path.join(GOLDEN, case + ".shots"), np.uint64)
    assert ref.size == meta["shots"]
    rng = meta["seed"]
    got = []
    outcomes = np.zeros(64, np.int32)
    for _ in range(meta["shots"]):
        st = O.State(circ.n)                      # endGate: back to |0..0>
        outcomes[:] = 0
        rng = st.run(circ, rng_state=rng, outcomes=outcomes)
        bits = 0
        for c in range(circ.n):
            bits |= int(outcomes[c] & 1) << c
        got.append(bits)
    assert np.array_equal(np.array(got, np.uint64), ref)


def test_shot_sampler_matches_reference_loop():
    meta = manifest()["cfg0_random_n16_d20"]
    circ = C.Circuit.load(os.path.join(GOLDEN, "cfg0_random_n16_d20.qsvc"))
    pre = np.fromfile(os.path.join(GOLDEN, "cfg0_random_n16_d20.state"), np.complex128)
    ref = np.fromfile(os.path.join(GOLDEN, "cfg0_random_n16_d20.shots"), np.uint64)
    ngates = int((circ.gates["kind"] != C.MEASURE).sum())
    bits, _ = O.sample_shots(pre, circ.n, list(range(circ.n)), ngates, meta["seed"], meta["shots"])
    assert np.array_equal(bits, ref)


def test_generator_reproduces_stored_circuit():
    circ = C.Circuit.load(os.path.join(GOLDEN, "cfg0_random_n16_d20.qsvc"))
    again = C.random_layers(16, 20, 20240601, measure_all=True)
    assert np.array_equal(circ.gates, again.gates)


def test_hadamard_measure_statistics():
    # testCPUQuantumGates.cpp:335-358: |p0_hat - 0.5| < 1e-2.
    shots = np.fromfile(os.path.join(GOLDEN, "hadamard_measure_n4.shots"), np.uint64)
    assert abs((shots == 0).mean() - 0.5) < 1e-2


# ----------------------------------------------- reference KATs restated --
def test_kat_hadamard():
    n = 10
    st = O.State(n)
    for i in range(n):
        st.h(i)
        assert abs(st.psi[0].real - SQ2) < ZERO and abs(st.psi[1 << i].real - SQ2) < ZERO
        st.h(i)
    for i in range(n):
        st.x(i)
        st.h(i)
        assert abs(st.psi[0].real - SQ2) < ZERO and abs(st.psi[1 << i].real + SQ2) < ZERO
        st.h(i)
        st.x(i)


@pytest.mark.parametrize("theta", [0.3, 1.7, 4.0])
def test_kat_rotations(theta):
    n = 10
    st = O.State(n)
    for i in range(n):  # testCPUQuantumGates.cpp:93-118
        st.rx(i, theta, 0)
        assert abs(st.psi[0].real - math.cos(theta / 2)) < ZERO
        assert abs(st.psi[1 << i].imag + math.sin(theta / 2)) < ZERO
        st.rx(i, theta, 1)
    for i in range(n):  # :120-145
        st.ry(i, theta, 0)
        assert abs(st.psi[0].real - math.cos(theta / 2)) < ZERO
        assert abs(st.psi[1 << i].real - math.sin(theta / 2)) < ZERO
        st.ry(i, theta, 1)
    for i in range(n):  # :147-176 pins diag(e^{-i t/2}, e^{i t/2})
        st.h(i)
        st.rz(i, theta, 0)
        assert abs(st.psi[0] - SQ2 * np.exp(-0.5j * theta)) < ZERO
        assert abs(st.psi[1 << i] - SQ2 * np.exp(0.5j * theta)) < ZERO
        st.rz(i, theta, 1)
        st.h(i)
    assert abs(st.psi[0] - 1) < 1e-12


def test_kat_cnot_cr_iswap_sqiswap():
    n = 6
    for a in range(n):
        for b in range(n):
            if a == b:
                continue
            st = O.State(n)  # CNOT, :178-210
            st.x(a)
            st.cnot(a, b)
            assert abs(st.psi[(1 << a) | (1 << b)] - 1) < ZERO
            st = O.State(n)  # CR, :212-248: e^{i t} on |11>
            st.x(a); st.x(b)
            st.cr(a, b, 0.9, 0)
            assert abs(st.psi[(1 << a) | (1 << b)] - np.exp(0.9j)) < ZERO
            st = O.State(n)  # iSWAP +i, :250-289
            st.x(a)
            st.iswap(a, b, 0)
            assert abs(st.psi[1 << b] - 1j) < ZERO
            st = O.State(n)  # sqiSWAP, :291-333
            st.x(a)
            st.sqiswap(a, b, 0)
            assert abs(st.psi[1 << a] - SQ2) < ZERO and abs(st.psi[1 << b] - 1j * SQ2) < ZERO


def test_kat_toffoli_truth_table():
    n = 4
    for c1 in range(n):
        for c2 in range(n):
            for t in range(n):
                if len({c1, c2, t}) < 3:
                    continue
                for v in range(8):
                    st = O.State(n)
                    bits = [(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1]
                    for q, bit in zip((c1, c2, t), bits):
                        if bit:
                            st.x(q)
                    st.toffoli(c1, c2, t)
                    exp_t = bits[2] ^ (bits[0] & bits[1])
                    idx = (bits[0] << c1) | (bits[1] << c2) | (exp_t << t)
                    assert abs(st.psi[idx] - 1) < ZERO


def test_kat_controlled_forms_match_masked_matrices():
    # :448-731 -- controls then target; compare with a brute-force controlled matrix.
    rng = np.random.default_rng(5)
    n = 5
    for op in ("H", "X", "RX", "RY", "RZ"):
        for _ in range(6):
            qs = rng.permutation(n)[:3].tolist()
            theta = float(rng.uniform(0, 2 * PI))
            dag = int(rng.integers(0, 2))
            psi = C.random_gaussian_state(n, int(rng.integers(1 << 30)))
            st = O.State(n)
            st.set(psi)
            st.controlled(op, qs, theta, dag)
            u = single_matrix(op, theta, dag)
            exp = apply_controlled_dense(psi, n, qs[:-1], qs[-1], u)
            assert np.abs(st.psi - exp).max() < ZERO


def test_kat_shor_order_finding():
    # testCPUQuantumGates.cpp:733-772: pMeasure{5,4} = 0.25 x 4.
    st = O.State(12)
    st.x(0)
    for i in range(4, 12):
        st.h(i)
    for i in range(4):
        st.cnot(i, 4)
    for c, a, b in ((4, 2, 1), (4, 2, 3), (4, 3, 0), (5, 2, 0), (5, 3, 1)):
        st.cswap(c, a, b)
    for i in range(11, 3, -1):
        st.h(i)
        for j in range(i - 1, 3, -1):
            st.cr(i, j, -PI / (1 << (i - j)), 0)
    p = st.marginals([5, 4])
    assert np.abs(p - 0.25).max() < ZERO


def test_kat_pmeasure_uniform():
    st = O.State(10)
    for i in range(10):
        st.h(i)
    p = st.marginals([2, 1, 3])
    assert np.abs(p - 1 / 8).max() < ZERO


# --------------------------------------------------------- closed forms ----
def test_ghz_closed_form():
    n = 12
    st, _ = O.run_circuit(C.ghz(n))
    exp = np.zeros(1 << n, complex)
    exp[0] = exp[-1] = SQ2
    assert np.abs(st.psi - exp).max() < ZERO


@pytest.mark.parametrize("n", [6, 11])
def test_qft_is_inverse_dft_bit_reversed(n):
    psi = C.random_gaussian_state(n, 100 + n)
    st = O.State(n, exact=True)
    st.set(psi)
    st.run(C.qft(n))
    N = 1 << n
    rev = np.array([int(format(k, f"0{n}b")[::-1], 2) for k in range(N)])
    assert np.abs(st.psi[rev] - np.fft.ifft(psi) * math.sqrt(N)).max() < ZERO


def kron_op(n, qubits, u):
    """Dense 2^n x 2^n operator of a k-qubit matrix (row bit j = bit(qubits[j]))."""
    N = 1 << n
    k = len(qubits)
    M = np.zeros((N, N), complex)
    for col in range(N):
        c_sub = sum(((col >> q) & 1) << j for j, q in enumerate(qubits))
        rest = col
        for q in qubits:
            rest &= ~(1 << q)
        for r_sub in range(1 << k):
            row = rest
            for j, q in enumerate(qubits):
                row |= ((r_sub >> j) & 1) << q
            M[row, col] += u[r_sub, c_sub]
    return M


def single_matrix(op, theta, dag):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    if op == "H":
        return np.array([[SQ2, SQ2], [SQ2, -SQ2]], complex)
    if op == "X":
        return np.array([[0, 1], [1, 0]], complex)
    if op == "RX":
        m = np.array([[c, -1j * s], [-1j * s, c]])
    elif op == "RY":
        m = np.array([[c, -s], [s, c]], complex)
    else:
        m = np.diag([np.exp(-0.5j * theta), np.exp(0.5j * theta)])
    return m.conj().T if dag else m


def apply_controlled_dense(psi, n, ctrls, t, u):
    out = psi.copy()
    for i in range(1 << n):
        if all((i >> c) & 1 for c in ctrls) and not (i >> t) & 1:
            j = i | (1 << t)
            out[i] = u[0, 0] * psi[i] + u[0, 1] * psi[j]
            out[j] = u[1, 0] * psi[i] + u[1, 1] * psi[j]
    return out


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_block_matches_kronecker(k):
    n = 7
    rng = C.Xoshiro256(k)
    u = C.haar_unitary(rng, 1 << k)
    g = np.random.default_rng(k)
    qubits = g.permutation(n)[:k].tolist()
    psi = C.random_gaussian_state(n, k)
    st = O.State(n)
    st.set(psi)
    st.block(qubits, u)
    assert np.abs(st.psi - kron_op(n, qubits, u) @ psi).max() < ZERO


def test_u2_convention_against_reference_gates():
    # row = 2*bit(qa) + bit(qb): CNOT(qa -> qb) is [[1,0,0,0],[0,1,0,0],[0,0,0,1],[0,0,1,0]].
    n = 6
    cnot = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], complex)
    psi = C.random_gaussian_state(n, 9)
    a = O.State(n); a.set(psi); a.u2(4, 1, cnot)
    b = O.State(n); b.set(psi); b.cnot(4, 1)
    assert np.abs(a.psi - b.psi).max() < ZERO
    # kron(RX on qa, RY on qb) equals RX then RY.
    u = np.kron(single_matrix("RX", 0.7, 0), single_matrix("RY", 1.3, 0))
    a = O.State(n); a.set(psi); a.u2(2, 5, u)
    b = O.State(n); b.set(psi); b.rx(2, 0.7); b.ry(5, 1.3)
    assert np.abs(a.psi - b.psi).max() < ZERO


def test_expect_zz_closed_forms():
    n = 9
    st = O.State(n)
    assert abs(st.expect_zz_chain() - (n - 1)) < ZERO
    st.x(0)
    assert abs(st.expect_zz_chain() - (n - 3)) < ZERO
    st = O.State(n)
    for q in range(n):
        st.h(q)
    assert abs(st.expect_zz_chain()) < 1e-12
    st, _ = O.run_circuit(C.ghz(n))
    assert abs(st.expect_zz_chain() - (n - 1) * 2 * SQ2 * SQ2) < ZERO
    psi = C.random_gaussian_state(n, 4)
    st = O.State(n); st.set(psi)
    p = np.abs(psi) ** 2
    x = np.arange(1 << n)
    e = sum(np.where(((x >> i) ^ (x >> (i + 1))) & 1, -1.0, 1.0) for i in range(n - 1))
    assert abs(st.expect_zz_chain() - float(p @ e)) < 1e-12


def test_measure_collapse_renormalises():
    n = 6
    psi = C.random_gaussian_state(n, 21)
    st = O.State(n); st.set(psi)
    p0 = st.prob0(3)
    out = st.measure(3, 0.999)  # (float)0.999 > p0 unless p0 is ~1
    assert out == int(np.float32(0.999) > p0)
    assert abs(st.norm2() - 1) < 1e-12
    assert np.all(st.psi[(np.arange(1 << n) >> 3 & 1) != out] == 0)


# ------------------------------------------- general 2x2 and exact modes ----
def _random_2x2(g, unitary):
    if unitary:
        return C.haar_unitary(C.Xoshiro256(int(g.integers(1 << 30))), 2)
    return g.standard_normal((2, 2)) + 1j * g.standard_normal((2, 2))


@pytest.mark.parametrize("unitary", [True, False])
def test_u1_matches_kronecker_and_k1_block(unitary):
    """orc_u1 (the general 2x2, a correct unitarySingleQubitGate; the
    reference's, CPUQuantumGates.cpp:1170-1212, addresses qubit qn - 1) against
    the dense Kronecker operator and the independent k = 1 block loop, on
    unitary and on general complex matrices (a transposed or conjugated
    entry fails both)."""
    g = np.random.default_rng(31 + unitary)
    n = 6
    for q in range(n):
        u = _random_2x2(g, unitary)
        psi = C.random_gaussian_state(n, 40 + q)
        a = O.State(n); a.set(psi); a.u1(q, u)
        assert np.abs(a.psi - kron_op(n, [q], u) @ psi).max() < 1e-12
        b = O.State(n); b.set(psi); b.block([q], u)
        assert np.abs(a.psi - b.psi).max() < 1e-12


def test_controlled_u1_matches_masked_matrix():
    """Controlled U1 (orc_ctrl_1q op 5, controls then target as in
    Instruction.cpp:62) against the brute-force controlled matrix."""
    g = np.random.default_rng(77)
    n = 6
    for trial in range(12):
        m = int(g.integers(1, 4))
        qs = g.permutation(n)[:m + 1].tolist()
        u = _random_2x2(g, trial % 2 == 0)
        psi = C.random_gaussian_state(n, 500 + trial)
        st = O.State(n); st.set(psi)
        st.controlled("U1", qs, u=u)
        exp = apply_controlled_dense(psi, n, qs[:-1], qs[-1], u)
        assert np.abs(st.psi - exp).max() < 1e-12


def test_exact_mode_cz_sqiswap_h_closed_forms():
    """The ground-truth mode's exact gates against Kronecker operators of
    their closed-form matrices: CZ = diag(1, 1, 1, -1); sqiSWAP with 1/sqrt(2)
    and +i (dagger -i) on {|01>, |10>} (CPUQuantumGates.cpp:978-1039 with the
    exact constant); H with 1/sqrt(2).  The same records run through the
    oracle's circuit runner in exact mode."""
    n = 5
    r = 1 / math.sqrt(2)
    cz = np.diag([1, 1, 1, -1]).astype(complex)
    for a in range(n):
        for b in range(n):
            if a == b:
                continue
            psi = C.random_gaussian_state(n, 10 * a + b)
            st = O.State(n, exact=True); st.set(psi); st.cz_exact(a, b)
            assert np.abs(st.psi - kron_op(n, [a, b], cz) @ psi).max() < 1e-13
            for dag in (0, 1):
                k = -1j if dag else 1j
                sq = np.array([[1, 0, 0, 0], [0, r, k * r, 0], [0, k * r, r, 0], [0, 0, 0, 1]], complex)
                st = O.State(n, exact=True); st.set(psi); st.sqiswap(a, b, dag)
                assert np.abs(st.psi - kron_op(n, [a, b], sq) @ psi).max() < 1e-13
    h = np.array([[r, r], [r, -r]], complex)
    psi = C.random_gaussian_state(n, 99)
    st = O.State(n, exact=True); st.set(psi); st.h(2)
    assert np.abs(st.psi - kron_op(n, [2], h) @ psi).max() < 1e-13
    # circuit records in exact mode: CZ(1,3), SQISWAP(0,4), SQISWAP+(2,1), H(3)
    b = C._Builder(n)
    b.add(C.CZ, 1, 3)
    b.add(C.SQISWAP, 0, 4)
    b.add(C.SQISWAP, 2, 1, dagger=1)
    b.add(C.H, 3)
    st = O.State(n, exact=True); st.set(psi); st.run(b.build())
    sq = lambda k: np.array([[1, 0, 0, 0], [0, r, k * r, 0], [0, k * r, r, 0], [0, 0, 0, 1]], complex)
    exp = kron_op(n, [3], h) @ (kron_op(n, [2, 1], sq(-1j)) @ (kron_op(n, [0, 4], sq(1j)) @ (kron_op(n, [1, 3], cz) @ psi)))
    assert np.abs(st.psi - exp).max() < 1e-13
    # plain mode: CZ runs as CR(pi), which leaves the sin(pi) residue the
    # exact mode removes
    st2 = O.State(n); st2.set(psi); st2.run(b.build())
    assert 1e-17 < np.abs(st2.psi - st.psi).max() < 1e-12


def test_u1_records_through_the_circuit_runner():
    """U1 records (matrix pool entries) in the oracle's circuit runner equal
    the Kronecker chain of their matrices."""
    n = 5
    circ = C.all_gate_kinds(n, 8, reps=1, u1=True)
    sel = circ.gates["kind"] == C.U1
    assert sel.sum() == 2
    b = C._Builder(n)
    for g in circ.gates[sel]:
        b.add(C.U1, int(g["q0"]), mat=circ.mats[g["mat"]:g["mat"] + 4].reshape(2, 2))
    psi = C.random_gaussian_state(n, 3)
    st = O.State(n); st.set(psi); st.run(b.build())
    exp = psi
    for g in circ.gates[sel]:
        exp = kron_op(n, [int(g["q0"])], circ.mats[g["mat"]:g["mat"] + 4].reshape(2, 2)) @ exp
    assert np.abs(st.psi - exp).max() < 1e-12
