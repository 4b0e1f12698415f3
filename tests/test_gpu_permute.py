"""Qubit permutation (qsv_permute_qubits): the amplitude of basis state x
moves to the index whose bit perm[q] is bit q of x.  The sweeps only move
data, so the result must equal numpy's index permutation bit for bit, for
permutations that keep qubits 0..3, exchange them with high qubits, and
permute everything (cycles longer than one 13-qubit tile, which take several
sweeps)."""
import sys

import numpy as np
import pytest

from qpanda_amd import circuits as C
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected(psi, perm):
    n = len(perm)
    idx = np.arange(1 << n, dtype=np.int64)
    dst = np.zeros_like(idx)
    for q in range(n):
        dst |= ((idx >> q) & 1) << perm[q]
    out = np.empty_like(psi)
    out[dst] = psi
    return out


def c64_state(n, seed):
    # complex64-representable amplitudes, so set/get round-trips exactly
    psi = C.random_gaussian_state(n, seed)
    return psi.astype(np.complex64).astype(np.complex128)


def perms(n, rng):
    yield "identity", list(range(n))
    p = list(range(n))
    for j in range(4):
        p[j], p[n - 1 - j] = p[n - 1 - j], p[j]
    yield "low<->high", p
    p = list(range(n))
    p[0], p[5] = p[5], p[0]
    yield "swap 0,5", p
    yield "rotate", [(q + 1) % n for q in range(n)]
    yield "random", list(rng.permutation(n))
    yield "random high", list(range(4)) + list(4 + rng.permutation(n - 4))


@pytest.mark.parametrize("n", [13, 14, 18, 22])
def test_permute_qubits_bit_exact(n):
    rng = np.random.default_rng(n)
    psi = c64_state(n, 100 + n)
    with E.StateVector(n, E.C64) as sv:
        for name, p in perms(n, rng):
            say("permute", n, name, p)
            sv.set(psi)
            sv.permute_qubits(p)
            got = sv.get()
            assert np.array_equal(got, expected(psi, p)), (n, name)
            inv = np.argsort(p)
            sv.permute_qubits(inv)
            assert np.array_equal(sv.get(), psi), (n, name, "inverse")


def test_permute_qubits_rejects_bad_input():
    with E.StateVector(14, E.C64) as sv:
        with pytest.raises(RuntimeError):
            sv.permute_qubits([0] * 14)
        with pytest.raises(RuntimeError):
            sv.permute_qubits(list(range(1, 15)))
    with E.StateVector(12, E.C64) as sv:
        with pytest.raises(RuntimeError):
            sv.permute_qubits(list(range(12)))
    with E.StateVector(14, E.C128) as sv:
        with pytest.raises(RuntimeError):
            sv.permute_qubits(list(range(14)))
