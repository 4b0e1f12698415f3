"""Shot sampling without a state copy per shot (qsv_sample_shots) against the
per-shot copy path shaped as the reference's loop (qsv_sample_shots_copy),
which the golden fixtures pin bit-exactly (test_gpu_parity.test_shots_bit_exact).

Bar: identical outcome words and final generator state for every case -- all
qubits in order, reversed, random orders and subsets, gate-draw skips, states
where a decision sits on the threshold (forcing the fallback), and complex64
states (compared with the copy path on their complex128 conversion).  Cases
with more than 16 measured qubits exercise the per-shot LDS levels below the
marginal tree."""
import math
import sys

import numpy as np
import pytest

from qpanda_amd import circuits as C
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def prepared(n, seed, prec=E.C128, depth=6):
    sv = E.StateVector(n, prec)
    sv.run(C.random_layers(n, depth, seed, haar_frac=0.1), fuse_k=4 if prec == E.C64 else 0)
    return sv


ORDERS = {
    "ascending": lambda n, g: list(range(n)),
    "descending": lambda n, g: list(range(n - 1, -1, -1)),
    "permuted": lambda n, g: [int(x) for x in g.permutation(n)],
    "subset": lambda n, g: [int(x) for x in g.permutation(n)[: max(1, n // 2)]],
}


@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("n", [6, 12, 18, 20])
def test_tree_equals_copy(n, order):
    g = np.random.default_rng(n * 7 + len(order))
    qs = ORDERS[order](n, g)
    shots = 300 if n <= 12 else 120
    say("sampler", n, order, "m", len(qs), "shots", shots)
    with prepared(n, 40 + n) as sv:
        a, ra, ta = sv.sample_shots(qs, 37, 12345, shots, method="tree")
        b, rb, tb = sv.sample_shots(qs, 37, 12345, shots, method="copy")
    say("ties tree", ta, "copy", tb)
    assert ra == rb
    assert np.array_equal(a, b)


def test_tree_many_shots_distribution():
    # 20000 shots of a 10-qubit state: outcome frequencies match |a|^2 (chi-square)
    n = 10
    with prepared(n, 5) as sv:
        p = np.abs(sv.get()) ** 2
        w, _, _ = sv.sample_shots(list(range(n)), 0, 777, 20000)
    counts = np.bincount(w.astype(np.int64), minlength=1 << n)
    exp = p / p.sum() * len(w)
    keep = exp > 5
    chi2 = float(np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep]))
    dof = int(keep.sum()) - 1
    say("chi2", chi2, "dof", dof)
    assert chi2 < dof + 6 * math.sqrt(2 * dof)


def test_threshold_decision_falls_back():
    # P(q0 = 0) set to the first measurement draw of the first shot: that
    # decision is inside the tie band, so the shot is re-run by the copy path
    rng0 = 424242
    r = E.pm_next(rng0)
    fr = float(np.float32(r / 2147483647.0))
    theta = 2 * math.acos(math.sqrt(fr))
    b = C._Builder(3)
    b.add(C.RY, 0, theta=theta)
    b.add(C.H, 1)
    b.add(C.CNOT, 1, 2)
    with E.StateVector(3) as sv:
        sv.run(b.build())
        a, ra, ta = sv.sample_shots([0, 1, 2], 0, rng0, 50, method="tree")
        c, rc, tc = sv.sample_shots([0, 1, 2], 0, rng0, 50, method="copy")
    say("ties", ta, tc)
    assert ta >= 1
    assert ra == rc and np.array_equal(a, c)


@pytest.mark.parametrize("n", [14, 19])
def test_complex64_matches_fp64_copy_path(n):
    qs = list(range(n - 1, -1, -1))
    with prepared(n, 90 + n, prec=E.C64) as s64, E.StateVector(n, E.C128) as s128:
        s128.copy_from(s64)
        a, ra, _ = s64.sample_shots(qs, 5, 999, 150, method="tree")
        b, rb, _ = s128.sample_shots(qs, 5, 999, 150, method="copy")
    assert ra == rb and np.array_equal(a, b)


def test_state_unchanged_and_zero_shots():
    n = 8
    with prepared(n, 3) as sv:
        before = sv.get()
        w, r, _ = sv.sample_shots(list(range(n)), 3, 55, 0)
        assert w.shape == (0,) and r == 55
        sv.sample_shots(list(range(n)), 3, 55, 40)
        assert np.array_equal(sv.get(), before)
        with pytest.raises(E.QsvError):
            sv.sample_shots([0, 0], 0, 1, 4)


def test_gather_matches_full_readout():
    n = 11
    with prepared(n, 8) as s128, prepared(n, 8, prec=E.C64) as s64:
        idx = np.array([0, 1, 5, 1023, 2047, 77, 77], np.uint64)
        assert np.array_equal(s128.gather(idx), s128.get()[idx.astype(np.int64)])
        assert np.array_equal(s64.gather(idx), s64.get()[idx.astype(np.int64)])
        with pytest.raises(E.QsvError):
            s128.gather(np.array([1 << n], np.uint64))
