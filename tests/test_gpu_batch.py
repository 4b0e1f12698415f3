"""Batched variational circuits in one persistent launch, against the FP64 oracle.

Each instance's sum of <Z_q Z_{q+1}> must match the oracle's value for the same
gate list (tests restate the instance with qpanda_amd.engine.hea_circuit) within
the complex64 state's rounding (absolute 2e-4 on a sum of up to 21 terms).
Instances of up to 14 qubits run inside LDS, 15..22 through the two-pass HBM
path.  hipGraph replays must reproduce the eager results exactly, replay after
replay."""
import sys

import numpy as np
import pytest

import oracle as O
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_zz(n, angles):
    st = O.State(n)
    st.run(E.hea_circuit(n, angles))
    return st.expect_zz_chain()


def test_batch_matches_oracle_lds_and_hbm_paths():
    nq, ly, ang = E.hea_batch(40, 7, nmin=10, nmax=17)
    nq[0], ly[0] = 17, 8                     # make sure both paths and the largest sizes appear
    ang[0] = np.random.default_rng(1).uniform(-np.pi, np.pi, (8, 17, 2))
    nq[1], ly[1] = 2, 3
    ang[1] = np.random.default_rng(2).uniform(-np.pi, np.pi, (3, 2, 2))
    say("batch sizes", sorted(set(nq)))
    b = E.Batch(nq, ly, ang)
    zz, lat, ms = b.run(512)
    ref = np.array([oracle_zz(n, a) for n, a in zip(nq, ang)])
    err = np.abs(zz - ref).max()
    say("max |zz - oracle|", err, "batch ms", ms, "p50 latency us", np.median(lat))
    assert err < 2e-4
    assert np.all(lat > 0)


def test_batch_thread_counts_agree_and_autotune():
    nq, ly, ang = E.hea_batch(64, 11, nmin=10, nmax=16)
    b = E.Batch(nq, ly, ang)
    outs = [b.run(t)[0] for t in (256, 512, 1024)]
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() < 1e-5
    best, ms = b.autotune()
    say("autotuned threads", best, "ms", ms)
    assert best in (256, 512, 1024) and ms > 0


def test_graph_replays_match_eager():
    nq, ly, ang = E.hea_batch(96, 13, nmin=10, nmax=16)
    b = E.Batch(nq, ly, ang)
    eager, _, _ = b.run(512)
    for _ in range(3):
        zz, ms = b.graph_run(replays=2)
        assert np.array_equal(zz, eager)
        eager2, _, _ = b.run(512)          # eager work between replays
        assert np.array_equal(eager2, eager)


def test_full_configuration_4096_instances():
    nq, ly, ang = E.hea_batch(4096, 2024)
    b = E.Batch(nq, ly, ang)
    best, _ = b.autotune()
    zz, lat, ms = b.run()
    zg, gms = b.graph_run(replays=3)
    say(f"4096 instances (n 10..22): {ms:.2f} ms eager, {gms:.2f} ms per graph replay, threads {best}, "
        f"p50 {np.percentile(lat, 50):.1f} us, p99 {np.percentile(lat, 99):.1f} us")
    assert np.array_equal(zg, zz)
    # spot-check the smallest instances of each size against the oracle (n <= 16)
    for n in range(10, 17):
        idx = [i for i in range(len(nq)) if nq[i] == n]
        if idx:
            i = idx[0]
            assert abs(zz[i] - oracle_zz(n, ang[i])) < 2e-4


def test_two_pass_path_at_the_largest_sizes():
    """The second pass's tile holds qubits 13..n-1 plus the 27 - n lowest (5 of
    them at 22 qubits).  Every size of the HBM path, and the LDS path's
    largest, against the FP64 oracle at every thread count."""
    nq = [15, 16, 18, 19, 20, 21, 22, 14]
    ly = [3, 2, 2, 1, 2, 1, 2, 2]
    g = np.random.default_rng(31)
    ang = [g.uniform(-np.pi, np.pi, (l, n, 2)) for n, l in zip(nq, ly)]
    b = E.Batch(nq, ly, ang)
    for t in (256, 512, 1024):
        zz, _, _ = b.run(t)
        say("threads", t, "sizes", nq)
        for n, a, z in zip(nq, ang, zz):
            assert abs(z - oracle_zz(n, a)) < 2e-4, (t, n)


def test_split_instances_match_oracle_and_replay_exactly():
    """Instances of 20, 21 and 22 qubits run as 2, 4 and 8 cooperating
    workgroups (a per-instance barrier after every pass, per-part ZZ sums added
    in part order): oracle values, and bit-identical eager and graph results."""
    nq = [22, 21, 20, 22, 12, 17]
    ly = [3, 2, 1, 1, 4, 2]
    g = np.random.default_rng(77)
    ang = [g.uniform(-np.pi, np.pi, (l, n, 2)) for n, l in zip(nq, ly)]
    b = E.Batch(nq, ly, ang)
    zz, lat, _ = b.run(512)                 # the graph's thread count (untuned: 512)
    say("split sizes", nq, "latency us", lat)
    for n, a, z in zip(nq, ang, zz):
        assert abs(z - oracle_zz(n, a)) < 2e-4, n
    for _ in range(2):
        zg, _ = b.graph_run(replays=2)
        assert np.array_equal(zg, zz)
        assert np.array_equal(b.run(512)[0], zz)
