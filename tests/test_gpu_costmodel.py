"""The cost-model kernels (forward, backward, Adam on the device) against the
FP64 oracle (oracle/mlp_oracle.c, itself pinned by finite differences), the
engine's timing log, the learned scheduler's independence of results, and the
trained network's measured gain over the greedy schedule it replaces."""
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


def batch(seed, n=300):
    g = np.random.default_rng(seed)
    x = g.uniform(0, 1, (n, 12)).astype(np.float32)
    t = (np.sin(3 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * x[:, 7]).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    m = E.CostModel(11)
    x, _ = batch(1)
    y = m.forward(x)
    ref = O.mlp_forward(m.get_params().astype(np.float64), x.astype(np.float64))
    assert np.abs(y - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())


def test_gradients_match_oracle():
    m = E.CostModel(12)
    x, t = batch(2)
    g, loss = m.gradients(x, t)
    gref, lref = O.mlp_grad(m.get_params().astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    scale = np.abs(gref).max()
    err = np.abs(g - gref).max()
    say("max |grad - oracle|", err, "scale", scale)
    assert err < 1e-5 * scale + 1e-7
    assert abs(loss - lref) < 1e-5 * max(1.0, lref)


def test_device_adam_steps_match_oracle():
    # three full-batch Adam steps on the device against the FP64 restatement
    m = E.CostModel(21)
    x, t = batch(5, 300)
    p0 = m.get_params().astype(np.float64)
    m.train(x, t, 3, 3e-3)
    got = m.get_params().astype(np.float64)
    lr = 3e-3
    ref = O.mlp_train(p0, x.astype(np.float64), t.astype(np.float64), 3, lr)
    # where the gradient is tiny, g / (|g| + eps) is ill-conditioned between the
    # device's FP32 gradient and the FP64 one: check those only against Adam's
    # step bound (lr (1 - b1) / sqrt(1 - b2) ~ 3.2 lr per step); the rest to 1%
    # of one step
    g0, _ = O.mlp_grad(p0, x.astype(np.float64), t.astype(np.float64))
    well = np.abs(g0) > 1e-3 * np.abs(g0).max()
    err = np.abs(got - ref)
    say("max |params - oracle| after 3 Adam steps", err[well].max(), "on", int(well.sum()), "of", err.size)
    assert well.sum() > err.size // 4
    assert err[well].max() < 3e-5
    assert np.abs(got - p0).max() <= 3 * 3.17 * lr


def test_training_on_device_reduces_loss():
    m = E.CostModel(13)
    x, t = batch(3, 600)
    l0 = m.train(x, t, 0)
    l1 = m.train(x, t, 800, 3e-3)
    say("loss", l0, "->", l1)
    assert l1 < 0.1 * l0
    # the loss the device reports is the oracle's loss at the device's parameters
    lref = O.mlp_loss(m.get_params().astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(l1 - lref) < 1e-4 * max(1.0, lref)


def test_timing_log_has_one_record_per_block():
    n = 20
    circ = C.random_layers(n, 8, 3, haar_frac=0.1)
    with E.StateVector(n, E.C64) as sv:
        st = sv.run(circ, fuse_k=4, profile=True)
        f, ms = sv.timing_log()
    assert f.shape == (st.blocks, 12) and ms.shape == (st.blocks,)
    assert np.all(ms > 0)


def test_learned_schedule_does_not_change_results():
    n = 18
    circ = C.random_layers(n, 12, 21, haar_frac=0.1)
    m = E.CostModel(99)                      # untrained: arbitrary choices
    with E.StateVector(n, E.C128) as ref, E.StateVector(n, E.C64) as a, E.StateVector(n, E.C64) as b:
        ref.run(circ)
        a.attach_costmodel(m)
        st = a.run(circ, fuse_k=4, scheduler=1)
        b.run(circ, fuse_k=4, scheduler=0)
        fa, fb = ref.fidelity(a), ref.fidelity(b)
    say("learned blocks", st.blocks, "fidelity", fa, "greedy", fb)
    assert fa >= 1 - 1e-6 and fb >= 1 - 1e-6


@pytest.mark.skipif(not os.path.exists(E.DEFAULT_COSTMODEL), reason="no trained cost model in the tree")
def test_learned_window_plans_across_measurements():
    """The learned scheduler plans its candidates as shapes and rebuilds the
    winner with matrices, replaying the window plan's qubit masks; with
    MEASURE and RESET records the windows restart per segment.  Same random
    stream as the plain FP64 path (one draw per record), same outcomes here,
    and the state at the fused-path bar."""
    n = 20
    b = C._Builder(n)
    for seed, rec in ((31, (C.MEASURE, 5, 0)), (32, (C.RESET, 11)), (33, None)):
        for g in C.random_layers(n, 8, seed).gates:
            b.add(int(g["kind"]), int(g["q0"]), int(g["q1"]), theta=float(g["theta"]))
        if rec:
            b.add(*rec)
    circ = b.build()
    out_f = np.zeros(8, np.int32)
    out_p = np.zeros(8, np.int32)
    with E.StateVector(n, E.C64) as f, E.StateVector(n, E.C128) as p:
        f.attach_costmodel(E.CostModel.load())
        st = f.run(circ, fuse_k=4, split=3, scheduler=1, rng_state=4242, outcomes=out_f)
        p.run(circ, rng_state=4242, outcomes=out_p)
        say("learned planner", st.planner, "k", st.fuse_k, "blocks", st.blocks, "outcomes", out_f[:2], out_p[:2])
        assert st.planner == 1                   # window plans, rebuilt from replayed masks
        assert f.last_rng == p.last_rng
        assert np.array_equal(out_f[:2], out_p[:2])
        assert f.fidelity(p) >= 1 - 1e-6


@pytest.mark.skipif(not os.path.exists(E.DEFAULT_COSTMODEL), reason="no trained cost model in the tree")
@pytest.mark.parametrize("split", [2, 3])
def test_trained_model_gain_over_greedy(split):
    m = E.CostModel.load()
    rows = []
    for circ in (C.random_layers(28, 24, 4242, haar_frac=0.1), C.hea(28, 6, 77), C.qft(26)):
        n = circ.n
        with E.StateVector(n, E.C64) as a, E.StateVector(n, E.C64) as b:
            a.attach_costmodel(m)
            tl, tg = [], []
            for _ in range(3):
                a.reset()
                tl.append(a.run(circ, fuse_k=4, split=split, scheduler=1).device_ms)
                b.reset()
                tg.append(b.run(circ, fuse_k=4, split=split, scheduler=0).device_ms)
            fid = a.fidelity(b)
        rows.append((circ.name, np.median(tg), np.median(tl), fid))
        say(f"{circ.name}: greedy k=4 {np.median(tg):.1f} ms, learned {np.median(tl):.1f} ms, "
            f"gain {np.median(tg) / np.median(tl):.3f}x, fidelity {fid:.12f}")
    # the learned schedule groups blocks into shared sweeps where the model
    # predicts a gain; greedy k = 4 sweeps once per block (held-out circuits:
    # 2.1-4.6x when the model was trained)
    gains = [tg / tl for _, tg, tl, _ in rows]
    for (name, tg, tl, fid), g in zip(rows, gains):
        assert g >= 1.5, name
        assert fid >= 1 - 1e-6
    assert np.exp(np.mean(np.log(gains))) >= 2.0


@pytest.mark.skipif(not os.path.exists(E.DEFAULT_COSTMODEL), reason="no trained cost model in the tree")
def test_trained_model_against_the_heuristic_schedulers():
    """Held-out 28-qubit circuits, five repetitions round-robin over the
    methods (device drift hits every method alike; run-to-run spread of the
    medians < 1%).  Random circuit: the learned schedule beats greedy maximal
    grouping of k = 4 blocks and is the window planner's schedule up to one
    tile-size choice (within 2%).  Hardware-efficient ansatz: the window
    planner puts 16 blocks in its first sweep; the model splits that window
    into two sweeps of 12 and 13 tile bits and folds the lone block after it
    into the second (9.4 ms against 10.0 ms when trained), so it must be at
    least 3% faster than the window heuristic there."""
    m = E.CostModel.load()
    methods = (("passes", dict(scheduler=2, group=8)), ("window", dict(scheduler=3)),
               ("learned", dict(scheduler=1)))
    res = {}
    for circ in (C.random_layers(28, 24, 5151, haar_frac=0.1), C.hea(28, 6, 77)):
        ms = {name: [] for name, _ in methods}
        with E.StateVector(28, E.C64) as sv:
            sv.attach_costmodel(m)
            for _ in range(5):
                for name, kw in methods:
                    sv.reset()
                    st = sv.run(circ, fuse_k=4, split=3, **kw)
                    ms[name].append(st.device_ms)
        t = {name: float(np.median(v)) for name, v in ms.items()}
        say("%s: passes %.2f ms, window %.2f ms, learned %.2f ms" % (circ.name, t["passes"], t["window"], t["learned"]))
        res[circ.name] = t
    r = res["random_n28_d24_s5151"]
    assert r["learned"] < r["passes"]
    assert r["learned"] <= r["window"] * 1.02
    h = res["hea_n28_l6_s77"]
    assert h["learned"] * 1.03 <= h["window"]
