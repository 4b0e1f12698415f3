"""Train the fused-path cost model from timings the engine records itself.

    python -m qpanda_amd.train_costmodel [--out qpanda_amd/costmodel_params.f32]
                                         [--report report.json] [--quick]
                                         [--evaluate PARAMS]

1. Collect: seeded circuits (random layers, QFT, hardware-efficient ansatz) at
   qubit counts 20..30 (and 32, with the products and shapes that matter
   there) run on the fused path under every decision the
   scheduler can make -- block size k = 2..5, the four products (FP32, FP32
   + BF16 lo correction, BF16x6, FP16x3), single-block launches at every tile
   shape (6, 7, 8 column bits), multi-block sweeps of up to 2, 3, 4, 8 or 16
   blocks at minimal / 12 / 13 tile bits, and the window planner's sweeps
   (k = 4, 5; 12 and 13 tile bits) -- with per-launch event timing; the
   engine's timing log gives one (features, ms) record per launch.
2. Train: full-batch Adam on the device (the engine's own forward / backward
   kernels), target log(ms * 2^(30 - n)).
3. Evaluate on held-out circuits: per-launch prediction error, and the measured
   device time of the learned schedule against the greedy k = 4 schedule (one
   block per sweep; the method without the model), against greedy maximal
   grouping and against the window planner with one sweep per window; the
   states of all schedules agree to FP32 rounding.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import circuits as C
from . import engine as E

SINGLE = [dict(fuse_k=k, split=s, tile=t, scheduler=0) for k in (2, 3, 4, 5) for s in (0, 1, 2, 3)
          for t in (6, 7, 8)]
PASSES = [dict(fuse_k=k, split=s, scheduler=2, group=g, pass_tb=tb) for k in (3, 4, 5) for s in (1, 2, 3)
          for g in (2, 3, 4, 8, 16) for tb in (0, 12, 13)]
WINDOWS = [dict(fuse_k=k, split=s, scheduler=3, pass_tb=tb) for k in (4, 5) for s in (1, 2, 3) for tb in (12, 13)]
# above 30 qubits (32 GiB states) only the decisions that matter at that size:
# the FP16x3 / BF16x6 products, k = 4 and 5, full tiles
LARGE = ([dict(fuse_k=k, split=s, tile=7, scheduler=0) for k in (4, 5) for s in (2, 3)]
         + [dict(fuse_k=k, split=s, scheduler=2, group=g, pass_tb=tb) for k in (4, 5) for s in (2, 3)
            for g in (4, 8, 16) for tb in (0, 13)]
         + [dict(fuse_k=k, split=s, scheduler=3, pass_tb=13) for k in (4, 5) for s in (2, 3)])


def workloads(ns, seed0, depth):
    out = []
    for i, n in enumerate(ns):
        out.append(C.random_layers(n, depth, seed0 + i, haar_frac=0.1))
        out.append(C.hea(n, 3, seed0 + 100 + i))
        if n <= 26:
            out.append(C.qft(n))
    return out


def collect(ns, seed0=7, depth=6, log=print):
    """Run every workload under every configuration; return features, log-targets."""
    feats, targ = [], []
    by_n = {}
    for circ in workloads(ns, seed0, depth):
        by_n.setdefault(circ.n, []).append(circ)
    for n, circs in sorted(by_n.items()):
        with E.StateVector(n, E.C64) as sv:
            for circ in circs:
                for cfg in (SINGLE + PASSES + WINDOWS) if n <= 30 else LARGE:
                    sv.reset()
                    sv.run(circ, **cfg)                           # warm
                    sv.reset()
                    sv.run(circ, profile=True, **cfg)
                    f, ms = sv.timing_log()
                    keep = ms > 0
                    feats.append(f[keep])
                    targ.append(np.log(ms[keep].astype(np.float64) * 2.0 ** (30 - n)))
                log(f"  {circ.name}: {sum(x.shape[0] for x in feats)} records so far")
        log(f"collected n={n}: {sum(x.shape[0] for x in feats)} records so far")
    return np.concatenate(feats).astype(np.float32), np.concatenate(targ).astype(np.float32)


def train(x, t, epochs=4000, lr=3e-3, seed=1, log=print):
    m = E.CostModel(seed)
    p = m.get_params()
    p[-1] = float(np.mean(t))                    # output bias at the mean target
    m.set_params(p)
    loss0 = m.train(x, t, 0)
    loss = m.train(x, t, epochs, lr)
    log(f"trained on {x.shape[0]} records: loss {loss0:.4f} -> {loss:.5f}")
    return m, loss0, loss


METHODS = (("greedy_k4", dict(fuse_k=4, scheduler=0)),
           ("passes_k4", dict(fuse_k=4, scheduler=2, group=8)),
           ("window_k4", dict(fuse_k=4, scheduler=3)),
           ("learned_fixed_low", dict(fuse_k=4, scheduler=1, fixed_low=True)),
           ("learned", dict(fuse_k=4, scheduler=1)))


def compare(m, circ, reps=3, split=3):
    """Median device ms of each method on one circuit, plus fidelities to greedy."""
    n = circ.n
    res = {}
    with E.StateVector(n, E.C64) as ref, E.StateVector(n, E.C64) as sv:
        ref.run(circ, fuse_k=4, split=split)
        sv.attach_costmodel(m)
        # repetitions round-robin over the methods, so that drift of the
        # device's speed over the run affects every method alike
        times = {name: [] for name, _ in METHODS}
        plans = {name: [] for name, _ in METHODS}
        for rep in range(reps):
            for name, kw in METHODS:
                sv.reset()
                st = sv.run(circ, split=split, **kw)
                times[name].append(st.device_ms)
                plans[name].append(st.plan_ms)
        for name, kw in METHODS:
            sv.reset()
            st = sv.run(circ, split=split, profile=True, **kw)     # launch counts, timing log
            cl = st.classes()
            r = {"ms": float(np.median(times[name])),
                 "launches": int(cl.get("fused", {}).get("launches", 0) + cl.get("pass", {}).get("launches", 0)),
                 "blocks": int(st.blocks), "plan_ms": float(np.median(plans[name])), "fidelity": sv.fidelity(ref)}
            if name == "learned":
                r["planner"] = "window" if st.planner else "block-first"
                r["k"] = int(st.fuse_k)
                f, ms = sv.timing_log()
                pred = np.exp(m.forward(f).astype(np.float64)) * 2.0 ** (n - 30)
                r["launch_mape"] = float(np.mean(np.abs(pred - ms) / ms))
                r["predicted_ms"] = float(st.model_ms_pred)
            res[name] = r
    return res


def evaluate(m, ns=(26, 28, 30), seed0=1000, depth=24, split=3, log=print):
    rows = []
    for n in ns:
        circs = [C.random_layers(n, depth, seed0 + n, haar_frac=0.1), C.hea(n, 6, seed0 + 50 + n)]
        if n <= 28:
            circs.append(C.qft(n))
        for circ in circs:
            res = compare(m, circ, split=split)
            row = {"circuit": circ.name, "n": n, "gates": circ.ngates, **{
                f"{k}_{f}": v[f] for k, v in res.items() for f in ("ms", "launches", "blocks")},
                "learned_predicted_ms": res["learned"]["predicted_ms"],
                "learned_launch_mape": res["learned"]["launch_mape"],
                "learned_plan_ms": res["learned"]["plan_ms"],
                "learned_planner": res["learned"]["planner"], "learned_k": res["learned"]["k"],
                "gain_vs_greedy": res["greedy_k4"]["ms"] / res["learned"]["ms"],
                "gain_vs_passes": res["passes_k4"]["ms"] / res["learned"]["ms"],
                "gain_vs_window": res["window_k4"]["ms"] / res["learned"]["ms"],
                "gain_of_relabelling": res["learned_fixed_low"]["ms"] / res["learned"]["ms"],
                "min_fidelity": min(v["fidelity"] for v in res.values())}
            log(json.dumps(row))
            rows.append(row)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=E.DEFAULT_COSTMODEL)
    ap.add_argument("--report", default=None)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--split", type=int, default=3, help="product of the held-out comparison (0, 1, 2, 3)")
    ap.add_argument("--evaluate", default=None,
                    help="skip collection and training: evaluate these trained parameters and merge the "
                         "held-out rows into --report (which must hold the training figures)")
    a = ap.parse_args(argv)

    def log(*s):
        print(*s, file=sys.stderr, flush=True)

    if a.evaluate:
        with open(a.report) as f:
            report = json.load(f)
        rows = evaluate(E.CostModel.load(a.evaluate), ns=(24, 28) if a.quick else (26, 28, 30), split=a.split,
                        log=log)
        report.update({"heldout_split": a.split, "heldout": rows,
                       "geomean_gain_vs_greedy": float(np.exp(np.mean([np.log(r["gain_vs_greedy"]) for r in rows]))),
                       "geomean_gain_vs_passes": float(np.exp(np.mean([np.log(r["gain_vs_passes"]) for r in rows]))),
                       "geomean_gain_vs_window": float(np.exp(np.mean([np.log(r["gain_vs_window"]) for r in rows]))),
                       "geomean_gain_of_relabelling": float(np.exp(np.mean([np.log(r["gain_of_relabelling"])
                                                                            for r in rows]))),
                       "min_fidelity": float(min(r["min_fidelity"] for r in rows))})
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
        print(json.dumps({k: v for k, v in report.items() if k != "heldout"}), flush=True)
        return

    t0 = time.time()
    ns = (20, 24, 28) if a.quick else (20, 22, 24, 26, 28, 30, 32)
    x, t = collect(ns, log=log)
    g = np.random.default_rng(0)
    perm = g.permutation(x.shape[0])
    cut = int(0.85 * x.shape[0])
    tr, va = perm[:cut], perm[cut:]
    m, loss0, loss = train(x[tr], t[tr], epochs=a.epochs, log=log)
    yv = m.forward(x[va])
    val_rmse = float(np.sqrt(np.mean((yv - t[va]) ** 2)))
    log(f"validation rmse (log ms): {val_rmse:.4f}  (~{100 * (math.exp(val_rmse) - 1):.1f}% relative)")
    m.save(a.out)
    rows = evaluate(m, ns=(24, 28) if a.quick else (26, 28, 30), split=a.split, log=log)
    report = {"records": int(x.shape[0]), "heldout_split": a.split, "train_loss_initial": loss0, "train_loss": loss,
              "val_rmse_log": val_rmse, "seconds": time.time() - t0, "heldout": rows,
              "geomean_gain_vs_greedy": float(np.exp(np.mean([np.log(r["gain_vs_greedy"]) for r in rows]))),
              "geomean_gain_vs_passes": float(np.exp(np.mean([np.log(r["gain_vs_passes"]) for r in rows]))),
              "geomean_gain_vs_window": float(np.exp(np.mean([np.log(r["gain_vs_window"]) for r in rows]))),
              "min_fidelity": float(min(r["min_fidelity"] for r in rows))}
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps({k: v for k, v in report.items() if k != "heldout"}), flush=True)


if __name__ == "__main__":
    main()
