"""Train the fused-block cost model from timings the engine records itself.

    python -m qpanda_amd.train_costmodel [--out qpanda_amd/costmodel_params.f32]
                                         [--report report.json] [--quick]

1. Collect: seeded circuits (random layers, QFT, hardware-efficient ansatz) at
   several qubit counts run on the fused path at every block size (k = 2..5),
   both split settings and every tile shape, with per-launch event timing; the
   engine's timing log gives one (features, ms) record per block launch.
2. Train: full-batch Adam on the device (the engine's own forward / backward
   kernels), target log(ms * 2^(30 - n)).
3. Evaluate on held-out circuits: per-block prediction error, and the measured
   time of the learned schedule against the greedy k = 4 schedule (the method
   without the model) on the same circuits; the circuit results of the two
   schedules agree to FP32 rounding.
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


def workloads(ns, seed0, depth):
    out = []
    for i, n in enumerate(ns):
        out.append(C.random_layers(n, depth, seed0 + i, haar_frac=0.1))
        out.append(C.hea(n, 3, seed0 + 100 + i))
        if n <= 26:
            out.append(C.qft(n))
    return out


def collect(ns, seed0=7, depth=6, log=print):
    """Run every workload at every (k, split, tile); return features, log-targets."""
    feats, targ = [], []
    by_n = {}
    for circ in workloads(ns, seed0, depth):
        by_n.setdefault(circ.n, []).append(circ)
    for n, circs in sorted(by_n.items()):
        with E.StateVector(n, E.C64) as sv:
            for circ in circs:
                for k in (2, 3, 4, 5):
                    for split in (0, 1):
                        for tile in (6, 7, 8):
                            sv.reset()
                            sv.run(circ, fuse_k=k, split=bool(split), tile=tile)        # warm
                            sv.reset()
                            sv.run(circ, fuse_k=k, split=bool(split), tile=tile, profile=True)
                            f, ms = sv.timing_log()
                            keep = ms > 0
                            feats.append(f[keep])
                            targ.append(np.log(ms[keep].astype(np.float64) * 2.0 ** (30 - n)))
            log(f"collected n={n}: {sum(x.shape[0] for x in feats)} records so far")
    return np.concatenate(feats).astype(np.float32), np.concatenate(targ).astype(np.float32)


def train(x, t, epochs=3000, lr=3e-3, seed=1, log=print):
    m = E.CostModel(seed)
    p = m.get_params()
    p[-1] = float(np.mean(t))                    # output bias at the mean target
    m.set_params(p)
    loss0 = m.train(x, t, 0)
    loss = m.train(x, t, epochs, lr)
    log(f"trained on {x.shape[0]} records: loss {loss0:.4f} -> {loss:.5f}")
    return m, loss0, loss


def evaluate(m, ns=(26, 28, 30), seed0=1000, depth=24, log=print):
    """Held-out circuits: per-block error and learned vs greedy measured time."""
    rows = []
    for n in ns:
        circs = [C.random_layers(n, depth, seed0 + n, haar_frac=0.1), C.hea(n, 6, seed0 + 50 + n)]
        if n <= 28:
            circs.append(C.qft(n))
        with E.StateVector(n, E.C64) as sv, E.StateVector(n, E.C64) as sv2:
            sv.attach_costmodel(m)
            for circ in circs:
                res = {}
                for name, kw in (("greedy_k4", dict(fuse_k=4, scheduler=0)),
                                 ("learned", dict(fuse_k=4, scheduler=1))):
                    target = sv if name == "learned" else sv2
                    times = []
                    for rep in range(3):
                        target.reset()
                        st = target.run(circ, split=True, profile=(rep == 2), **kw)
                        times.append(st.device_ms)
                    res[name] = {"ms": float(np.median(times)), "blocks": int(st.blocks),
                                 "predicted_ms": float(st.model_ms_pred)}
                    if name == "learned":
                        f, ms = target.timing_log()
                        pred = np.exp(m.forward(f).astype(np.float64)) * 2.0 ** (n - 30)
                        res[name]["block_mape"] = float(np.mean(np.abs(pred - ms) / ms))
                fid = sv.fidelity(sv2)
                row = {"circuit": circ.name or f"n{n}", "n": n, "gates": circ.ngates,
                       "greedy_ms": res["greedy_k4"]["ms"], "learned_ms": res["learned"]["ms"],
                       "learned_blocks": res["learned"]["blocks"], "greedy_blocks": res["greedy_k4"]["blocks"],
                       "learned_predicted_ms": res["learned"]["predicted_ms"],
                       "block_mape": res["learned"]["block_mape"],
                       "gain": res["greedy_k4"]["ms"] / res["learned"]["ms"],
                       "fidelity_learned_vs_greedy": fid}
                log(json.dumps(row))
                rows.append(row)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=E.DEFAULT_COSTMODEL)
    ap.add_argument("--report", default=None)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=4000)
    a = ap.parse_args(argv)

    def log(*s):
        print(*s, file=sys.stderr, flush=True)

    t0 = time.time()
    ns = (20, 24, 28) if a.quick else (20, 22, 24, 26, 28, 30)
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
    rows = evaluate(m, ns=(24, 28) if a.quick else (26, 28, 30), log=log)
    report = {"records": int(x.shape[0]), "train_loss_initial": loss0, "train_loss": loss,
              "val_rmse_log": val_rmse, "seconds": time.time() - t0, "heldout": rows,
              "mean_gain": float(np.mean([r["gain"] for r in rows])),
              "min_fidelity_learned_vs_greedy": float(min(r["fidelity_learned_vs_greedy"] for r in rows))}
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps({k: v for k, v in report.items() if k != "heldout"}), flush=True)


if __name__ == "__main__":
    main()
