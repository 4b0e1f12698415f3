#!/usr/bin/env python3
"""Benchmark: gate-applications/s and achieved HBM bandwidth at 30 qubits.

Default workload (--workload random30): a 30-qubit random circuit of depth 200
(per layer: H/RX/RY/RZ on every qubit with angles uniform in [0, 2*pi); a seeded
perfect pairing with CNOT or CZ, 10% of the pairs replaced by Haar-random
two-qubit unitaries), 9,000 logical gates, synthetic, generated from a fixed seed.

Paths:
  plain  per-gate complex128 sweeps (one launch per gate), the FP64 reference path;
  fused  complex64 state, gates fused into <= k-qubit blocks run as complex
         GEMMs on the matrix cores -- by default the FP16x3 product (matrix and
         power-of-two-scaled state split into fp16 hi + lo pieces, three piece
         products accumulated in FP32; --product bf16x6: three bf16 pieces,
         six products; --product split: FP32 MFMA plus a BF16 correction of
         the matrix's lo part; --product fp32) -- scheduled by the trained cost
         model (planner: block-first or 13-qubit windows; block size, grouping
         of blocks into shared HBM sweeps, tile shape); --scheduler greedy
         gives one sweep per block, --scheduler window one sweep per window.
A step is one full circuit execution from |0...0> on the headline path (fused).
The plain path (also once with gate tiling) and the greedy k=4 fused path with
the FP32 split product each run once for comparison; fidelity |<plain|fused>|^2
is computed on the device in FP64.

Other workloads (one JSON line each; not the default line):
  --workload large33  33-qubit random circuit of depth 60, complex64 (64 GiB),
                      learned schedule (and one run of k = 5 blocks in window
                      sweeps for comparison), FP64 norm, 10^5 measurement shots from
                      the final state (linear cross-entropy of the samples), and
                      the fidelity of the same pipeline against the FP64 path on
                      the 28-qubit restriction of the same seeded generator;
  --workload batch    4096 hardware-efficient-ansatz instances (10..22 qubits,
                      1..8 layers) in one persistent launch under hipGraph
                      replay; instances/s with per-instance p50 / p99 latency;
  --workload qrunes   the random-circuit generator at n=26, depth 20 (no Haar
                      pairs; CZ written as CR(pi)) written as a QRunes program
                      with MEASURE on every qubit, run through the QPandaSDK
                      facade (load, setComputeUnit(GPU), run(10^5)): shots/s.
                      The reference re-runs the program per shot; the facade
                      runs it once and samples the shots, and the bench checks
                      on a short run that both give the same result strings;
  --workload lowq     gates on the lowest qubits of a 30-qubit complex128 state
                      (CZ(0,1), CNOT(0,1), H(0); CZ(20,21) for comparison), 32
                      of each: ms per gate one sweep per gate, and with gate
                      tiling (the 32 gates in one sweep over LDS tiles);
  --workload stateio  whole-state host copies and binary state files of a
                      28-qubit complex128 state (4 GiB) and its complex64 form:
                      GB/s of get / set / save / load, checked bit-exact.

Roofline: the dominant kernel class of the headline path, timed per launch with
HIP events on the state's stream inside the timed steps; achieved = algorithmic
bytes (every amplitude a gate may change, read and written once) / time.  A
fused k=4 block performs 8 complex-GEMM flop per byte moved; a multi-block
sweep ("pass") 8 flop/B per block.  The matrix-core peak in those flops is
the FP32 MFMA peak (157.3 TF, ridge ~20 flop/B) for the FP32 products, the
BF16 dense peak / 6 (417 TF, ridge ~52 flop/B) for BF16x6, which executes
six bf16 piece products per flop, and the FP16 dense peak / 3 (833 TF, ridge
~104 flop/B) for FP16x3; a sweep is MFMA-bound once its intensity crosses the
ridge, else HBM-bound.  The matrix-core rate is always reported
("mfma").  "copy_gbs" is a
state-to-state copy of the same size timed in the same run: the practical
streaming rate of the device, for reference.

CPU baseline: QPanda's own CPU simulator (oracle/_ref, OpenMP build) on a bounded
strided sample of the same circuit at the same size, on this host's cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--workload", choices=["random30", "large33", "batch", "qrunes", "stateio", "lowq"],
                   default="random30")
    p.add_argument("--n", type=int, default=None, help="qubits (random30: 30, large33: 33)")
    p.add_argument("--depth", type=int, default=None, help="layers (random30: 200, large33: 60)")
    p.add_argument("--shots", type=int, default=100000, help="large33, qrunes: measurement shots")
    p.add_argument("--count", type=int, default=4096, help="batch: instances")
    p.add_argument("--haar", type=float, default=0.1)
    p.add_argument("--seed", type=int, default=2025)
    p.add_argument("--fuse-k", type=int, default=4, help="block size for greedy / passes (learned picks its own)")
    p.add_argument("--scheduler", choices=["greedy", "passes", "learned", "window"], default="learned",
                   help="greedy: one block per sweep; passes: greedy multi-block sweeps; "
                        "learned: cost-model schedule (needs a trained model)")
    p.add_argument("--group", type=int, default=8, help="blocks per sweep for --scheduler passes")
    p.add_argument("--pass-tb", type=int, default=0,
                   help="tile bits of multi-block sweeps (passes, window; 0 = minimal / 13)")
    p.add_argument("--costmodel", default=None, help="trained cost-model parameters (float32)")
    p.add_argument("--path", choices=["fused", "plain"], default="fused")
    p.add_argument("--product", choices=["fp32", "split", "bf16x6", "fp16x3"], default="fp16x3",
                   help="matrix product of the fused blocks")
    p.add_argument("--no-plain", action="store_true", help="skip the FP64 and greedy comparison runs")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--cpu-stride", type=int, default=97)
    p.add_argument("--probe", action="store_true",
                   help="counter probe: read-only, write-only and copy passes (known bytes) and three RX sweeps")
    a = p.parse_args()
    a.split = {"fp32": 0, "split": 1, "bf16x6": 2, "fp16x3": 3}[a.product]
    if a.n is None:
        a.n = {"large33": 33, "qrunes": 26, "stateio": 28}.get(a.workload, 30)
    if a.depth is None:
        a.depth = {"large33": 60, "qrunes": 20}.get(a.workload, 200)
    return a


def probe(a):
    """Short fixed launch sequence for hardware-counter runs and the device's
    streaming ceiling: the copy kernel moves exactly 2 * 16 * 2^n bytes, which
    scales FETCH_SIZE / WRITE_SIZE; a read-only pass (the FP64 norm), a
    write-only pass (|0...0>) and a copy, three of each, timed on the host
    around a synchronised launch; then RX sweeps on qubits 0, n/2 and n-1 (same
    algorithmic bytes as the copy)."""
    from qpanda_amd import circuits as C
    from qpanda_amd import engine as E
    S = 16 * (1 << a.n)
    out = {"n": a.n, "copy_bytes": 2 * S}
    with E.StateVector(a.n, E.C128) as s1, E.StateVector(a.n, E.C128) as s2:
        s2.copy_from(s1)
        rates = {}
        for name, fn, nbytes in (("read_only_norm", s1.norm2, S), ("write_only_fill", s1.reset, S),
                                 ("copy", lambda: s2.copy_from(s1), 2 * S)):
            best = None
            for _ in range(3):
                s1.sync()
                s2.sync()
                t0 = time.perf_counter()
                fn()
                s1.sync()
                s2.sync()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            rates[name] = {"ms": best * 1e3, "gbs": nbytes / best / 1e9}
        out["streaming"] = rates
        b = C._Builder(a.n)
        for q in (0, a.n // 2, a.n - 1):
            b.add(C.RX, q, theta=0.5)
        st = s1.run(b.build(), profile=True)
        out["rx_ms"] = st.class_ms[0] / max(1, st.class_launches[0])
        out["rx_algorithmic_bytes"] = 2 * S
    print(json.dumps(out), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel_class, a, headline, k):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (profiles/hbm_traffic_per_launch.txt, scaled by the copy-kernel
    calibration), when they cover this kernel at this size; else None."""
    key = None
    if kernel_class in ("fused", "pass"):
        key = f"{kernel_class}_k{k}_{a.product}@{a.n}"
    elif kernel_class == "dense1" and headline == "plain":
        key = f"dense1_c128@{a.n}"
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_per_launch.txt")) as f:
            t = json.load(f)
        return t[key]["bytes_per_launch"] if key in t else None
    except (OSError, ValueError, KeyError):
        return None


def physical_cores() -> int:
    """Physical cores among the CPUs this process may run on (distinct
    package/core pairs of its affinity set)."""
    try:
        cores = set()
        for c in os.sched_getaffinity(0):
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        return max(1, len(cores))
    except (OSError, AttributeError):
        return os.cpu_count() or 1


def cpu_threads():
    """Threads for the OpenMP baseline: the physical cores this process may use,
    capped by OMP_NUM_THREADS when the environment sets it (the share of the
    host a job is granted)."""
    phys = physical_cores()
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(env, phys) if env > 0 else phys), phys


def cpu_baseline(circ, seconds, stride):
    """QPanda's CPU simulator (reference build) on records 0, stride, 2*stride, ...
    of the benchmark circuit: the OpenMP build on the physical cores (threads
    bound close), the as-shipped 1-thread build, and a check that the OpenMP
    build's state equals the 1-thread build's (its CNOT / Toffoli loops race,
    CPUQuantumGates.cpp:678/688, :1391/1414).  Falls back to the oracle port
    when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from qpanda_amd import circuits as Cc
    threads, phys = cpu_threads()
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "bench.qsvc")
        circ.save(path)
        if O.ref_available():
            r = O.ref_time(path, stride, seconds, omp=True, threads=threads)
            r1 = O.ref_time(path, stride, max(8.0, seconds / 2), omp=False)
            # the same gate distribution at 22 qubits, first 300 records, both builds
            chk = os.path.join(td, "check.qsvc")
            Cc.random_layers(22, 20, 2222, haar_frac=0.1).save(chk)
            d_omp = O.ref_state_digest(chk, 300, omp=True, threads=threads)
            d_one = O.ref_state_digest(chk, 300, omp=False)
            return {"value": r["gates"] / r["seconds"], "unit": "gate-apps/s", "cores": threads,
                    "kind": "reference",
                    "sample": (f"QPanda CPUQuantumGates (OpenMP build, OMP_NUM_THREADS={threads}, "
                               f"OMP_PROC_BIND=close) at n={circ.n}, c128: records 0,{stride},{2 * stride},... "
                               f"of the benchmark circuit, {r['gates']} gates ({r['reference_calls']} reference "
                               f"calls; CZ as CR(pi), Haar pairs as an 18-call KAK-shaped template) in "
                               f"{r['seconds']:.1f} s, state allocation excluded"),
                    "physical_cores_visible": phys,
                    "one_thread_as_shipped": {
                        "value": r1["gates"] / r1["seconds"], "unit": "gate-apps/s", "gates": r1["gates"],
                        "seconds": r1["seconds"],
                        "build": "no OpenMP, -O3 (QPandaSDK/CMakeLists.txt:65-71)"},
                    "omp_state_equals_one_thread": d_omp == d_one,
                    "omp_check_sample": ("n=22, first 300 records of the benchmark gate distribution "
                                         f"(seed 2222), FNV-1a of the state bytes: OpenMP {d_omp}, "
                                         f"1 thread {d_one}")}
        # port: the oracle restatement, single thread
        from qpanda_amd.circuits import Circuit
        st = O.State(circ.n)
        t0 = time.time()
        done = 0
        for i in range(0, circ.ngates, stride):
            st.run(Circuit(circ.n, circ.gates[i:i + 1].copy(), circ.mats))
            done += 1
            if time.time() - t0 > seconds:
                break
        dt = time.time() - t0
        del st
        return {"value": done / dt, "unit": "gate-apps/s", "cores": 1, "kind": "port",
                "sample": f"oracle restatement, 1 thread, n={circ.n}, {done} strided gates in {dt:.1f} s"}


def tiled_plain_run(E, circ, sv_plain, plain):
    """The plain path with gate tiling on (queued per-gate sweeps applied over
    64 KiB LDS tiles): one run, its device time, and a strided sample of its
    amplitudes compared bit for bit with the plain path's."""
    log("plain path with gate tiling: one run")
    with E.StateVector(circ.n, E.C128) as sv:
        sv.set_gate_tiling(True)
        st = sv.run(circ)
        idx = np.arange(0, 1 << circ.n, max(1, (1 << circ.n) // 65536), dtype=np.uint64)
        same = bool(np.array_equal(sv.gather(idx), sv_plain.gather(idx)))
    G = circ.ngates
    return {"gate_apps_per_s": G / (st.device_ms / 1e3), "device_ms": st.device_ms, "sweeps": int(st.launches),
            "speedup_vs_plain": plain.device_ms / st.device_ms, "sample_equals_plain": same,
            "sample": f"{idx.size} amplitudes at stride {int(idx[1] - idx[0]) if idx.size > 1 else 1}"}


def copy_rate(n, reps=5):
    """State-to-state copy of a complex64 state of n qubits (read + write), GB/s."""
    from qpanda_amd import engine as E
    with E.StateVector(n, E.C64) as x, E.StateVector(n, E.C64) as y:
        y.copy_from(x)
        y.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            y.copy_from(x)
        y.sync()
        dt = (time.perf_counter() - t0) / reps
    return 2 * 8 * (1 << n) / dt / 1e9


def load_costmodel(a):
    from qpanda_amd import engine as E
    path = a.costmodel or E.DEFAULT_COSTMODEL
    if not os.path.exists(path):
        raise SystemExit(f"--scheduler learned needs a trained cost model ({path})")
    return E.CostModel.load(path)


def fused_dtype(a):
    return {"bf16x6": "c64 (fp32 state; bf16x6 matrix-core product, fp32 accumulation; fp64 reductions)",
            "fp16x3": "c64 (fp32 state; fp16x3 matrix-core product, fp32 accumulation; fp64 reductions)",
            "split": "c64 (fp32 mfma + bf16 correction; fp64 reductions)",
            "fp32": "c64 (fp32 mfma; fp64 reductions)"}[a.product]


def roofline_of(stats_list, a, headline, result):
    """Dominant kernel class over the given runs: roofline object, MFMA rate."""
    agg = {}
    for st in stats_list:
        for name, c in st.classes().items():
            d = agg.setdefault(name, {"launches": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0})
            for key in d:
                d[key] += c[key]
    dom = max(agg, key=lambda nm: agg[nm]["ms"])
    d = agg[dom]
    achieved = d["bytes"] / (d["ms"] * 1e6)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if dom in ("fused", "pass"):
        # flops: the complex GEMM's own (8 per amplitude and block qubit set);
        # BF16x6 executes them six times over on the BF16 matrix cores, so its
        # peak in these flops is the BF16 dense peak / 6
        tf = d["flops"] / (d["ms"] * 1e9)
        intensity = d["flops"] / d["bytes"]
        # (FP16x3: three fp16 piece products, the FP16 dense peak / 3)
        pieces = {"bf16x6": 6, "fp16x3": 3}.get(a.product)
        peak = BF16_PEAK_TFLOPS / pieces if pieces else FP32_PEAK_TFLOPS
        result["mfma"] = {"tflops": tf, "peak": peak, "frac": tf / peak, "flop_per_byte": intensity,
                          "product": a.product}
        if pieces:
            result["mfma"]["half_tflops_executed"] = pieces * tf
        if intensity * HBM_PEAK_GBS / 1e3 >= peak:                  # right of the ridge
            roof = {"bound": "mfma", "achieved": tf, "peak": peak, "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    k_used = int(stats_list[-1].fuse_k) or a.fuse_k
    roof["traffic"] = measured_traffic(dom, a, headline, k_used)
    roof["kernel_class"] = dom
    roof["avg_launch_ms"] = d["ms"] / d["launches"]
    roof["algorithmic_bytes_per_launch"] = d["bytes"] / d["launches"]
    total_bytes = sum(v["bytes"] for v in agg.values())
    total_ms = sum(v["ms"] for v in agg.values())
    result["hbm_gbs"] = total_bytes / (total_ms * 1e6)
    result["hbm_frac"] = result["hbm_gbs"] / HBM_PEAK_GBS
    n_runs = max(1, len(stats_list))
    result["kernel_classes"] = {k2: {"launches": int(v["launches"] / n_runs), "ms": v["ms"] / n_runs,
                                     "GBs": v["bytes"] / (v["ms"] * 1e6) if v["ms"] else 0.0}
                                for k2, v in agg.items()}
    return roof


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {a.gpus}); the multi-GPU design is "
              "described, not run", file=sys.stderr)
        sys.exit(2)
    if a.probe:
        probe(a)
        return
    if a.workload == "large33":
        return main_large(a)
    if a.workload == "batch":
        return main_batch(a)
    if a.workload == "qrunes":
        return main_qrunes(a)
    if a.workload == "stateio":
        return main_stateio(a)
    if a.workload == "lowq":
        return main_lowq(a)
    from qpanda_amd import circuits as C
    from qpanda_amd import engine as E

    circ = C.random_layers(a.n, a.depth, a.seed, haar_frac=a.haar)
    G = circ.ngates
    log(f"workload n={a.n} depth={a.depth} gates={G} counts={circ.counts()}")
    headline = a.path
    steps = a.steps if a.steps is not None else (3 if headline == "fused" else 1)
    warmup = a.warmup if a.warmup is not None else (1 if headline == "fused" else 0)
    sched = {"greedy": 0, "learned": 1, "passes": 2, "window": 3}[a.scheduler]
    cm = load_costmodel(a) if sched == 1 else None

    plain = None
    greedy = None
    result = {}
    sv_plain = None
    if headline == "plain" or not a.no_plain:
        sv_plain = E.StateVector(a.n, E.C128)
        log("plain path: one run (FP64 comparison)" if headline == "fused" else "plain path")

    def run_plain(profile):
        sv_plain.reset()
        return sv_plain.run(circ, profile=profile)

    if headline == "plain":
        for _ in range(warmup):
            run_plain(False)
        sv_plain.sync()
        t0 = time.perf_counter()
        stats = []
        for _ in range(steps):
            stats.append(run_plain(True))
        sv_plain.sync()
        wall = time.perf_counter() - t0
        head_stats = stats
    else:
        if sv_plain is not None:
            plain = run_plain(True)
            sv_plain.sync()
            result["plain_tiled"] = tiled_plain_run(E, circ, sv_plain, plain)
        sv = E.StateVector(a.n, E.C64)
        if cm is not None:
            sv.attach_costmodel(cm)
        k = a.fuse_k
        if sv_plain is not None and sched != 0:
            # the FP32 split-precision product (FP32 MFMA + BF16 correction of
            # the matrix's lo part), one greedy k = 4 block per sweep
            log("greedy k=4 fused path, FP32 split product: one run (comparison)")
            greedy = sv.run(circ, fuse_k=4, split=1, scheduler=0, profile=True)
            result["greedy_k4"] = {"gate_apps_per_s": G / (greedy.device_ms / 1e3), "device_ms": greedy.device_ms,
                                   "sweeps": int(greedy.launches), "product": "split",
                                   "fidelity_vs_fp64": sv_plain.fidelity(sv)}
        for _ in range(warmup):
            sv.reset()
            sv.run(circ, fuse_k=k, split=a.split, scheduler=sched, group=a.group, pass_tb=a.pass_tb)
        sv.sync()
        t0 = time.perf_counter()
        stats = []
        for _ in range(steps):
            sv.reset()
            stats.append(sv.run(circ, fuse_k=k, split=a.split, scheduler=sched, group=a.group,
                                pass_tb=a.pass_tb, profile=True))
        sv.sync()
        wall = time.perf_counter() - t0
        head_stats = stats
        if sv_plain is not None:
            result["fidelity_vs_fp64"] = sv_plain.fidelity(sv)
        cl = stats[-1].classes()
        result["fused"] = {"k": int(stats[-1].fuse_k), "product": a.product, "scheduler": a.scheduler,
                           "planner": "window" if stats[-1].planner else "block-first",
                           "sweeps": int(sum(cl.get(c, {}).get("launches", 0) for c in ("fused", "pass", "block64"))),
                           "plan_ms": stats[-1].plan_ms,
                           "blocks": int(stats[-1].blocks),
                           "device_ms": stats[-1].device_ms,
                           "predicted_ms": stats[-1].model_ms_pred}
        sv.close()

    ms_per_step = wall * 1e3 / steps
    value = G / (wall / steps)
    roof = roofline_of(head_stats, a, headline, result)
    if plain is not None:
        result["plain"] = {"gate_apps_per_s": G / (plain.device_ms / 1e3), "device_ms": plain.device_ms,
                           "hbm_gbs": sum(plain.class_bytes) / (sum(plain.class_ms) * 1e6)}
        result["speedup_vs_plain"] = plain.device_ms / head_stats[-1].device_ms
    if greedy is not None:
        result["speedup_vs_greedy_k4"] = greedy.device_ms / head_stats[-1].device_ms

    if headline == "fused":
        result["copy_gbs"] = copy_rate(a.n)
    cpu = None
    if not a.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(circ, a.cpu_seconds, a.cpu_stride)
    if sv_plain is not None:
        sv_plain.close()

    line = {
        "metric": "gate-apps/s at 30q (random circuit, depth 200)",
        "value": value,
        "unit": "gate-apps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": fused_dtype(a) if headline == "fused" else "c128 (fp64)",
        "data": "synthetic seeded circuit",
        "config": {"workload": f"random circuit n={a.n} depth={a.depth} haar_frac={a.haar} "
                               f"seed={a.seed} ({G} gates)",
                   "path": headline, "n_qubits": a.n, "depth": a.depth, "gates": G},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(result)
    print(json.dumps(line), flush=True)


def main_large(a):
    """33 qubits, complex64: learned schedule, FP64 norm, shots, 28-qubit fidelity."""
    from qpanda_amd import circuits as C
    from qpanda_amd import engine as E
    circ = C.random_layers(a.n, a.depth, a.seed, haar_frac=a.haar)
    G = circ.ngates
    log(f"workload n={a.n} depth={a.depth} gates={G}")
    cm = load_costmodel(a)
    steps = a.steps if a.steps is not None else 1
    warmup = a.warmup if a.warmup is not None else 0
    result = {}
    with E.StateVector(a.n, E.C64) as sv:
        sv.attach_costmodel(cm)
        # comparison leg: k = 5 blocks in window sweeps (the learned schedule
        # picks its own k; configuration 4 names k = 5)
        log("k=5 window sweeps: one run (comparison)")
        sv.reset()
        k5 = sv.run(circ, fuse_k=5, split=a.split, scheduler=3, profile=True)
        cl5 = k5.classes()
        result["k5_windows"] = {"device_ms": k5.device_ms, "gate_apps_per_s_device": G / (k5.device_ms / 1e3),
                                "blocks": int(k5.blocks),
                                "sweeps": int(sum(cl5.get(c, {}).get("launches", 0) for c in ("fused", "pass"))),
                                "ms_per_sweep": cl5.get("pass", {}).get("ms", 0.0) /
                                max(1, cl5.get("pass", {}).get("launches", 0))}
        for _ in range(warmup):
            sv.reset()
            sv.run(circ, fuse_k=5, split=a.split, scheduler=1)
        sv.sync()
        stats = []
        t0 = time.perf_counter()
        for i in range(steps):
            sv.reset()
            stats.append(sv.run(circ, fuse_k=5, split=a.split, scheduler=1, profile=True))
            log(f"step {i}: device {stats[-1].device_ms:.0f} ms")
        sv.sync()
        wall = time.perf_counter() - t0
        st = stats[-1]
        norm = sv.norm2()
        result["norm2_fp64"] = norm
        cl = st.classes()
        result["fused"] = {"k": int(st.fuse_k), "product": a.product, "scheduler": "learned",
                           "planner": "window" if st.planner else "block-first", "blocks": int(st.blocks),
                           "sweeps": int(sum(cl.get(c, {}).get("launches", 0) for c in ("fused", "pass", "block64"))),
                           "plan_ms": st.plan_ms, "device_ms": st.device_ms, "predicted_ms": st.model_ms_pred}
        # shots: all qubits in order from the final state, one draw per gate skipped per shot
        log(f"{a.shots} shots")
        t1 = time.perf_counter()
        words, _, ties = sv.sample_shots(list(range(a.n)), G, 20250101, a.shots)
        t_shots = time.perf_counter() - t1
        amps = sv.gather(words)
        p = np.abs(amps) ** 2 / norm
        result["shots"] = {"count": a.shots, "seconds": t_shots, "shots_per_s": a.shots / t_shots,
                           "fallback_shots": int(ties), "distinct": int(np.unique(words).size),
                           "linear_xeb": float((1 << a.n) * p.mean() - 1)}
    roof = roofline_of(stats, a, "fused", result)
    # fidelity of the same pipeline on the 28-qubit restriction of the generator
    nr = 28
    c28 = C.random_layers(nr, a.depth, a.seed, haar_frac=a.haar)
    log("28-qubit restriction: learned c64 vs plain c128")
    with E.StateVector(nr, E.C64) as f, E.StateVector(nr, E.C128) as pl:
        f.attach_costmodel(cm)
        f.run(c28, fuse_k=5, split=a.split, scheduler=1)
        pl.run(c28)
        result["fidelity_vs_fp64_28q"] = pl.fidelity(f)
    cpu = None
    if not a.no_cpu:
        log("cpu baseline (28-qubit restriction; the 33-qubit c128 state needs 128 GiB of host memory)")
        cpu = cpu_baseline(c28, a.cpu_seconds, a.cpu_stride)
    line = {
        "metric": f"gate-apps/s at {a.n}q (random circuit, depth {a.depth}, complex64)",
        "value": G / (wall / steps), "unit": "gate-apps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": fused_dtype(a), "data": "synthetic seeded circuit",
        "config": {"workload": f"random circuit n={a.n} depth={a.depth} haar_frac={a.haar} seed={a.seed} "
                               f"({G} gates), {a.shots} shots", "n_qubits": a.n, "depth": a.depth, "gates": G},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(result)
    print(json.dumps(line), flush=True)


def main_batch(a):
    """4096 hardware-efficient-ansatz instances in one persistent launch."""
    from qpanda_amd import engine as E
    nq, ly, ang = E.hea_batch(a.count, a.seed)
    steps = a.steps if a.steps is not None else 5
    warmup = a.warmup if a.warmup is not None else 1
    b = E.Batch(nq, ly, ang)
    best, _ = b.autotune()
    for _ in range(warmup):
        b.run()
    zz, lat, eager_ms = b.run()
    t0 = time.perf_counter()
    zg, graph_ms = b.graph_run(replays=steps)
    wall = time.perf_counter() - t0
    b.close()
    nq = np.asarray(nq)
    ly = np.asarray(ly)
    # algorithmic HBM bytes: instances above 14 qubits stream their state twice per layer,
    # less the read of |0...0> before the first pass and the store after the last
    big = (nq > 14) & (ly > 0)
    hbm_bytes = float(np.sum((4 * ly[big] - 2) * 8 * (2.0 ** nq[big])))
    gate_apps = float(np.sum(ly * (2 * nq - 1)))
    result = {
        "latency_us": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "eager_ms": eager_ms, "graph_replay_ms": graph_ms, "threads": best,
        "graph_equals_eager": bool(np.array_equal(zg, zz)),
        "gate_apps_per_s": gate_apps / (graph_ms / 1e3),
        "hbm_gbs_streamed_instances": hbm_bytes / (graph_ms * 1e6),
    }
    roof = {"bound": "hbm", "achieved": result["hbm_gbs_streamed_instances"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": result["hbm_gbs_streamed_instances"] / HBM_PEAK_GBS, "traffic": None,
            "kernel_class": "batch",
            "note": "instances of <= 14 qubits stay in LDS and move no HBM bytes; bytes count the two "
                    "LDS-tiled passes per layer of the larger ones ((4 L - 2) x 8 x 2^n per instance)"}
    cpu = None
    if not a.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline_batch(nq, ang, a.cpu_seconds)
    line = {
        "metric": f"instances/s ({a.count} HEA instances, 10-22 qubits, <Z_i Z_i+1> sums)",
        "value": a.count / (graph_ms / 1e3), "unit": "instances/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": graph_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 state, fp64 reductions)", "data": "synthetic seeded instances",
        "config": {"workload": f"{a.count} hardware-efficient-ansatz instances seed={a.seed}",
                   "instances": a.count}, "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(result)
    print(json.dumps(line), flush=True)


def main_qrunes(a):
    """A QRunes program through the QPandaSDK facade: shots/s of run(shots)."""
    from qpanda_amd import circuits as C
    from qpanda_amd import engine as E
    from qpanda_amd import sdk
    circ = C.random_layers(a.n, a.depth, a.seed, haar_frac=0.0)
    G = circ.ngates
    steps = a.steps if a.steps is not None else 3
    warmup = a.warmup if a.warmup is not None else 1
    log(f"qrunes workload n={a.n} depth={a.depth} gates={G} shots={a.shots}")
    result = {}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "program.txt")
        with open(path, "w") as f:
            f.write(C.to_qrunes(circ))

        def session(seed, resample_off=False):
            s = sdk.SDK()
            if s.load_file(path) != sdk.Q_ERROR_NONE or s.set_compute_unit(sdk.GPU) != sdk.Q_ERROR_NONE:
                raise SystemExit("QPandaSDK: loading the program failed")
            s.set_seed(seed)
            s.set_resample_disabled(resample_off)
            return s

        s = session(a.seed)
        for _ in range(warmup):
            if s.run(a.shots) != sdk.Q_ERROR_NONE:
                raise SystemExit("QPandaSDK: run failed")
        t0 = time.perf_counter()
        for _ in range(steps):
            if s.run(a.shots) != sdk.Q_ERROR_NONE:
                raise SystemExit("QPandaSDK: run failed")
        wall = time.perf_counter() - t0
        if not s.last_run_sampled():
            raise SystemExit("QPandaSDK: the program did not take the sampled path")
        result["distinct_outcomes"] = s.result().count("\n")
        s.close()
        # the reference's loop: the whole program again for every shot
        rerun_shots = 2
        s1, s2 = session(a.seed + 1), session(a.seed + 1, resample_off=True)
        s1.run(rerun_shots)
        t1 = time.perf_counter()
        s2.run(rerun_shots)
        rerun_s = (time.perf_counter() - t1) / rerun_shots
        result["rerun_equals_sampled"] = bool(s1.result() == s2.result() and
                                              np.array_equal(s1.amplitudes(), s2.amplitudes()))
        result["rerun_shots_per_s"] = 1.0 / rerun_s
        s1.close()
        s2.close()
    # the facade dispatches one FP64 kernel per gate; the same kernels, timed per
    # launch with HIP events, in a run of the same gates through qsv_run_circuit
    with E.StateVector(a.n, E.C128) as sv:
        st = sv.run(circ, profile=True)
        sv.sync()
    roof = roofline_of([st], a, "plain", result)
    roof["note"] = ("per-gate FP64 kernels of the program, timed in a run of the same gates through the "
                    "engine's FP64 path (the facade launches the same kernels, one call per gate)")
    result["gates"] = G
    result["speedup_vs_rerun"] = (steps * a.shots / wall) / result["rerun_shots_per_s"]
    cpu = None
    if not a.no_cpu:
        log("cpu baseline")
        cg = cpu_baseline(circ, a.cpu_seconds, 1)
        cpu = dict(cg)
        cpu["value"] = cg["value"] / G
        cpu["unit"] = "shots/s"
        cpu["sample"] = (cg["sample"] + f"; shots/s = gate-apps/s / {G} gates, since QPandaSDK::run "
                         "executes the whole program for every shot (its measurements not counted)")
    line = {
        "metric": f"shots/s (QRunes program through QPandaSDK: n={a.n}, depth {a.depth}, MEASURE all, "
                  f"{a.shots} shots per run)",
        "value": steps * a.shots / wall, "unit": "shots/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)", "data": "synthetic seeded circuit as QRunes text",
        "config": {"workload": f"QRunes program: random circuit n={a.n} depth={a.depth} seed={a.seed} "
                               f"({G} gates, CZ as CR(pi)) + {a.n} MEASURE, run({a.shots})",
                   "n_qubits": a.n, "depth": a.depth, "shots": a.shots},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(result)
    print(json.dumps(line), flush=True)


def main_lowq(a):
    """Per-gate sweeps of low-qubit gates against the same gates tiled."""
    from qpanda_amd import circuits as C
    from qpanda_amd import engine as E
    n, reps = a.n, 32
    N = float(1 << n)
    cases = {"CZ(0,1)": (C.CZ, 0, 1, N / 4), "CNOT(0,1)": (C.CNOT, 0, 1, N / 2), "H(0)": (C.H, 0, 0, N),
             "CZ(20,21)": (C.CZ, 20, 21, N / 4)}
    out = {}
    with E.StateVector(n, E.C128) as sv:
        for name, (kind, q0, q1, touched) in cases.items():
            b = C._Builder(n)
            for _ in range(reps):
                b.add(kind, q0, q1)
            circ = b.build()
            sv.set_gate_tiling(False)
            sv.run(circ)                                   # warm-up
            per = sv.run(circ, profile=True).device_ms / reps
            sv.set_gate_tiling(True)
            tiled = sv.run(circ).device_ms / reps
            sv.set_gate_tiling(False)
            alg = 2 * 16 * touched
            out[name] = {"per_gate_ms": per, "per_gate_alg_tbs": alg / (per * 1e9),
                         "tiled_ms_per_gate": tiled, "speedup": per / tiled}
            log(f"{name}: per-gate {per:.3f} ms ({alg / (per * 1e9):.2f} TB/s algorithmic), "
                f"tiled {tiled:.3f} ms per gate")
    line = {
        "metric": f"ms per low-qubit gate at {n}q (c128), per-gate sweep vs gate tiling",
        "value": out["CZ(0,1)"]["tiled_ms_per_gate"], "unit": "ms per gate", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": out["CZ(0,1)"]["tiled_ms_per_gate"] * reps, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic gate runs",
        "config": {"workload": f"{reps} x each gate, n={n}", "n_qubits": n},
        "roofline": None, "cpu_baseline": None, "cases": out,
    }
    print(json.dumps(line), flush=True)


def main_stateio(a):
    """Host copies and state files: GB/s of the state's FP64 bytes moved."""
    from qpanda_amd import engine as E
    steps = a.steps if a.steps is not None else 2
    n = a.n
    rng = np.random.default_rng(a.seed)
    v = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)) / np.sqrt(2 << n)
    gb = (16 << n) / 1e9
    rates = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        path = os.path.join(td, "state.bin")
        for prec, name in ((E.C128, "c128"), (E.C64, "c64")):
            want = v if prec == E.C128 else v.astype(np.complex64).astype(np.complex128)
            with E.StateVector(n, prec) as sv:
                t = {}
                for op in ("set", "get", "save", "load"):
                    best = None
                    for _ in range(steps):
                        t0 = time.perf_counter()
                        if op == "set":
                            sv.set(v)
                        elif op == "get":
                            got = sv.get()
                        elif op == "save":
                            if os.path.exists(path):
                                os.remove(path)        # every save creates the file
                            sv.save(path)
                        else:
                            sv.load(path)
                        dt = time.perf_counter() - t0
                        best = dt if best is None else min(best, dt)
                    t[op] = best
                    if op == "get" and not np.array_equal(got, want):
                        raise SystemExit(f"{name} get differs")
                if not np.array_equal(sv.get(), want):
                    raise SystemExit(f"{name} load differs")
                rates[name] = {op: gb / t[op] for op in t}
                log(f"{name}: " + " ".join(f"{op} {r:.1f} GB/s" for op, r in rates[name].items()))
    value = 2 * gb / (gb / rates["c128"]["save"] + gb / rates["c128"]["load"])
    line = {
        "metric": f"GB/s of a binary state file written and read back (n={n}, c128, {gb:.1f} GB)",
        "value": value, "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": 1e3 * (gb / rates["c128"]["save"] + gb / rates["c128"]["load"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 records",
        "data": "synthetic random state", "config": {"workload": f"state I/O n={n}", "n_qubits": n},
        "roofline": None, "cpu_baseline": None,
        "rates_gbs": rates,
        "note": "best of the timed repetitions; files in the host's temporary directory (page cache "
                "included); rates count the state's FP64 bytes (16 per amplitude) for both precisions",
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_batch(nq, ang, seconds):
    """QPanda's CPU simulator on instances 0, 1, 2, ... until `seconds` elapse."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from qpanda_amd import engine as E
    threads, _ = cpu_threads()
    if not O.ref_available():
        return None
    done, spent = 0, 0.0
    with tempfile.TemporaryDirectory() as td:
        for n, an in zip(nq, ang):
            path = os.path.join(td, "inst.qsvc")
            E.hea_circuit(int(n), an).save(path)
            r = O.ref_time(path, 1, 1e9, omp=True, threads=threads)
            spent += r["seconds"]
            done += 1
            if spent > seconds:
                break
    return {"value": done / spent, "unit": "instances/s", "cores": threads, "kind": "reference",
            "sample": f"QPanda CPUQuantumGates (OpenMP build, OMP_NUM_THREADS={threads}, OMP_PROC_BIND=close), c128, the first "
                      f"{done} instances of the batch one after another in {spent:.1f} s (allocation excluded)"}


if __name__ == "__main__":
    main()
