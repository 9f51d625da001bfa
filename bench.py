#!/usr/bin/env python3
"""IMM hot-path benchmark on one MI355X; prints one JSON line.

Workload (default): IMM under the LT model on a directed R-MAT graph of scale
20 (2^20 vertices, edge factor 16, a/b/c/d = .57/.19/.19/.05, deduplicated,
no self loops), in-weights drawn U(0,1] and normalised per target to 0.9,
k = 100, epsilon = 0.13.  A step is one sampling pass of `batch` RRR sets
(batch = the theta the IMM run ends with) into a reset HBM pool; the graph
is resident in HBM before the timed region starts.  IMM wall time is one full
rip_imm call (theta estimation + final sampling + selection).

The CPU baseline is the oracle's OpenMP restatement of the reference's CPU
path, timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import ripples_amd as R

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "rmat20_lt": dict(scale=20, model=R.LT, scheme=R.W_LT_NORM, wscale=0.9, k=100,
                      eps=0.13, desc="IMM LT, R-MAT scale 20, in-weights U(0,1] "
                                     "normalised to 0.9, k=100, eps=0.13"),
    "rmat22_wc": dict(scale=22, model=R.IC, scheme=R.W_CASCADE, wscale=1.0, k=200,
                      eps=0.1, desc="IMM IC, R-MAT scale 22, weighted cascade "
                                    "1/indeg, k=200, eps=0.1"),
    "rmat16_ic": dict(scale=16, model=R.IC, scheme=R.W_UNIFORM, wscale=0.1, k=50,
                      eps=0.5, desc="IMM IC, R-MAT scale 16, w ~ U[0,0.1], k=50, "
                                    "eps=0.5"),
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_graph(wl, seed):
    t = time.time()
    s, d = R.rmat_edges(wl["scale"])
    n = 1 << wl["scale"]
    index, edges, w = R.csr_from_edges(n, s, d, wl["scheme"], wl["wscale"], seed)
    log("graph scale %d: n=%d m=%d built in %.1fs" % (wl["scale"], n, len(edges),
                                                         time.time() - t))
    return index, edges, w


def ic_union_bytes(index, off, ver):
    """B_IC(L) = sum over the union of (8 + 8 indeg) + 4 per output id."""
    u = np.unique(ver)
    deg = (index[u.astype(np.int64) + 1] - index[u.astype(np.int64)]).astype(np.int64)
    return int(8 * len(u) + 8 * deg.sum() + 4 * len(ver))


def cpu_baseline(wl, index, edges, w, seed, batch, target_s):
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    threads = min(threads, 64)
    probe = max(1000, min(batch, 20000))
    t = time.time()
    O.sample(wl["model"], index, edges, w, seed, 0, probe, threads=threads)
    dt = max(time.time() - t, 1e-3)
    count = int(max(probe, probe * target_s / dt))
    t = time.time()
    O.sample(wl["model"], index, edges, w, seed, 0, count, threads=threads)
    dt = time.time() - t
    return {"value": count / dt, "unit": "RRR sets/s", "cores": threads,
            "kind": "port",
            "sample": "%d RRR sets (sample ids 0..%d) of the same graph and seed, "
                      "oracle FIFO reverse BFS/walk per sample, OpenMP over "
                      "samples, %.1f s" % (count, count - 1, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat20_lt", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0, help="sets per step (0 = theta)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--tune", default="", help='sampler tuning, e.g. "colors=1,block=256"')
    ap.add_argument("--skip-imm", action="store_true")
    a = ap.parse_args()
    tune = None
    if a.tune:
        tune = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in a.tune.split(",")}
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if R.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    wl = WORKLOADS[a.workload]
    index, edges, w = build_graph(wl, a.seed)
    g = R.Graph(index, edges, w, model=wl["model"])

    # IMM wall time: one warm-up run, one timed run
    imm_wall, rec = None, None
    if not a.skip_imm:
        seeds, rec = R.imm(g, wl["k"], wl["eps"], seed=a.seed, tune=tune)
        t = time.time()
        seeds, rec = R.imm(g, wl["k"], wl["eps"], seed=a.seed, tune=tune)
        imm_wall = time.time() - t
        log("IMM: theta=%d rounds=%d wall=%.3fs seeds[:5]=%s" % (
            rec["Theta"], rec["Rounds"], imm_wall, seeds[:5].tolist()))

    batch = a.batch or (int(rec["Theta"]) if rec else 1 << 20)
    pool = g.pool(a.seed)
    for _ in range(a.warmup):
        pool.reset()
        pool.generate_rrr_sets(batch, tune=tune)
    kms, algo = [], []
    t0 = time.time()
    for _ in range(a.steps):
        pool.reset()
        pool.generate_rrr_sets(batch, tune=tune)   # returns after the device finished
        kms.append(pool.last_kernel_ms()[0])
        algo.append(pool.last_algo_bytes())
    wall = time.time() - t0
    stats = pool.last_stats()
    log("kernel counters (last step): %s" % stats)
    value = batch * a.steps / wall
    n_sets, n_elem = pool.size()
    if wl["model"] == R.IC:
        off, ver = pool.export()
        per_launch = ic_union_bytes(index, off, ver)
    else:
        per_launch = float(np.mean(algo))
    kernel_ms = float(np.mean(kms))
    achieved = per_launch / (kernel_ms * 1e-3) / 1e9
    log("sampling: %d x %d sets in %.3fs; kernel %.3f ms; %.1f GB/s algorithmic" % (
        a.steps, batch, wall, kernel_ms, achieved))

    out = {
        "metric": "RRR sets/s + IMM wall s at fixed k,eps; p50/p99 latency of served IMM queries",
        "value": value,
        "unit": "RRR sets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids, f32 weights, u64 lcg64",
        "data": "synthetic R-MAT (generated in-process, seed %d)" % a.seed,
        "config": {"workload": a.workload, "description": wl["desc"],
                   "num_nodes": int(len(index) - 1), "num_edges": int(len(edges)),
                   "batch_sets": batch, "k": wl["k"], "epsilon": wl["eps"],
                   "parallelism": "single GPU"},
        "imm_wall_s": imm_wall,
        "imm": None if rec is None else {
            "theta": rec["Theta"], "rounds": rec["Rounds"], "num_seeds": rec["NumSeeds"],
            "final_f": rec["FinalF"], "estimation_ms": rec["ThetaEstimation"],
            "generate_ms": rec["GenerateRRRSets"],
            "select_ms": rec["FindMostInfluentialSet"]},
        "pool": {"sets": n_sets, "elements": n_elem},
        "kernel_counters": stats,
        "tune": tune,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "lt_walk_kernel" if wl["model"] == R.LT else "ic_fused_kernel",
                     "kernel_ms": kernel_ms, "algo_bytes_per_launch": per_launch},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, index, edges, w, a.seed, batch, a.cpu_seconds)
        log("cpu baseline: %.0f sets/s on %d threads" % (out["cpu_baseline"]["value"],
                                                          out["cpu_baseline"]["cores"]))
    pool.close()
    g.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
