#!/usr/bin/env python3
"""IMM hot-path benchmark on one MI355X; prints one JSON line.

Workload (default): IMM under the LT model on a directed R-MAT graph of scale
20 (2^20 vertices, edge factor 16, a/b/c/d = .57/.19/.19/.05, deduplicated,
no self loops), in-weights drawn U(0,1] and normalised per target to 0.9,
k = 100, epsilon = 0.13.  A step is one sampling pass of `batch` RRR sets
(batch = the theta the IMM run ends with) into a reset HBM pool; the graph
is resident in HBM before the timed region starts.  IMM wall time is one full
rip_imm call (theta estimation + final sampling + selection).

Serving (the "serving" object): IMM queries on resident R-MAT graphs (scales
18/20/22; per scale one LT graph with in-weights normalised to 0.9 and one IC
graph, U[0,0.1] at scale 18 and weighted cascade above, where U[0,0.1] pools
outgrow HBM), k in {10,50,100,200}, epsilon in
{0.5,0.2,0.13}, seed = arrival counter, answered by the continuous-batching
server, against the same queries solved back to back with rip_imm: queries/s
with every query submitted at once, and p50/p99 latency under Poisson arrivals
(the back-to-back figures are those of a FIFO queue in front of the measured
solve times), once with seed = arrival counter (no stream shared) and once
with seed 0 for every query (the reference imm tool's fixed seeding, so
queries on a graph share its RRR sets).  Every served answer is checked
against its solo solve.

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


def ic_plain_bytes(index, ver):
    """Sum over samples of sum over R_s of (8 + 8 indeg + 4): every sample
    reading its own rows, the one-BFS-per-sample form."""
    v = ver.astype(np.int64)
    deg = (index[v + 1] - index[v]).astype(np.int64)
    return int(12 * len(v) + 8 * deg.sum())


def cpu_baseline(wl, index, edges, w, seed, batch, target_s):
    from oracle import oracle as O
    # the host's CPU share: OMP_NUM_THREADS where the environment sets it,
    # else at most 16 threads
    threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
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


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc runs."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic.json")
    try:
        table = json.load(open(path))
    except (OSError, ValueError):
        return None
    entry = table.get("%s/%s" % (workload, kernel))
    return None if entry is None else entry.get("hbm_bytes_per_launch")


def measure(name, seed, steps, warmup, batch, tune, skip_imm, fused_steps=0):
    wl = WORKLOADS[name]
    index, edges, w = build_graph(wl, seed)
    g = R.Graph(index, edges, w, model=wl["model"])
    imm_wall, rec = None, None
    if not skip_imm:
        seeds, rec = R.imm(g, wl["k"], wl["eps"], seed=seed, tune=tune)   # warm-up
        t = time.time()
        seeds, rec = R.imm(g, wl["k"], wl["eps"], seed=seed, tune=tune)
        imm_wall = time.time() - t
        log("%s IMM: theta=%d rounds=%d wall=%.3fs seeds[:5]=%s" % (
            name, rec["Theta"], rec["Rounds"], imm_wall, seeds[:5].tolist()))
    batch = batch or (int(rec["Theta"]) if rec else 1 << 20)
    pool = g.pool(seed)
    for _ in range(warmup):
        pool.reset()
        pool.generate_rrr_sets(batch, tune=tune)
    kms, algo = [], []
    t0 = time.time()
    for _ in range(steps):
        pool.reset()
        pool.generate_rrr_sets(batch, tune=tune)   # returns after the device finished
        kms.append(pool.last_kernel_ms()[0])
        algo.append(pool.last_algo_bytes())
    wall = time.time() - t0
    stats = pool.last_stats()
    n_sets, n_elem = pool.size()
    kernel_ms = float(np.mean(kms))
    roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel_ms": kernel_ms}
    if wl["model"] == R.IC:
        off, ver = pool.export()
        per_launch = ic_union_bytes(index, off, ver)
        plain = ic_plain_bytes(index, ver)
        roof["kernel"] = "ic_wave_kernel"
        roof["algo_bytes_per_launch"] = per_launch
        roof["algo_bytes_formula"] = ("sum over the union of the launch's sets of "
                                      "(8 + 8 indeg) + 4 per output id")
        roof["plain_form_bytes_per_launch"] = plain
        roof["plain_form_achieved"] = plain / (kernel_ms * 1e-3) / 1e9
        roof["edges_read"] = stats["edges"]
    else:
        per_launch = float(np.mean(algo))
        roof["kernel"] = "lt_walk_kernel"
        roof["algo_bytes_per_launch"] = per_launch
        roof["algo_bytes_formula"] = ("sum over walks of 8 per visited vertex + 8 per "
                                      "in-edge the subtraction loop scans + 4 per output id")
        # what the breakpoint search itself reads: the 32-byte vertex record,
        # one 64-byte node per index level or breakpoint block (the `coins`
        # counter), 4 per chosen id, 4 per output id
        search = 32 * n_elem + 64 * stats["coins"] + 8 * n_elem
        roof["search_form_bytes_per_launch"] = search
        roof["search_form_achieved"] = search / (kernel_ms * 1e-3) / 1e9
        roof["note"] = ("algorithmic bytes are those of the reference's per-edge "
                        "subtraction scan; the kernel finds each crossing edge through "
                        "per-row breakpoints instead, so frac can pass 1: traffic is "
                        "the measured HBM bytes")
    achieved = per_launch / (kernel_ms * 1e-3) / 1e9
    roof["achieved"] = achieved
    roof["frac"] = achieved / HBM_PEAK_GBS
    roof["traffic"] = measured_traffic(roof["kernel"], name)
    log("%s sampling: %d x %d sets in %.3fs; kernel %.3f ms; %.1f GB/s algorithmic; "
        "counters %s" % (name, steps, batch, wall, kernel_ms, achieved, stats))
    res = {
        "value": batch * steps / wall, "ms_per_step": wall / steps * 1e3,
        "config": {"workload": name, "description": wl["desc"],
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
        "roofline": roof,
        "graph": (wl, index, edges, w, batch),
    }
    if wl["model"] == R.IC and fused_steps:
        # the 64-colour fused sampler on the same batch, for the fused-versus-
        # one-sample-per-wave comparison (the sets are the same)
        fz = {"sampler": 2}
        pool.reset()
        pool.generate_rrr_sets(batch, tune=fz)
        t0 = time.time()
        for _ in range(fused_steps):
            pool.reset()
            pool.generate_rrr_sets(batch, tune=fz)
        fwall = time.time() - t0
        res["fused_colour_sampler"] = {
            "value": batch * fused_steps / fwall, "ms_per_step": fwall / fused_steps * 1e3,
            "kernel_ms": pool.last_kernel_ms()[0],
            "wave_over_fused": (batch * steps / wall) / (batch * fused_steps / fwall)}
        log("%s fused colour sampler: %.3f s per step (%.2fx slower than one sample per wave)"
            % (name, fwall / fused_steps, res["fused_colour_sampler"]["wave_over_fused"]))
    pool.close()
    g.close()
    return res


SERVE_K = (10, 50, 100, 200)
SERVE_EPS = (0.5, 0.2, 0.13)


# Largest scale whose IC graph keeps the U[0,0.1] in-weights in the serving
# mix: measured mean RRR set sizes of 7.4K (scale 18), 25K (20) and 83K (22)
# vertices put one k=200, eps=0.13 query's pool at 25 GB, 107 GB and 457 GB
# (beyond HBM), so the larger IC graphs use weighted-cascade weights.
SERVE_IC_UNIFORM_MAX_SCALE = 18


def serve_graphs(scales, seed):
    """Resident graphs of the serving workload: R-MAT at each scale, one IC
    graph (U[0,0.1] up to SERVE_IC_UNIFORM_MAX_SCALE, weighted cascade above)
    and one LT graph (in-weights normalised to 0.9)."""
    out = []
    for sc in scales:
        t = time.time()
        s, d = R.rmat_edges(sc)
        n = 1 << sc
        ic = (R.W_UNIFORM, 0.1, "ic") if sc <= SERVE_IC_UNIFORM_MAX_SCALE else \
            (R.W_CASCADE, 1.0, "wc")
        for model, scheme, ws, tag in ((R.IC,) + ic, (R.LT, R.W_LT_NORM, 0.9, "lt")):
            index, edges, w = R.csr_from_edges(n, s, d, scheme, ws, seed)
            out.append(("rmat%d_%s" % (sc, tag), R.Graph(index, edges, w, model=model)))
        log("serving graphs scale %d built in %.1fs" % (sc, time.time() - t))
    return out


def fifo_latencies(arrivals, service):
    """Latencies of one FIFO server that solves the queries back to back."""
    free, lat = 0.0, []
    for t, d in zip(arrivals, service):
        start = max(t, free)
        free = start + d
        lat.append(free - t)
    return lat


def served_run(graphs, queries, arrivals, slots, workers, budget):
    """Submits `queries` at `arrivals` (seconds from the start; None = all at
    once) and returns (makespan s, latencies s, server counters)."""
    srv = R.Server(slots=slots, workers=workers, budget_bytes=budget)
    gids = [srv.add_graph(g) for _, g in graphs]
    t0 = time.time()
    ids = []
    for i, (gi, k, eps, qseed) in enumerate(queries):
        if arrivals is not None:
            dt = arrivals[i] - (time.time() - t0)
            if dt > 0:
                time.sleep(dt)
        ids.append(srv.submit(gids[gi], k, eps, qseed))
    lat, results = [], []
    for q in ids:
        seeds, _, ms = srv.poll(q)
        lat.append(ms * 1e-3)
        results.append(seeds)
    span = time.time() - t0
    st = srv.stats()
    srv.close()
    return span, lat, st, results


def pct(x, p):
    return float(np.percentile(np.asarray(x), p))


def serve_traffic(graphs, queries, rng, slots, workers, load, budget):
    """One traffic mix: back-to-back plain solves, everything submitted at
    once, and Poisson arrivals at `load` x the plain capacity."""
    n_queries = len(queries)
    service, plain = [], []
    per_graph = {}
    for gi, k, eps, qseed in queries:
        t = time.time()
        seeds, _ = R.imm(graphs[gi][1], k, eps, seed=qseed)
        service.append(time.time() - t)
        plain.append(seeds)
        per_graph[graphs[gi][0]] = per_graph.get(graphs[gi][0], 0.0) + service[-1]
    seq_qps = n_queries / sum(service)
    log("serving: %d plain solves back to back in %.2fs (%.2f queries/s); seconds per graph %s"
        % (n_queries, sum(service), seq_qps, {k: round(v, 2) for k, v in per_graph.items()}))
    span, _, st, res = served_run(graphs, queries, None, slots, workers, budget)
    assert all(np.array_equal(a, b) for a, b in zip(res, plain)), "served != solo"
    sat_qps = n_queries / span
    log("serving: all %d submitted at once, done in %.2fs (%.2f queries/s); %s" % (
        n_queries, span, sat_qps, st))
    out = {"plain_queries_per_s": seq_qps, "plain_seconds_per_graph": per_graph,
           "served_saturated_queries_per_s": sat_qps, "saturated_speedup": sat_qps / seq_qps,
           "saturated_counters": st, "results_equal_solo": True, "poisson": []}
    for f in load:
        lam = f * seq_qps
        arr = np.cumsum(rng.exponential(1.0 / lam, n_queries)).tolist()
        span, lat, st, res = served_run(graphs, queries, arr, slots, workers, budget)
        assert all(np.array_equal(a, b) for a, b in zip(res, plain)), "served != solo"
        fifo = fifo_latencies(arr, service)
        row = {"rate_per_s": lam, "load_vs_plain_capacity": f,
               "served_p50_ms": pct(lat, 50) * 1e3, "served_p99_ms": pct(lat, 99) * 1e3,
               "served_queries_per_s": n_queries / span,
               "plain_fifo_p50_ms": pct(fifo, 50) * 1e3, "plain_fifo_p99_ms": pct(fifo, 99) * 1e3}
        log("serving at %.2f x plain capacity: served p50/p99 %.1f/%.1f ms, plain FIFO %.1f/%.1f ms"
            % (f, row["served_p50_ms"], row["served_p99_ms"], row["plain_fifo_p50_ms"],
               row["plain_fifo_p99_ms"]))
        out["poisson"].append(row)
    return out


def serve_bench(scales, n_queries, seed, slots, workers, load, budget, modes):
    """Continuous-batching serving against back-to-back plain solves.
    Queries: graph uniform over the resident graphs, k in {10,50,100,200},
    eps in {0.5,0.2,0.13}, Poisson arrivals.  Seed modes: "arrival" gives
    query i the sampling seed i (no two queries share a stream); "shared"
    gives every query seed 0, the fixed seeding of the reference's imm tool,
    so queries on one graph reuse each other's RRR sets."""
    graphs = serve_graphs(scales, seed)
    rng = np.random.default_rng(seed + 1)
    mix = [(int(rng.integers(len(graphs))), int(rng.choice(SERVE_K)),
            float(rng.choice(SERVE_EPS))) for _ in range(n_queries)]
    for gi in range(len(graphs)):      # first-touch costs outside the timing
        R.imm(graphs[gi][1], 10, 0.5, seed=10 ** 6)
    out = {"graphs": [name for name, _ in graphs], "k": list(SERVE_K), "eps": list(SERVE_EPS),
           "queries": n_queries, "slots": slots, "workers": workers, "budget_bytes": budget}
    for mode in modes:
        log("serving traffic with %s seeds" % mode)
        queries = [(gi, k, eps, i if mode == "arrival" else 0) for i, (gi, k, eps) in
                   enumerate(mix)]
        out[mode + "_seeds"] = serve_traffic(graphs, queries, rng, slots, workers, load, budget)
    for _, g in graphs:
        g.close()
    return out


def copy_probe(gib):
    """One device-to-device copy of `gib` GiB: the reference kernel for scaling
    rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(gib * (1 << 30))
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat20_lt", choices=sorted(WORKLOADS))
    ap.add_argument("--secondary", default="rmat22_wc,rmat16_ic",
                    help="workloads reported beside the first, comma-separated ('' for none)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0, help="sets per step (0 = theta)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="CPU-baseline budget as estimated from a small probe (the full "
                         "run usually takes a third of it)")
    ap.add_argument("--tune", default="", help='sampler tuning, e.g. "sampler=2,block=256"')
    ap.add_argument("--skip-imm", action="store_true")
    ap.add_argument("--serve-scales", default="18,20,22",
                    help="R-MAT scales of the serving workload ('' skips it)")
    ap.add_argument("--serve-queries", type=int, default=48)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0,
                    help="HBM budget of the resident pools")
    ap.add_argument("--serve-seeds", default="arrival,shared",
                    help="seed modes of the serving traffic: arrival (query i has seed i), "
                         "shared (every query seed 0)")
    ap.add_argument("--serve-load", default="0.8,2.0",
                    help="Poisson rates as multiples of the plain back-to-back capacity")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device copy of this many GiB (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if a.copy_probe:
        return copy_probe(a.copy_probe)
    if R.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    tune = None
    if a.tune:
        tune = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in a.tune.split(",")}
    res = measure(a.workload, a.seed, a.steps, a.warmup, a.batch, tune, a.skip_imm)
    wl, index, edges, w, batch = res.pop("graph")
    out = {
        "metric": "RRR sets/s + IMM wall s at fixed k,eps; p50/p99 latency of served IMM queries",
        "value": res["value"],
        "unit": "RRR sets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids, f32 weights, u64 lcg64",
        "data": "synthetic R-MAT (generated in-process, seed %d)" % a.seed,
        "config": res["config"],
        "imm_wall_s": res["imm_wall_s"],
        "imm": res["imm"],
        "pool": res["pool"],
        "kernel_counters": res["kernel_counters"],
        "tune": tune,
        "roofline": res["roofline"],
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, index, edges, w, a.seed, batch, a.cpu_seconds)
        log("cpu baseline: %.0f sets/s on %d threads" % (out["cpu_baseline"]["value"],
                                                          out["cpu_baseline"]["cores"]))
    del index, edges, w
    others = [w for w in a.secondary.split(",") if w and w != a.workload]
    for i, name in enumerate(others):
        sec = measure(name, a.seed, max(1, a.steps // 3), 1, 0, None, False,
                      fused_steps=2 if i == 0 else 0)
        sec.pop("graph")
        out["secondary" if i == 0 else "secondary_" + name] = {
            k: sec[k] for k in ("value", "ms_per_step", "config", "imm_wall_s", "imm", "pool",
                                "kernel_counters", "roofline", "fused_colour_sampler")
            if k in sec}
    if a.serve_scales:
        out["serving"] = serve_bench([int(x) for x in a.serve_scales.split(",")],
                                     a.serve_queries, a.seed, a.serve_slots, a.serve_workers,
                                     [float(x) for x in a.serve_load.split(",")],
                                     int(a.serve_budget_gib * (1 << 30)),
                                     a.serve_seeds.split(","))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
