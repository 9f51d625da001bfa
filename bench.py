#!/usr/bin/env python3
"""IMM hot-path benchmark on one MI355X; prints one JSON line.

Workloads (R-MAT a/b/c/d = .57/.19/.19/.05, edge factor 16, deduplicated, no
self loops, R-MAT seed = scale; weights from lcg64(seed).split(2,0)):
  rmat24_ic  IMM IC, scale 24, w ~ U[0,0.02], k=100, eps=0.1
  rmat22_wc  IMM IC, scale 22, weighted cascade w = 1/indeg, k=200, eps=0.1
  rmat20_lt  IMM LT, scale 20, in-weights U(0,1] normalised to 0.9, k=100,
             eps=0.13
  rmat16_ic  IMM IC, scale 16, w ~ U[0,0.1], k=50, eps=0.5
The headline line is the largest workload, rmat24_ic.  A step is one
sampling pass of `batch` RRR sets into a reset HBM pool (batch = 100,000 sets
for rmat24_ic, whose theta of about 4.4M sets is reported by its IMM run;
the theta its IMM run ends with for the others); the graph is resident in
HBM before the timed region starts.  IMM wall time is one full rip_imm call
(theta estimation + final sampling + selection), after a warm-up call on the
smaller workloads and after a warm-up sampling call of 20,000 sets on
rmat24_ic.  graph_create_s (the device layout, including the LT breakpoints,
whose cost grows with the sum of squared in-degrees, reported as
sum_indeg_squared) is reported beside it.

Roofline: the dominant kernel's algorithmic bytes per launch (the bytes its
own algorithm reads and writes, counted on the device: IC gap samplers 64 per
row expanded, 4 per candidate threshold of rows with mixed thresholds, 4 per
live in-neighbour, 4 per output id; LT breakpoint walk 32 per visited vertex,
64 per index node or breakpoint block searched, 8 per output id) over its
launch time from HIP events on the pool's stream, against 8 TB/s.  The
reference's per-edge figures (one coin per in-edge, as its samplers draw) are reported beside it,
labelled, never as the fraction.

IC workloads sample with the autotuner's choice (rip_autotune: lane list
capacity x lane workgroup size, timed on a probe of 2^20 sets on the scale-22
graph and of ~20 ms on the others), reported in `tune`; --no-autotune keeps
the defaults.

CPU baselines are the oracle's OpenMP restatements of the reference's CPU
algorithms (IC: the 64-colour BatchedBFSNeighborColor sweep,
batched_add_rrrset.h:184-257; LT: one walk per sample), timed on a bounded
sample of the same workload on this host with the thread share it is given
(OMP_NUM_THREADS, else every CPU of the process's affinity); CPU IMM wall
times (orc_imm) for the smaller workloads.

Serving (the "serving" object): IMM queries on resident R-MAT graphs
(scales 18/20/22), k in {10,50,100,200}, eps in {0.5,0.2,0.13}, answered by
the continuous-batching server against the same queries solved back to back:
queries/s with every query submitted at once and p50/p99 latency under
Poisson arrivals, with seed = arrival counter (the served workload's own
seeding) and with seed 0 for every query.  Every served answer is checked
against its solo solve.
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
    "rmat24_ic": dict(scale=24, model=R.IC, scheme=R.W_UNIFORM, wscale=0.02, k=100, eps=0.1,
                      batch=100000, imm_warmup=False,
                      desc="IMM IC, R-MAT scale 24, w ~ U[0,0.02], k=100, eps=0.1"),
    "rmat22_wc": dict(scale=22, model=R.IC, scheme=R.W_CASCADE, wscale=1.0, k=200, eps=0.1,
                      desc="IMM IC, R-MAT scale 22, weighted cascade 1/indeg, k=200, eps=0.1"),
    "rmat20_lt": dict(scale=20, model=R.LT, scheme=R.W_LT_NORM, wscale=0.9, k=100, eps=0.13,
                      desc="IMM LT, R-MAT scale 20, in-weights U(0,1] normalised to 0.9, "
                           "k=100, eps=0.13"),
    "rmat16_ic": dict(scale=16, model=R.IC, scheme=R.W_UNIFORM, wscale=0.1, k=50, eps=0.5,
                      desc="IMM IC, R-MAT scale 16, w ~ U[0,0.1], k=50, eps=0.5"),
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS") or aff or 1)
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "threads_used": threads}


def build_graph(wl, seed):
    t = time.time()
    s, d = R.rmat_edges(wl["scale"])
    n = 1 << wl["scale"]
    index, edges, w = R.csr_from_edges(n, s, d, wl["scheme"], wl["wscale"], seed)
    log("graph scale %d: n=%d m=%d built on the host in %.1fs" % (
        wl["scale"], n, len(edges), time.time() - t))
    return index, edges, w


def lt_search_bytes(stats, n_elem):
    """What the LT breakpoint walk itself moves: the 32 bytes of a vertex
    record it uses per visited vertex, one 64-byte index node or breakpoint
    block per search probe, 4 for the chosen in-neighbour and 4 per output id."""
    return 32 * n_elem + 64 * stats["coins"] + 8 * n_elem


def ic_reference_bytes(index, off, ver):
    """The reference's per-edge cost: sum over the union of the launch's sets of
    (8 + 8 indeg) + 4 per output id (every in-edge of every expanded row read
    once, as the reference's per-edge coins need)."""
    u = np.unique(ver)
    deg = (index[u.astype(np.int64) + 1] - index[u.astype(np.int64)]).astype(np.int64)
    return int(8 * len(u) + 8 * deg.sum() + 4 * len(ver))


def cpu_sample_rate(wl, index, edges, w, seed, target_s, threads):
    """Oracle restatement of the reference's CPU sampler on a bounded sample."""
    from oracle import oracle as O
    algo = O.ALGO_BATCHED64 if wl["model"] == R.IC else O.ALGO_BFS
    probe = 640 if algo == O.ALGO_BATCHED64 else 20000
    t = time.time()
    O.sample(wl["model"], index, edges, w, seed, 0, probe, algo=algo, threads=threads)
    dt = max(time.time() - t, 1e-3)
    count = int(max(probe, probe * target_s / dt))
    if algo == O.ALGO_BATCHED64:
        count = (count + 63) // 64 * 64
    t = time.time()
    O.sample(wl["model"], index, edges, w, seed, 0, count, algo=algo, threads=threads)
    dt = time.time() - t
    what = ("64-colour BatchedBFSNeighborColor sweeps (batched_add_rrrset.h:184-257), "
            "OpenMP over batches of 64" if algo == O.ALGO_BATCHED64 else
            "one walk per sample (add_rrrset.h:136-150), OpenMP over samples")
    return {"value": count / dt, "unit": "RRR sets/s", "cores": threads, "kind": "port",
            "sample": "%d RRR sets (sample ids 0..%d) of the same graph and seed, oracle %s, "
                      "%.1f s" % (count, count - 1, what, dt)}


def cpu_imm_wall(wl, index, edges, w, seed, threads):
    from oracle import oracle as O
    t = time.time()
    seeds, rec = O.imm(wl["model"], index, edges, w, wl["k"], wl["eps"], seed, threads=threads)
    return time.time() - t, seeds, rec


def spread_report(g, wl, index, edges, w, seeds, seed, gpu_replicas, cpu_replicas, threads):
    """Monte Carlo spread of the chosen seeds (diffusion_simulation.h:73-178):
    the device simulator over gpu_replicas, the oracle's restatement over the
    first cpu_replicas of the same replica streams."""
    from oracle import oracle as O
    t = time.time()
    act, _ = g.spread(seeds, gpu_replicas, seed=seed)
    gdt = time.time() - t
    t = time.time()
    eact, _ = O.spread(wl["model"], index, edges, w, seeds, seed, 0, cpu_replicas,
                       threads=threads)
    cdt = time.time() - t
    same = bool(np.array_equal(act[:cpu_replicas], eact))
    return {"replicas": gpu_replicas, "mean": float(act.mean()),
            "stderr": float(act.std() / np.sqrt(len(act))), "device_s": gdt,
            "oracle_replicas": cpu_replicas, "oracle_mean": float(eact.mean()),
            "oracle_s": cdt, "replicas_equal_oracle": same}


def sampling_steps(pool, batch, steps, warmup, tune):
    for _ in range(warmup):
        pool.reset()
        pool.generate_rrr_sets(batch, tune=tune)
    kms, calls, kstats, grids, waves = [], [], [], [], []
    t0 = time.time()
    for _ in range(steps):
        pool.reset()
        pool.generate_rrr_sets(batch, tune=tune)   # returns after the device finished
        kms.append(pool.last_kernel_ms()[0])
        calls.append(pool.last_call_ms())
        kstats.append(pool.last_kernel_stats())
        grids.append(pool.last_grid())
        waves.append(pool.last_wave())
    wall = time.time() - t0
    return wall, kms, calls, kstats, grids, waves


def measure(name, seed, steps, warmup, batch, tune, host, cpu_s, with_cpu_imm, fused_steps,
            spread_replicas, run_imm=True, autotune=True):
    wl = WORKLOADS[name]
    index, edges, w = build_graph(wl, seed)
    t = time.time()
    g = R.Graph(index, edges, w, model=wl["model"])
    graph_create_s = time.time() - t
    log("%s: rip_graph_create %.2fs (device layout%s)" % (
        name, graph_create_s, ", LT breakpoints and index" if wl["model"] == R.LT else ""))
    tune_info = {"given": tune} if tune else None
    if autotune and not tune and wl["model"] == R.IC:
        # the graph samples with the autotuner's choice from here on (IMM,
        # steps); probe: 2^20 sets on the scale-22 graph, whose steps of
        # 2.7M small sets a probe of ~20 ms (16K sets) misrepresents; ~20 ms
        # elsewhere
        t = time.time()
        best, probe_ms, tried = g.autotune(probe_sets=(1 << 20) if wl["scale"] == 22 else 0,
                                           force=True)
        tune_info = {"autotuned": best, "probe_ms": probe_ms, "tried": tried,
                     "seconds": time.time() - t}
        log("%s autotune: %s (%d settings, %.2f ms per probe, %.1fs)" % (
            name, best, tried, probe_ms, time.time() - t))
    if not run_imm:
        warm = g.pool(seed + 1)
        warm.generate_rrr_sets(20000, tune=tune)
        warm.close()
        seeds, rec, imm_wall = None, None, None
    else:
        if wl.get("imm_warmup", True):
            seeds, rec = R.imm(g, wl["k"], wl["eps"], seed=seed, tune=tune)   # warm-up
        else:
            warm = g.pool(seed + 1)
            warm.generate_rrr_sets(20000, tune=tune)
            warm.close()
        t = time.time()
        seeds, rec = R.imm(g, wl["k"], wl["eps"], seed=seed, tune=tune)
        imm_wall = time.time() - t
        log("%s IMM: theta=%d rounds=%d wall=%.3fs seeds[:5]=%s" % (
            name, rec["Theta"], rec["Rounds"], imm_wall, seeds[:5].tolist()))
    batch = batch or wl.get("batch") or int(rec["Theta"])
    pool = g.pool(seed)
    wall, kms, calls, kstats, grids, waves = sampling_steps(pool, batch, steps, warmup, tune)
    stats = pool.last_stats()
    n_sets, n_elem = pool.size()
    kernel_ms = float(np.mean(kms))
    call_ms = float(np.mean(calls))
    ks = {k: float(np.mean([s[k] for s in kstats])) for k in kstats[0]}
    roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel_ms": kernel_ms,
            "call_ms": call_ms}
    if wl["model"] == R.IC:
        per_launch = ks["algo_bytes"]
        roof["kernel"] = "ic_lane_kernel" if (tune or {}).get("sampler", 0) in (0, 3) \
            else "ic_wave_kernel"
        grid_ms = float(np.mean([gr[0] for gr in grids]))
        grid_levels = float(np.mean([gr[1] for gr in grids]))
        grid_bytes = float(np.mean([gr[2]["algo_bytes"] for gr in grids]))
        wave_ms = float(np.mean([wv[0] for wv in waves]))
        wave_bytes = float(np.mean([wv[1]["algo_bytes"] for wv in waves]))
        roof["lane_kernel_ms"] = kernel_ms
        roof["wave_stage_ms"] = wave_ms
        roof["grid_levels_ms"] = grid_ms
        roof["grid_levels"] = grid_levels
        # the dominant kernel of the call: the lane kernel, the wave stage's
        # kernel (samples past the lane capacity) or the grid sampler's level
        # kernels (timed over all levels)
        stage_ms, stage_kernel, stage_bytes = max(
            (kernel_ms, roof["kernel"], per_launch),
            (wave_ms, "ic_wave_kernel (vertex-parallel wave stage)", wave_bytes),
            (grid_ms, "grid_level_kernel (all BFS levels of the call)", grid_bytes))
        roof["kernel"] = stage_kernel
        per_launch = stage_bytes
        kernel_ms = stage_ms
        roof["kernel_ms"] = stage_ms
        roof["algo_bytes_formula"] = ("64 per row expanded (vertex record) + 4 per candidate "
                                      "edge of a row with mixed thresholds + 4 per live edge "
                                      "(in-neighbour) + 4 per output id, counted on the device")
        if n_elem <= (1 << 28):
            off, ver = pool.export()
            roof["reference_per_edge_bytes"] = ic_reference_bytes(index, off, ver)
        else:
            # a pool of billions of elements is not copied to the host: the
            # figure is taken over its first 4096 sets
            off, ver = pool.export(0, 4096)
            roof["reference_per_edge_bytes_first_4096_sets"] = ic_reference_bytes(index, off, ver)
        roof["reference_per_edge_note"] = (
            "reference per-edge cost: sum over the union of the launch's sets of (8 + 8 indeg) + 4 per "
            "output id, the bytes of one coin per in-edge; the gap sampler draws only the "
            "candidates, so this figure is not the fraction's numerator")
        del off, ver
    else:
        per_launch = lt_search_bytes(ks, n_elem)
        roof["kernel"] = "lt_walk_kernel"
        roof["algo_bytes_formula"] = ("32 per visited vertex (record) + 64 per index node or "
                                      "breakpoint block searched + 4 per chosen in-neighbour + "
                                      "4 per output id")
        roof["reference_subtraction_scan_bytes"] = ks["algo_bytes"]
        roof["reference_subtraction_scan_note"] = (
            "reference cost: 8 per visited vertex + 8 per in-edge the reference's subtraction loop "
            "scans + 4 per output id; the walk searches precomputed breakpoints instead")
    roof["algo_bytes_per_launch"] = per_launch
    achieved = per_launch / (kernel_ms * 1e-3) / 1e9
    roof["achieved"] = achieved
    roof["frac"] = achieved / HBM_PEAK_GBS
    roof["traffic"] = measured_traffic(roof["kernel"], name)
    log("%s sampling: %d x %d sets in %.3fs; main kernel %.3f ms, call %.3f ms; %.1f GB/s "
        "algorithmic; kernel counters %s" % (name, steps, batch, wall, kernel_ms, call_ms,
                                             achieved, ks))
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
            "generate_ms": rec["GenerateRRRSets"], "select_ms": rec["FindMostInfluentialSet"],
            "rrr_elements": rec["RRRElements"], "pool_device_bytes": rec["PoolBytes"],
            "theta_prime": rec["ThetaPrime"],
            "round_generate_ms": rec["ThetaEstimationGenerateRRR"]},
        "tune": tune_info,
        "graph_create_s": graph_create_s,
        "sum_indeg_squared": int(np.sum(np.diff(index.astype(np.int64)) ** 2)),
        "pool": {"sets": n_sets, "elements": n_elem, "device_bytes": pool.bytes()},
        "kernel_counters": ks, "call_counters": stats,
        "roofline": roof,
    }
    if host is not None and cpu_s > 0:
        res["cpu_baseline"] = cpu_sample_rate(wl, index, edges, w, seed, cpu_s,
                                              host["threads_used"])
        log("%s cpu baseline: %.0f sets/s on %d threads" % (
            name, res["cpu_baseline"]["value"], host["threads_used"]))
    if with_cpu_imm and host is not None and run_imm:
        cw, cseeds, crec = cpu_imm_wall(wl, index, edges, w, seed, host["threads_used"])
        res["cpu_imm"] = {"wall_s": cw, "threads": host["threads_used"], "kind": "port",
                          "seeds_equal_gpu": cseeds.tolist() == seeds.tolist(),
                          "theta": crec.theta}
        log("%s cpu IMM: %.2fs, seeds equal: %s" % (name, cw, res["cpu_imm"]["seeds_equal_gpu"]))
    if spread_replicas and host is not None and run_imm:
        res["spread"] = spread_report(g, wl, index, edges, w, seeds, seed, spread_replicas[0],
                                      spread_replicas[1], host["threads_used"])
        log("%s spread of the seeds: device %.1f (%d replicas), oracle %.1f (%d replicas)" % (
            name, res["spread"]["mean"], spread_replicas[0], res["spread"]["oracle_mean"],
            spread_replicas[1]))
    if wl["model"] == R.IC and fused_steps:
        # the same batch with one BFS per wave (the plain one-sample-per-BFS
        # form) and with 64-colour groups for every sample: same sets
        for label, kind in (("plain_wave", 1), ("colour_groups", 2)):
            tz = dict(tune or {})
            tz["sampler"] = kind
            pw, pk, _, _, _, _ = sampling_steps(pool, batch, fused_steps, 1, tz)
            res[label] = {"value": batch * fused_steps / pw, "ms_per_step": pw / fused_steps * 1e3,
                          "kernel_ms": float(np.mean(pk))}
            res[label]["batched_lanes_over_this"] = res["value"] / res[label]["value"]
            log("%s %s: %.3f s per step; lane sampler %.2fx faster" % (
                name, label, pw / fused_steps, res[label]["batched_lanes_over_this"]))
    pool.close()
    g.close()
    return res


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


SERVE_K = (10, 50, 100, 200)
SERVE_EPS = (0.5, 0.2, 0.13)

# Largest scale whose IC graph keeps the U[0,0.1] in-weights in the serving
# mix: measured mean RRR set sizes of 7.4K (scale 18), 25K (20) and 83K (22)
# vertices put one k=200, eps=0.13 query's pool at 25 GB, 107 GB and 457 GB
# as sorted lists, so the larger IC graphs use weighted-cascade weights.
SERVE_IC_UNIFORM_MAX_SCALE = 18


def serve_graphs(scales, seed, autotune=True):
    """Resident graphs of the serving workload: R-MAT at each scale, one IC
    graph (U[0,0.1] up to SERVE_IC_UNIFORM_MAX_SCALE, weighted cascade above)
    and one LT graph (in-weights normalised to 0.9).  With autotune, each IC
    graph samples with the autotuner's choice, in plain and served solves
    alike; returns the graphs and those choices."""
    out, tuned = [], {}
    for sc in scales:
        t = time.time()
        s, d = R.rmat_edges(sc)
        n = 1 << sc
        ic = (R.W_UNIFORM, 0.1, "ic") if sc <= SERVE_IC_UNIFORM_MAX_SCALE else \
            (R.W_CASCADE, 1.0, "wc")
        for model, scheme, ws, tag in ((R.IC,) + ic, (R.LT, R.W_LT_NORM, 0.9, "lt")):
            index, edges, w = R.csr_from_edges(n, s, d, scheme, ws, seed)
            name = "rmat%d_%s" % (sc, tag)
            g = R.Graph(index, edges, w, model=model)
            if autotune and model == R.IC:
                best, _, _ = g.autotune(probe_sets=(1 << 20) if sc == 22 else 0)
                tuned[name] = best
            out.append((name, g))
        log("serving graphs scale %d built in %.1fs" % (sc, time.time() - t))
    if tuned:
        log("serving graphs autotuned: %s" % tuned)
    return out, tuned


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


def serve_bench(scales, n_queries, seed, slots, workers, load, budget, modes, autotune=True):
    """Continuous-batching serving against back-to-back plain solves.
    Queries: graph uniform over the resident graphs, k in {10,50,100,200},
    eps in {0.5,0.2,0.13}, Poisson arrivals.  Seed modes: "arrival" gives
    query i the sampling seed i (the served workload's seeding); "shared"
    gives every query seed 0, the fixed seeding of the reference's imm tool,
    so queries on one graph reuse each other's RRR sets."""
    graphs, tuned = serve_graphs(scales, seed, autotune)
    rng = np.random.default_rng(seed + 1)
    mix = [(int(rng.integers(len(graphs))), int(rng.choice(SERVE_K)),
            float(rng.choice(SERVE_EPS))) for _ in range(n_queries)]
    for gi in range(len(graphs)):      # first-touch costs outside the timing
        R.imm(graphs[gi][1], 10, 0.5, seed=10 ** 6)
    out = {"graphs": [name for name, _ in graphs], "k": list(SERVE_K), "eps": list(SERVE_EPS),
           "queries": n_queries, "slots": slots, "workers": workers, "budget_bytes": budget,
           "tune": tuned or None}
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


SECONDARY_KEYS = ("value", "ms_per_step", "config", "imm_wall_s", "imm", "tune", "graph_create_s",
                  "sum_indeg_squared", "pool", "kernel_counters", "roofline", "cpu_baseline", "cpu_imm", "spread",
                  "plain_wave", "colour_groups")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat24_ic", choices=sorted(WORKLOADS))
    ap.add_argument("--secondary", default="rmat22_wc,rmat20_lt,rmat16_ic",
                    help="workloads reported beside the first, comma-separated ('' for none)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0, help="sets per step (0 = theta)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU-baseline budget per workload as estimated from a small probe")
    ap.add_argument("--tune", default="", help='sampler tuning, e.g. "sampler=1,queue_cap=256"')
    ap.add_argument("--fused-steps", type=int, default=2,
                    help="steps of the plain and colour-group samplers on the rmat22_wc batch")
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
    ap.add_argument("--no-imm", action="store_true",
                    help="skip the IMM runs (profiling runs of the sampling steps; needs --batch "
                         "for workloads whose step is theta)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="sample with the default settings instead of the autotuner's choice")
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
    host = host_info()
    cpu_s = 0.0 if a.no_cpu else a.cpu_seconds
    big = WORKLOADS[a.workload]["scale"] >= 24
    res = measure(a.workload, a.seed, a.steps, a.warmup, a.batch, tune, host, cpu_s,
                  WORKLOADS[a.workload]["scale"] <= 20,
                  a.fused_steps if a.workload == "rmat22_wc" else 0,
                  (256, 32) if big else (4096, 64), not a.no_imm, not a.no_autotune)
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
        "dtype": "u32 vertex ids, integer coins (u64 lcg64 draws, fixed-point gaps)",
        "data": "synthetic R-MAT (generated in-process, seed %d)" % a.seed,
        "config": res["config"],
        "tune": res["tune"],
        "host": host,
    }
    for k in ("imm_wall_s", "imm", "graph_create_s", "pool", "kernel_counters", "call_counters",
              "roofline", "cpu_baseline", "cpu_imm", "spread", "plain_wave", "colour_groups"):
        if k in res:
            out[k] = res[k]
    out.setdefault("cpu_baseline", None)
    others = [x for x in a.secondary.split(",") if x and x != a.workload]
    for name in others:
        sec = measure(name, a.seed, max(1, a.steps // 2), 1, 0, None, host, cpu_s,
                      WORKLOADS[name]["scale"] <= 20,
                      a.fused_steps if name == "rmat22_wc" else 0, None, not a.no_imm,
                      not a.no_autotune)
        out["secondary_" + name] = {k: sec[k] for k in SECONDARY_KEYS if k in sec}
    if a.serve_scales:
        out["serving"] = serve_bench([int(x) for x in a.serve_scales.split(",")],
                                     a.serve_queries, a.seed, a.serve_slots, a.serve_workers,
                                     [float(x) for x in a.serve_load.split(",")],
                                     int(a.serve_budget_gib * (1 << 30)),
                                     a.serve_seeds.split(","), not a.no_autotune)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
