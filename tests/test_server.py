"""Serving loop (rip_server_*): IMM queries against resident graphs with
continuous batching.  A served query must return what imm() returns alone on
the same graph with seed = the graph's pool seed (imm.h:768-860): the
queries only select on the pool prefixes a solo run draws."""
import numpy as np
import pytest

import ripples_amd as R
from graphs import lt_normalise, random_graph, transposed_csr


def test_argument_errors():
    with pytest.raises(R.RipplesError, match="slot"):
        R.Server(slots=0)
    with pytest.raises(R.RipplesError, match="workers"):
        R.Server(workers=0)
    s = R.Server(slots=2, workers=2)
    with pytest.raises(R.RipplesError, match="no such graph"):
        s.submit(0, 3, 0.5)
    with pytest.raises(R.RipplesError, match="no such query"):
        s.poll(12345, wait_ms=0)
    assert s.stats() == dict.fromkeys(R.SERVER_STATS, 0)
    s.close()
    s.close()


def rmat_csr(scale, scheme, scale_w, seed=0):
    s, d = R.rmat_edges(scale, seed=seed)
    return R.csr_from_edges(1 << scale, s, d, scheme, scale_w, seed)


SOLO_KEYS = ["Theta", "NumSets", "Rounds", "NumSeeds", "LB", "FinalF", "ThetaPrime", "F"]


def lt_random():
    n, s, d, w = random_graph(400, 4000, 5)
    return transposed_csr(n, s, d, lt_normalise(n, s, d, w))


@pytest.mark.gpu
def test_mixed_queries_equal_solo_runs():
    graphs = [(R.IC, rmat_csr(10, R.W_UNIFORM, 0.1)), (R.LT, rmat_csr(10, R.W_LT_NORM, 0.9, 3)),
              (R.LT, lt_random())]
    devs = [R.Graph(*csr, model=m) for m, csr in graphs]
    srv = R.Server(slots=4, workers=3)
    gids = [srv.add_graph(g) for g in devs]
    # (graph, k, epsilon, seed): repeated keys exercise pool reuse
    plan = [(0, 5, 0.5, 7), (1, 8, 0.3, 11), (0, 20, 0.2, 7), (2, 4, 0.5, 0), (1, 2, 0.9, 11),
            (0, 5, 0.5, 8), (2, 10, 0.25, 0), (1, 8, 0.3, 11), (0, 1, 0.7, 7), (2, 3, 0.4, 1)]
    qids = [srv.submit(gids[gi], k, eps, seed) for gi, k, eps, seed in plan]
    results = [srv.poll(q, wait_ms=-1) for q in qids]
    need = {}
    for (gi, k, eps, seed), (seeds, rec, lat) in zip(plan, results):
        solo_seeds, solo = R.imm(devs[gi], k, eps, seed=seed)
        assert seeds.tolist() == solo_seeds.tolist(), (gi, k, eps, seed)
        for key in SOLO_KEYS:
            assert rec[key] == solo[key], (gi, k, eps, seed, key)
        assert lat > 0 and rec["Total"] > 0
        need[gi, seed] = max(need.get((gi, seed), 0), rec["NumSets"])
    st = srv.stats()
    # each pool only ever grows to the largest prefix any of its queries used
    assert st["sets_drawn"] == sum(need.values())
    assert st["pools"] == len(need) and st["reused"] == len(plan) - len(need)
    assert st["selections"] == sum(r[1]["Rounds"] + 1 for r in results)
    assert st["launches"] < st["selections"] and st["pool_bytes"] > 0
    assert st["evicted"] == 0 and st["recycled"] == 0
    with pytest.raises(R.RipplesError, match="no such query"):
        srv.poll(qids[0], wait_ms=0)
    srv.close()
    for g in devs:
        g.close()


@pytest.mark.gpu
def test_dense_pools_served_equal_solo_runs():
    """R-MAT scale 16 with U[0,0.1] weights: a fifth of the sets reach the
    giant component, so the pools take the dense selector (no inverted index)
    and the colour-group fallback; served answers still equal solo runs."""
    g = R.Graph(*rmat_csr(16, R.W_UNIFORM, 0.1, 3), model=R.IC)
    srv = R.Server(slots=2, workers=2)
    gid = srv.add_graph(g)
    plan = [(50, 0.5, 3), (10, 0.5, 3), (20, 0.7, 4)]
    qids = [srv.submit(gid, k, eps, seed) for k, eps, seed in plan]
    for (k, eps, seed), q in zip(plan, qids):
        seeds, rec, _ = srv.poll(q, wait_ms=-1)
        solo_seeds, solo = R.imm(g, k, eps, seed=seed)
        assert seeds.tolist() == solo_seeds.tolist(), (k, eps, seed)
        for key in SOLO_KEYS:
            assert rec[key] == solo[key], (k, eps, seed, key)
    srv.close()
    g.close()


@pytest.mark.gpu
def test_budget_and_single_slot_give_same_results():
    """Admission limits and eviction change the schedule, never the answers."""
    g = R.Graph(*rmat_csr(10, R.W_UNIFORM, 0.1), model=R.IC)
    plan = [(4, 0.5, 1), (12, 0.3, 2), (4, 0.5, 1), (2, 0.6, 3), (6, 0.4, 2)]
    out = []
    for slots, workers, budget in [(1, 1, 0), (4, 2, 1)]:
        srv = R.Server(slots=slots, workers=workers, budget_bytes=budget)
        gid = srv.add_graph(g)
        qids = [srv.submit(gid, k, eps, seed) for k, eps, seed in plan]
        out.append([srv.poll(q)[0].tolist() for q in qids])
        srv.close()
    # one query at a time under a 1-byte budget: every new key takes over the
    # idle pool's buffers
    srv = R.Server(slots=4, workers=2, budget_bytes=1)
    gid = srv.add_graph(g)
    out.append([srv.poll(srv.submit(gid, k, eps, seed))[0].tolist() for k, eps, seed in plan])
    st = srv.stats()
    assert st["recycled"] == 4 and st["evicted"] == 0 and st["pools"] == 1
    assert st["reused"] == 0
    srv.close()
    assert out[0] == out[1] == out[2]
    assert out[0] == [R.imm(g, k, eps, seed=seed)[0].tolist() for k, eps, seed in plan]
    g.close()


@pytest.mark.gpu
def test_query_errors_are_reported_per_query():
    g = R.Graph(*rmat_csr(8, R.W_UNIFORM, 0.1), model=R.IC)
    srv = R.Server()
    gid = srv.add_graph(g)
    with pytest.raises(R.RipplesError, match="epsilon"):
        srv.submit(gid, 3, 0.0)
    q = srv.submit(gid, 3, 0.5)
    assert srv.poll(q, wait_ms=-1)[0].tolist() == R.imm(g, 3, 0.5)[0].tolist()
    assert srv.poll(srv.submit(gid, 3, 0.5), wait_ms=-1, cap=3) is not None
    q = srv.submit(gid, 3, 0.5)
    with pytest.raises(R.RipplesError, match="too small"):
        srv.poll(q, wait_ms=-1, cap=1)
    assert srv.poll(q, wait_ms=0, cap=3) is not None  # still there after the refusal
    srv.close()
    g.close()


@pytest.mark.gpu
def test_many_seeds_recycle_pools():
    """More distinct seeds than slots on one graph: pools are re-keyed, at most
    `slots` stay resident, and every answer is still the solo one."""
    g = R.Graph(*rmat_csr(10, R.W_LT_NORM, 0.9, 1), model=R.LT)
    srv = R.Server(slots=3, workers=3)
    gid = srv.add_graph(g)
    plan = [(k, eps, seed) for seed in range(12) for k, eps in [(4, 0.5)] ] + [(7, 0.3, 2)]
    qids = [srv.submit(gid, k, eps, seed) for k, eps, seed in plan]
    got = [srv.poll(q)[0].tolist() for q in qids]
    st = srv.stats()
    assert st["pools"] <= 3 and st["recycled"] >= 9
    assert got == [R.imm(g, k, eps, seed=seed)[0].tolist() for k, eps, seed in plan]
    srv.close()
    g.close()
