"""GPU parity at the benchmark workloads' full sizes, through the C ABI.

Each test builds the workload's R-MAT graph (a/b/c/d = .57/.19/.19/.05, edge
factor 16, deduplicated, R-MAT seed = scale) and compares the HIP path with
the CPU oracle on the same seed:
  - LT, scale 20, in-weights U(0,1] normalised to 0.9, k=100, eps=0.13: the
    whole final pool (theta sets) and the full IMM run (seeds, theta, the
    theta' and coverage of every estimation round, LB, final coverage)
    against orc_imm (imm.h:145-260, 768-860).
  - IC weighted cascade, scale 22, k=200, eps=0.1: 30,000 sample ids in ten
    ranges spread over the final pool, and the k=200 selection on the
    exported pool against orc_select (find_most_influential.h:85-174).
  - Served queries (one IC and one LT graph per scale 18/20/22, seed = the
    arrival counter) submitted together to the continuous-batching server,
    each against orc_imm.
  - IC, scale 24, w ~ U[0,0.02]: 120,000 sets drawn into a pool that keeps
    large colour groups in its column store, 16 sample ids at each of eight
    offsets against orc_sample, and the column-store coverage of a seed set
    against the coverage counted on the exported sets.
The oracle runs on every host thread the environment allows."""
import os

import numpy as np
import pytest

import ripples_amd as R

pytestmark = pytest.mark.gpu


def threads():
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def rmat_csr(scale, scheme, scale_w, seed=0):
    s, d = R.rmat_edges(scale)
    return R.csr_from_edges(1 << scale, s, d, scheme, scale_w, seed)


def assert_same_sets(got, exp):
    assert np.array_equal(got[0], exp[0]), "set sizes differ"
    assert np.array_equal(got[1], exp[1]), "set contents differ"


def assert_imm_equal(seeds, rec, es, erec):
    assert seeds.tolist() == es.tolist()
    r = erec.rounds
    assert rec["Theta"] == erec.theta and rec["Rounds"] == r
    assert rec["ThetaPrime"] == list(erec.theta_prime[:r])
    assert rec["F"] == list(erec.f[:r])
    assert rec["LB"] == erec.lb and rec["FinalF"] == erec.final_f
    assert rec["NumSets"] == erec.num_sets


def test_lt_scale20_full_theta_and_imm(orc):
    index, edges, w = rmat_csr(20, R.W_LT_NORM, 0.9)
    g = R.Graph(index, edges, w, model=R.LT)
    seeds, rec = R.imm(g, 100, 0.13, seed=0)
    es, erec = orc.imm(R.LT, index, edges, w, 100, 0.13, 0, threads=threads())
    assert_imm_equal(seeds, rec, es, erec)
    theta = rec["Theta"]
    assert theta > 500000
    p = g.pool(0)
    p.generate_rrr_sets(theta)
    exp = orc.sample(R.LT, index, edges, w, 0, 0, theta, threads=threads())
    assert_same_sets(p.export(), exp)
    assert p.size() == (theta, int(exp[0][-1]))
    p.close()
    g.close()


def test_ic_cascade_scale22_spread_ids_and_selection(orc):
    index, edges, w = rmat_csr(22, R.W_CASCADE, 1.0)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.IC)
    seeds, rec = R.imm(g, 200, 0.1, seed=0)
    theta = rec["Theta"]
    assert theta > 2000000
    p = g.pool(0)
    p.generate_rrr_sets(theta)
    for first in np.linspace(0, theta - 3000, 10).astype(np.int64).tolist():
        exp = orc.sample(R.IC, index, edges, w, 0, first, 3000, threads=threads())
        assert_same_sets(p.export(first, 3000), exp)
    off, ver = p.export()
    es, ec = orc.select(n, off, ver, 200)
    got, cov = p.find_most_influential_set(200)
    assert got.tolist() == es.tolist() and cov == ec
    # rip_imm's final selection is the same greedy on the same theta sets
    assert seeds.tolist() == es.tolist() and rec["FinalF"] == ec / theta
    # the setting the autotuner picks on this graph (lanes up to 64 vertices,
    # the wave stage with its LDS table, 32 colours per grid group) draws
    # the same theta sets
    q = g.pool(0)
    q.generate_rrr_sets(theta, tune={"queue_cap": 64, "wave_stage": 1, "colors": 32})
    off2, ver2 = q.export()
    assert np.array_equal(off2, off) and np.array_equal(ver2, ver)
    assert q.last_wave()[0] > 0
    q.close()
    p.close()
    g.close()


SERVED = [  # (scale, model, k, epsilon): one IC and one LT query per scale
    (18, R.IC, 50, 0.5), (18, R.LT, 100, 0.2),
    (20, R.IC, 200, 0.5), (20, R.LT, 10, 0.13),
    (22, R.IC, 10, 0.5), (22, R.LT, 50, 0.2),
]


def served_csr(scale, model):
    """The serving workload's graphs: LT in-weights normalised to 0.9; IC
    U[0,0.1] at scale 18, weighted cascade above."""
    if model == R.LT:
        return rmat_csr(scale, R.W_LT_NORM, 0.9)
    if scale <= 18:
        return rmat_csr(scale, R.W_UNIFORM, 0.1)
    return rmat_csr(scale, R.W_CASCADE, 1.0)


def test_served_queries_match_oracle_imm(orc):
    csrs = [served_csr(sc, m) for sc, m, _, _ in SERVED]
    graphs = [R.Graph(*c, model=m) for c, (_, m, _, _) in zip(csrs, SERVED)]
    srv = R.Server(slots=8, workers=4)
    gids = [srv.add_graph(g) for g in graphs]
    # seed = arrival counter, every query submitted before the first finishes
    qids = [srv.submit(gids[i], k, eps, i) for i, (_, _, k, eps) in enumerate(SERVED)]
    results = [srv.poll(q, wait_ms=-1) for q in qids]
    srv.close()
    for i, ((sc, model, k, eps), (seeds, rec, _)) in enumerate(zip(SERVED, results)):
        index, edges, w = csrs[i]
        es, erec = orc.imm(model, index, edges, w, k, eps, i, threads=threads())
        assert_imm_equal(seeds, rec, es, erec)
    for g in graphs:
        g.close()


def test_ic_uniform_scale24_spread_ids(orc):
    index, edges, w = rmat_csr(24, R.W_UNIFORM, 0.02)
    g = R.Graph(index, edges, w, model=R.IC)
    count = 120000
    p = g.pool(0)
    p.generate_rrr_sets(count)
    groups, hot_n, _ = p.columns()
    assert groups > 0, "scale 24 U[0,0.02] groups keep their hot members as columns"
    # the column store holds the pool in fewer bytes than its lists would take
    n_sets, n_elem = p.size()
    assert n_sets == count and 0 < p.bytes() < 4 * n_elem
    for first in np.linspace(0, count - 16, 8).astype(np.int64).tolist():
        off, ver = p.export(first, 16)
        exp = orc.sample(R.IC, index, edges, w, 0, first, 16, threads=threads())
        assert_same_sets((off, ver), exp)
    # coverage of the first 2048 sets through the column store against the
    # count on their exported lists
    seeds = np.array([0, 1, 2, 3, 5, 8, 13, 21, 34, 55], np.uint32)
    num = 2048
    off, ver = p.export(0, num)
    member = np.isin(ver, seeds)
    hits = np.add.reduceat(member.astype(np.int64), off[:-1].astype(np.int64))
    hits[np.diff(off.astype(np.int64)) == 0] = 0
    assert p.coverage(seeds, num) == int((hits > 0).sum())
    p.close()
    g.close()
