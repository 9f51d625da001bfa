"""GPU parity: the HIP path against the CPU oracle, bit for bit, through the
C ABI.  RRR sets are compared as sorted vertex lists; seeds, coverage and
theta must be identical."""
import numpy as np
import pytest

import ripples_amd as R
from graphs import (karate_csr, lt_normalise, random_graph, sets_of,
                    transposed_csr)

pytestmark = pytest.mark.gpu


def rmat_csr(scale, scheme, scale_w, seed=0):
    s, d = R.rmat_edges(scale)
    return R.csr_from_edges(1 << scale, s, d, scheme, scale_w, seed)


GRAPHS = {
    "karate_ic": lambda: (R.IC,) + karate_csr(),
    "karate_lt": lambda: (R.LT,) + karate_csr(0.2),
    "random_ic": lambda: (R.IC,) + transposed_csr(*random_graph(500, 6000, 1, 0.0, 0.25)),
    "random_lt": lambda: (R.LT,) + transposed_csr(
        *(lambda n, s, d, w: (n, s, d, lt_normalise(n, s, d, w)))(*random_graph(500, 6000, 2))),
    "rmat10_ic": lambda: (R.IC,) + rmat_csr(10, R.W_UNIFORM, 0.1),
    "rmat12_wc": lambda: (R.IC,) + rmat_csr(12, R.W_CASCADE, 1.0),
    "rmat12_lt": lambda: (R.LT,) + rmat_csr(12, R.W_LT_NORM, 0.9),
}


@pytest.fixture(scope="module", params=sorted(GRAPHS))
def case(request):
    model, index, edges, w = GRAPHS[request.param]()
    g = R.Graph(index, edges, w, model=model)
    yield request.param, model, index, edges, w, g
    g.close()


def assert_same_sets(got, exp):
    assert np.array_equal(got[0], exp[0]), "set sizes differ"
    assert np.array_equal(got[1], exp[1]), "set contents differ"


def test_sampling_matches_oracle(case, orc):
    name, model, index, edges, w, g = case
    seed, count = 11, 3000
    p = g.pool(seed)
    p.generate_rrr_sets(1000)
    p.generate_rrr_sets(count - 1000)       # appended call continues the ids
    got = p.export()
    exp = orc.sample(model, index, edges, w, seed, 0, count)
    assert_same_sets(got, exp)
    part = p.export(1500, 700)
    assert_same_sets(part, orc.sample(model, index, edges, w, seed, 1500, 700))
    p.close()


def test_selection_matches_oracle(case, orc):
    name, model, index, edges, w, g = case
    p = g.pool(3)
    p.generate_rrr_sets(4000)
    n = len(index) - 1
    off, ver = p.export()
    for k in (1, 5, 50):
        for num in (4000, 1234):
            seeds, cov = p.find_most_influential_set(k, num)
            es, ec = orc.select(n, off[:num + 1], ver[:int(off[num])], k)
            assert seeds.tolist() == es.tolist() and cov == ec
    assert np.array_equal(p.count_occurrences(), orc.count(n, off, ver))
    p.close()


@pytest.mark.parametrize("sampler", [0, 1, 3])
def test_ic_bimodal_sets_match_oracle(sampler, orc):
    """R-MAT scale 16 with U[0,0.1] weights: most sets are small, a fifth reach
    the giant component of ~12K vertices (the wave kernel hands those to the
    colour groups or, with WAVE, to the large-set kernel); appended in three
    calls of uneven size."""
    index, edges, w = rmat_csr(16, R.W_UNIFORM, 0.1, 3)
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(6)
    for c in (700, 1, 1299):
        p.generate_rrr_sets(c, tune={"sampler": sampler})
    got = p.export()
    assert_same_sets(got, orc.sample(R.IC, index, edges, w, 6, 0, 2000))
    assert (np.diff(got[0]) > 4096).sum() > 0
    p.close()
    g.close()


def dense_pool_sets(n, count, seed):
    """Sorted, duplicate-free sets over [0, n): mostly a few hundred to a few
    thousand vertices, some empty or single, 1% the whole vertex set, and a
    few pinned to both sides of the 8192/16384 vertex range boundaries."""
    rng = np.random.default_rng(seed)
    pinned = [[8191, 8192], [16383, 16384],
              [b + o for b in range(8192, n, 8192) for o in (-1, 0)]]
    sets = [[], [0], [n - 1], list(range(n))] + [[v for v in x if v < n] for x in pinned]
    while len(sets) < count:
        u = rng.random()
        if u < 0.2:
            size = int(rng.integers(0, 4))
        elif u < 0.99:
            size = int(rng.integers(200, 3000))
        else:
            size = n
        sets.append(np.unique(rng.integers(0, n, size)).tolist() if size < n
                    else list(range(n)))
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in sets])
    ver = np.fromiter((v for x in sets for v in x), np.uint32, int(off[-1]))
    return off, ver


@pytest.mark.parametrize("n", [5000, 40000, 70001])
def test_dense_selection_matches_oracle(n, orc):
    """Pools of long sets (4M+ elements, mean length at least 64 per range of
    16K vertices) count and index per vertex range in LDS; shorter prefixes
    take the per-element kernels.  Counts, seeds and coverage equal the
    oracle's on every prefix."""
    off, ver = dense_pool_sets(n, 5000, n)
    assert off[-1] >= 4 << 20
    g = R.Graph(np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    p = g.pool(0)
    p.append(off[:2001], ver[:int(off[2000])])
    p.append(off[2000:] - off[2000], ver[int(off[2000]):])
    assert np.array_equal(p.count_occurrences(), orc.count(n, off, ver))
    for num in (5000, 3001, 700, 9):
        for k in (1, 40):
            seeds, cov = p.find_most_influential_set(k, num)
            es, ec = orc.select(n, off[:num + 1], ver[:int(off[num])], k)
            assert seeds.tolist() == es.tolist() and cov == ec, (num, k)
    p.close()
    g.close()


def test_dense_selection_on_ic_pool(orc):
    """R-MAT scale 16, U[0,0.1]: a fifth of the sets reach the ~12K-vertex
    giant component, so the 6000-set pool takes the dense kernels."""
    index, edges, w = rmat_csr(16, R.W_UNIFORM, 0.1, 3)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(9)
    p.generate_rrr_sets(6000)
    off, ver = p.export()
    assert off[-1] >= 4 << 20
    assert np.array_equal(off, orc.sample(R.IC, index, edges, w, 9, 0, 6000)[0])
    for num, k in ((6000, 100), (4500, 7)):
        seeds, cov = p.find_most_influential_set(k, num)
        es, ec = orc.select(n, off[:num + 1], ver[:int(off[num])], k)
        assert seeds.tolist() == es.tolist() and cov == ec, (num, k)
    p.close()
    g.close()


def test_selector_graph_replays_across_pool_growth(orc):
    """The seed rounds are one captured graph per (pool, k): replayed after the
    pool's buffers grew and moved, and on different prefixes, each replay
    equals the oracle."""
    index, edges, w = rmat_csr(10, R.W_UNIFORM, 0.1, 2)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(2)
    for total in (500, 3000, 20000, 60000):
        p.generate_rrr_sets(total - p.size()[0])
        off, ver = p.export()
        for num in (total, total // 3, 1):
            seeds, cov = p.find_most_influential_set(8, num)
            es, ec = orc.select(n, off[:num + 1], ver[:int(off[num])], 8)
            assert seeds.tolist() == es.tolist() and cov == ec, (total, num)
    p.close()
    g.close()


@pytest.mark.parametrize("tune", [dict(colors=1, sampler=2), dict(colors=17, block=256, sampler=2),
                                  dict(colors=64, block=512, sampler=2),
                                  dict(colors=32, block=1024, slots=3, sampler=2),
                                  dict(colors=64, block=256, slots=1, sampler=2),
                                  dict(colors=64, heavy_degree=1, sampler=2),
                                  dict(colors=64, heavy_degree=1 << 30, sampler=2),
                                  dict(colors=7, heavy_degree=3, sampler=2),
                                  dict(sampler=1), dict(sampler=2), dict(sampler=3),
                                  dict(queue_cap=64), dict(queue_cap=256, sampler=3),
                                  dict(queue_cap=64, out_chunk=3),
                                  dict(sampler=2, colors=64, heavy_degree=64),
                                  dict(out_chunk=1), dict(out_chunk=5, sampler=1),
                                  dict(out_chunk=1 << 20),
                                  dict(frontier_cap=7), dict(frontier_cap=64, queue_cap=64),
                                  dict(block=512), dict(block=1024, queue_cap=64),
                                  dict(queue_cap=4096), dict(queue_cap=4096, sampler=3),
                                  dict(queue_cap=2048, block=512), dict(colors=32),
                                  dict(colors=5, queue_cap=64), dict(wave_stage=1),
                                  dict(wave_stage=1, queue_cap=64),
                                  dict(wave_stage=1, queue_cap=256, block=512, colors=9)])
def test_ic_tuning_does_not_change_results(tune, orc):
    index, edges, w = rmat_csr(12, R.W_UNIFORM, 0.1, 1)
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(5)
    p.generate_rrr_sets(2500, tune=tune)
    exp = orc.sample(R.IC, index, edges, w, 5, 0, 2500)
    assert_same_sets(p.export(), exp)
    assert p.size() == (2500, int(exp[0][-1]))
    p.close()
    g.close()


def test_selector_reference_fixture():
    sets = [[2, 3] if i % 2 else [2, 5] for i in range(1000)]
    sets += [[[3, 4, 5], [1, 3], [1, 5]][i % 3] for i in range(99)]
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in sets])
    ver = np.array([v for s in sets for v in s], np.uint32)
    index = np.zeros(7, np.uint64)
    g = R.Graph(index, np.zeros(0, np.uint32), np.zeros(0, np.float32))
    p = g.pool(0)
    p.append(off, ver)
    assert p.count_occurrences().tolist() == [0, 66, 1000, 566, 33, 566]
    seeds, cov = p.find_most_influential_set(1)
    assert seeds.tolist() == [2] and cov == 1000
    seeds, cov = p.find_most_influential_set(10)
    assert seeds.tolist() == [2, 1, 3] and cov == 1099
    got = p.export()
    assert np.array_equal(got[0], off) and np.array_equal(got[1], ver)


def test_pool_growth_retry(orc):
    # every set is the whole reverse closure: far beyond the first estimate
    n, s, d, _ = random_graph(3000, 30000, 4)
    index, edges, w = transposed_csr(n, s, d, np.ones(len(s), np.float32))
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(1)
    p.generate_rrr_sets(600)
    assert_same_sets(p.export(), orc.sample(R.IC, index, edges, w, 1, 0, 600))
    p.close()
    g.close()


@pytest.mark.parametrize("sampler", [0, 1, 3, 4])
@pytest.mark.parametrize("n", [1500, 3000, 6000, 20000, 70000])
def test_ic_large_sets(n, sampler, orc):
    """Sets of ~n vertices: LDS sort (<= 1024), HBM sort (<= 4096) and, for
    sets a wave's list cannot hold, the colour groups (AUTO) or the large-set
    kernel (WAVE: bitmap visited, summary walk over several passes of 64
    summary words at n = 70000).  4: AUTO with the wave stage (lanes up to
    64 vertices, then one BFS per wave 64 vertices at a time, then the grid)."""
    rng = np.random.default_rng(n)
    s = rng.integers(0, n, 8 * n)
    d = rng.integers(0, n, 8 * n)
    keep = s != d
    key = np.unique(s[keep] * n + d[keep])
    s, d = key // n, key % n
    w = np.where(rng.random(len(s)) < 0.9, 1.0, 0.0).astype(np.float32)
    index, edges, w = transposed_csr(n, s, d, w)
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(4)
    cnt = 300 if n <= 20000 else 80
    tune = {"sampler": sampler} if sampler < 4 else {"wave_stage": 1, "queue_cap": 64}
    p.generate_rrr_sets(cnt, tune=tune)
    got = p.export()
    assert_same_sets(got, orc.sample(R.IC, index, edges, w, 4, 0, cnt))
    assert np.diff(got[0]).max() > min(n - 100, 4096) or n < 4096
    p.close()
    g.close()


def test_lt_long_walks(orc):
    # a 300-vertex chain with unit weights: walks far longer than the LDS list
    n = 300
    s = np.arange(1, n)
    d = np.arange(0, n - 1)
    index, edges, w = transposed_csr(n, s, d, np.ones(n - 1, np.float32))
    g = R.Graph(index, edges, w, model=R.LT)
    p = g.pool(2)
    p.generate_rrr_sets(500)
    got = p.export()
    assert_same_sets(got, orc.sample(R.LT, index, edges, w, 2, 0, 500))
    assert max(np.diff(got[0])) > 64
    p.close()
    g.close()


def lt_hub_graph(kind, n=2500, hubs=8, seed=7):
    """Hubs that are in- and out-neighbours of every vertex, so walks keep
    scanning rows of ~n in-edges, plus sparse random edges.  Weight kinds:
    normalised uniform, dyadic k*2^-q per row (exact grid values and halfway
    ties), a wide exponent range with subnormals, zeros and -0, and rows with
    inf / NaN weights that must be scanned one edge at a time."""
    rng = np.random.default_rng(seed)
    allv = np.arange(n)
    src, dst = [], []
    for h in range(hubs):
        others = allv[allv != h]
        src += [others, np.full(len(others), h)]
        dst += [np.full(len(others), h), others]
    s = rng.integers(0, n, 6 * n)
    d = rng.integers(0, n, 6 * n)
    src.append(s)
    dst.append(d)
    s, d = np.concatenate(src), np.concatenate(dst)
    keep = s != d
    key = np.unique(s[keep] * n + d[keep])
    s, d = key // n, key % n
    deg = np.bincount(d, minlength=n)
    if kind == "dyadic":
        q = np.floor(np.log2(deg[d] * 4.0)).astype(np.int64)
        w = np.ldexp(rng.integers(0, 8, len(s)).astype(np.float64), -q).astype(np.float32)
    elif kind == "wide":
        w = np.ldexp(rng.random(len(s)), -rng.integers(0, 40, len(s))).astype(np.float32)
        w = lt_normalise(n, s, d, w, 0.95)
        r = rng.random(len(s))
        w[r < 0.01] = np.ldexp(np.float32(1.0), -130 - rng.integers(0, 19, int((r < 0.01).sum())))
        w[(r >= 0.01) & (r < 0.03)] = 0.0
        w[(r >= 0.03) & (r < 0.04)] = -0.0
    else:
        w = lt_normalise(n, s, d, rng.random(len(s)).astype(np.float32), 0.97)
    index, edges, w = transposed_csr(n, s, d, w)
    if kind == "scalar":
        for h, bad in ((0, np.inf), (1, np.nan)):
            b, e = int(index[h]), int(index[h + 1])
            w[b + (e - b) // 2] = bad
    return index, edges, w


@pytest.mark.parametrize("kind", ["uniform", "dyadic", "wide", "scalar"])
def test_lt_hub_rows(kind, orc):
    """Breakpoint searches on long rows give the float chain's exact choice;
    rows with inf / NaN weights keep the subtraction loop."""
    index, edges, w = lt_hub_graph(kind)
    g = R.Graph(index, edges, w, model=R.LT)
    p = g.pool(9)
    p.generate_rrr_sets(6000)
    stats = p.last_stats()
    assert_same_sets(p.export(), orc.sample(R.LT, index, edges, w, 9, 0, 6000))
    assert stats["coins"] > 0          # breakpoints were searched
    p.close()
    g.close()


@pytest.mark.parametrize("model,k,eps", [(R.IC, 5, 0.5), (R.LT, 10, 0.3)])
def test_imm_matches_oracle(model, k, eps, orc):
    if model == R.IC:
        index, edges, w = rmat_csr(10, R.W_UNIFORM, 0.1)
    else:
        index, edges, w = rmat_csr(10, R.W_LT_NORM, 0.9)
    g = R.Graph(index, edges, w, model=model)
    seeds, rec = R.imm(g, k, eps, seed=0)
    es, erec = orc.imm(model, index, edges, w, k, eps, 0)
    assert seeds.tolist() == es.tolist()
    assert rec["Theta"] == erec.theta and rec["Rounds"] == erec.rounds
    assert rec["ThetaPrime"] == list(erec.theta_prime[:erec.rounds])
    assert rec["F"] == list(erec.f[:erec.rounds])
    assert rec["FinalF"] == erec.final_f
    # elements of the pool's sets (the LT pool's data also holds unused gaps)
    pool_off = orc.sample(model, index, edges, w, 0, 0, rec["NumSets"])[0]
    assert rec["RRRElements"] == int(pool_off[-1])
    assert rec["PoolBytes"] >= 12 * rec["NumSets"]   # at least the per-set offsets and lengths
    g.close()


@pytest.mark.parametrize("chunk", [1, 7, 100, 4096])
def test_lt_output_chunks_do_not_change_results(chunk, orc):
    """LT waves reserve pool space a chunk at a time and leave the rest of
    their last chunk unused: sets, element count and seeds are the same for
    every chunk size, across appended calls and a host append."""
    index, edges, w = rmat_csr(12, R.W_LT_NORM, 0.9)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.LT)
    p = g.pool(5)
    p.generate_rrr_sets(3000, tune={"out_chunk": chunk})
    p.generate_rrr_sets(1500, tune={"out_chunk": chunk})
    exp = orc.sample(R.LT, index, edges, w, 5, 0, 4500)
    assert_same_sets(p.export(), exp)
    assert p.size() == (4500, int(exp[0][-1]))
    p.append(np.array([0, 2, 3], np.uint64), np.array([1, 7, 4], np.uint32))
    off, ver = p.export()
    assert p.size() == (4502, int(exp[0][-1]) + 3)
    assert ver[-3:].tolist() == [1, 7, 4]
    seeds, cov = p.find_most_influential_set(20)
    es, ec = orc.select(n, off, ver, 20)
    assert seeds.tolist() == es.tolist() and cov == ec
    p.close()
    g.close()


def test_scale16_properties(orc):
    """R-MAT scale 16, IC U[0,0.1]: oracle parity on a sample of ids."""
    index, edges, w = rmat_csr(16, R.W_UNIFORM, 0.1)
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(0)
    p.generate_rrr_sets(20000)
    off, ver = p.export()
    sizes = np.diff(off)
    assert sizes.min() >= 1
    exp = orc.sample(R.IC, index, edges, w, 0, 17000, 3000)
    assert_same_sets(p.export(17000, 3000), exp)
    p.close()
    g.close()


@pytest.mark.parametrize("name", ["karate_ic", "karate_lt", "random_ic", "random_lt",
                                  "rmat12_wc", "rmat12_lt"])
def test_spread_matches_oracle(name, orc):
    """Monte Carlo spread (diffusion_simulation.h:73-178) replica by replica."""
    model, index, edges, w = GRAPHS[name]()
    g = R.Graph(index, edges, w, model=model)
    n = len(index) - 1
    deg = np.diff(index.astype(np.int64))
    seeds = np.argsort(-deg)[:5].astype(np.uint32)
    seeds = np.concatenate([seeds, seeds[:2]])        # duplicates are allowed
    act, lev = g.spread(seeds, 400, seed=3, first=50)
    ea, el = orc.spread(model, index, edges, w, seeds, 3, 50, 400)
    assert np.array_equal(act, ea) and np.array_equal(lev, el)
    act2, lev2 = g.spread(seeds, 100, seed=3, first=150)   # replica ids are stable
    assert np.array_equal(act2, ea[100:200]) and np.array_equal(lev2, el[100:200])
    assert act.min() >= len(set(seeds.tolist())) and act.max() <= n
    g.close()


def test_spread_unit_weights_and_errors():
    index, edges, w = rmat_csr(12, R.W_CONST, 1.0)
    g = R.Graph(index, edges, w, model=R.IC)
    act, lev = g.spread([0], 20)
    assert len(set(act.tolist())) == 1 and len(set(lev.tolist())) == 1   # w = 1: deterministic
    with pytest.raises(R.RipplesError, match="out of range"):
        g.spread([1 << 12], 5)
    g.close()


@pytest.mark.parametrize("model", [R.IC, R.LT])
@pytest.mark.parametrize("k,eps", [(5, 0.3), (10, 0.15)])
def test_opimc_matches_oracle(model, k, eps, orc):
    """OPIM-C (opimc.h:150-241): seeds, rounds and every round's coverages and
    bounds equal the oracle's."""
    if model == R.IC:
        index, edges, w = rmat_csr(10, R.W_UNIFORM, 0.1)
    else:
        index, edges, w = rmat_csr(10, R.W_LT_NORM, 0.9)
    g = R.Graph(index, edges, w, model=model)
    seeds, rec = R.opimc(g, k, eps, 0.05, seed=3)
    es, erec = orc.opimc(model, index, edges, w, k, eps, 0.05, 3)
    assert seeds.tolist() == es.tolist()
    r = erec.rounds
    assert rec["Rounds"] == r and rec["Theta"] == erec.theta
    assert (rec["ThetaZero"], rec["ThetaMax"]) == (erec.theta_zero, erec.theta_max)
    assert rec["Coverage1"] == list(erec.coverage1[:r])
    assert rec["Coverage2"] == list(erec.coverage2[:r])
    assert rec["RRRSetsGenerated"] == list(erec.rrr_sets_generated[:r])
    assert rec["Alpha"] == list(erec.alpha[:r])
    assert rec["Lower"] == list(erec.lower[:r]) and rec["Upper"] == list(erec.upper[:r])
    g.close()


def test_pool_coverage(orc):
    index, edges, w = rmat_csr(10, R.W_UNIFORM, 0.1)
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(4)
    p.generate_rrr_sets(5000)
    off, ver = p.export()
    n = len(index) - 1
    rng = np.random.default_rng(0)
    for seeds in ([], [0], rng.choice(n, 7, replace=False), rng.choice(n, 300, replace=False)):
        for ns in (5000, 1234, 0):
            want = orc.coverage(n, off[:ns + 1], ver[:int(off[ns])], seeds)
            assert p.coverage(seeds, ns) == want
    with pytest.raises(R.RipplesError, match="out of range"):
        p.coverage([n], 10)
    p.close()
    g.close()


@pytest.mark.parametrize("model", [R.IC, R.LT])
def test_autotune_keeps_results(model, orc):
    """rip_autotune times every candidate setting and the graph then samples
    with the fastest; the sets stay those of the sampling contract, and a
    graph of the same (model, skew, size) key takes the cached choice."""
    if model == R.IC:
        index, edges, w = rmat_csr(12, R.W_UNIFORM, 0.1, 1)
    else:
        index, edges, w = rmat_csr(12, R.W_LT_NORM, 0.9, 1)
    g = R.Graph(index, edges, w, model=model)
    best, ms, tried = g.autotune(probe_sets=2048, force=True)
    assert tried == (19 if model == R.IC else 4) and ms > 0
    if model == R.IC:
        assert best["queue_cap"] in (64, 256, 1024, 2048, 4096) and best["block"] in (256, 512, 1024)
    p = g.pool(5)
    p.generate_rrr_sets(2500)
    assert_same_sets(p.export(), orc.sample(model, index, edges, w, 5, 0, 2500))
    p.close()
    g2 = R.Graph(index, edges, w, model=model)
    best2, ms2, tried2 = g2.autotune(probe_sets=2048)
    assert tried2 == 0 and best2 == best
    g2.close()
    g.close()
