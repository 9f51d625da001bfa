"""Pins the oracle's RRR sampling (add_rrrset.h:105-157 semantics) with special
cases, textbook reachability, exhaustive live-edge enumeration and agreement
of three independently written traversals."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import breadth_first_order

from graphs import (karate_csr, lt_normalise, random_graph, sets_of,
                    transposed_csr)


def check_well_formed(orc, off, ver, n, seed, first=0):
    for i, st in enumerate(sets_of(off, ver)):
        assert st == sorted(set(st)), "sets are sorted and unique"
        assert all(0 <= v < n for v in st)
        assert orc.sample_root(seed, first + i, n) in st


@pytest.mark.parametrize("algo", [1, 2])
def test_ic_traversals_agree_on_karate(orc, algo):
    index, src, w = karate_csr()
    a = orc.sample(orc.IC, index, src, w, 0, 0, 500, algo=0)
    b = orc.sample(orc.IC, index, src, w, 0, 0, 500, algo=algo)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    check_well_formed(orc, a[0], a[1], len(index) - 1, 0)


@pytest.mark.parametrize("seed", [0, 7])
def test_ic_traversals_agree_on_random_graphs(orc, seed):
    n, s, d, w = random_graph(300, 3000, seed, 0.0, 0.3)
    index, src, w = transposed_csr(n, s, d, w)
    ref = orc.sample(orc.IC, index, src, w, seed, 40, 300, algo=0)
    for algo in (1, 2):
        got = orc.sample(orc.IC, index, src, w, seed, 40, 300, algo=algo)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    check_well_formed(orc, ref[0], ref[1], n, seed, first=40)


@pytest.mark.parametrize("whi", [1.0, 0.05])
def test_ic_traversals_agree_on_multi_block_rows(orc, whi):
    """Hub rows of ~400 in-edges span many gap blocks (16 edges at weights
    near 1, 512 at 0.05); some weights are exactly 0 or 1."""
    rng = np.random.default_rng(5)
    n = 500
    s = np.concatenate([rng.integers(0, n, 4000), np.arange(n), np.arange(n)])
    d = np.concatenate([rng.integers(0, n, 4000), np.full(n, 3), np.full(n, 77)])
    keep = s != d
    key = np.unique(s[keep] * n + d[keep])
    s, d = key // n, key % n
    w = rng.uniform(0.0, whi, len(s)).astype(np.float32)
    w[rng.random(len(s)) < 0.05] = 0.0
    if whi == 1.0:
        w[rng.random(len(s)) < 0.05] = 1.0
    index, src, w = transposed_csr(n, s, d, w)
    ref = orc.sample(orc.IC, index, src, w, 4, 0, 400, algo=0)
    for algo in (1, 2):
        got = orc.sample(orc.IC, index, src, w, 4, 0, 400, algo=algo)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    check_well_formed(orc, ref[0], ref[1], n, 4)


def test_lt_definition_agrees(orc):
    n, s, d, w = random_graph(200, 2000, 3)
    w = lt_normalise(n, s, d, w, 0.9)
    index, src, w = transposed_csr(n, s, d, w)
    a = orc.sample(orc.LT, index, src, w, 3, 0, 400, algo=0)
    b = orc.sample(orc.LT, index, src, w, 3, 0, 400, algo=2)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    check_well_formed(orc, a[0], a[1], n, 3)


def test_threads_do_not_change_results(orc):
    index, src, w = karate_csr()
    a = orc.sample(orc.IC, index, src, w, 1, 0, 1000, threads=1)
    b = orc.sample(orc.IC, index, src, w, 1, 0, 1000, threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_samples_depend_only_on_their_id(orc):
    index, src, w = karate_csr()
    a = orc.sample(orc.IC, index, src, w, 9, 0, 200)
    b = orc.sample(orc.IC, index, src, w, 9, 150, 50)
    assert sets_of(a[0], a[1])[150:] == sets_of(b[0], b[1])


@pytest.mark.parametrize("model", [0, 1])
def test_zero_weights_give_root_only(orc, model):
    index, src, _ = karate_csr()
    w = np.zeros(len(src), np.float32)
    n = len(index) - 1
    off, ver = orc.sample(model, index, src, w, 5, 0, 300)
    for i, st in enumerate(sets_of(off, ver)):
        assert st == [orc.sample_root(5, i, n)]


def test_unit_weights_give_reverse_reachability(orc):
    n, s, d, _ = random_graph(120, 300, 11)
    index, src, w = transposed_csr(n, s, d, np.ones(len(s), np.float32))
    off, ver = orc.sample(orc.IC, index, src, w, 2, 0, 120)
    # in-neighbour graph: edge v -> u for every input edge u -> v
    g = sp.csr_matrix((np.ones(len(s)), (d, s)), shape=(n, n))
    for i, st in enumerate(sets_of(off, ver)):
        r = orc.sample_root(2, i, n)
        reach = breadth_first_order(g, r, directed=True, return_predecessors=False)
        assert st == sorted(reach.tolist())


def test_ic_inclusion_matches_exhaustive_enumeration(orc):
    # 5 vertices, 8 edges: P(u in RRR(r)) by summing over all 2^8 live-edge sets
    s = np.array([0, 1, 2, 3, 0, 2, 4, 1])
    d = np.array([1, 2, 3, 4, 2, 0, 1, 4])
    w = np.array([0.5, 0.3, 0.8, 0.2, 0.6, 0.4, 0.7, 0.1], np.float32)
    n = 5
    exact = np.zeros((n, n))
    for live in itertools.product([0, 1], repeat=len(s)):
        p = np.prod([float(w[e]) if live[e] else 1.0 - float(w[e])
                     for e in range(len(s))])
        for r in range(n):
            seen, stack = {r}, [r]
            while stack:
                v = stack.pop()
                for e in range(len(s)):
                    if live[e] and d[e] == v and s[e] not in seen:
                        seen.add(int(s[e]))
                        stack.append(int(s[e]))
            for u in seen:
                exact[r, u] += p
    index, src, ww = transposed_csr(n, s, d, w)
    count = 200000
    off, ver = orc.sample(orc.IC, index, src, ww, 17, 0, count)
    sizes = np.diff(off).astype(np.int64)
    owner = np.repeat(np.arange(count), sizes)
    roots = np.array([orc.sample_root(17, i, n) for i in range(0, count)])
    freq = np.zeros((n, n))
    np.add.at(freq, (roots[owner], ver.astype(np.int64)), 1)
    per_root = np.bincount(roots, minlength=n)
    freq /= per_root[:, None]
    exact = np.clip(exact, 0.0, 1.0)
    sigma = np.sqrt(exact * (1 - exact) / per_root[:, None]) + 1e-9
    assert np.all(np.abs(freq - exact) <= 5 * sigma + 1e-6)


def test_lt_single_in_edges_give_deterministic_path(orc):
    # chain 0 <- 1 <- 2 <- 3 (edges u -> v = (1,0), (2,1), (3,2)), weights 1
    s = np.array([1, 2, 3])
    d = np.array([0, 1, 2])
    index, src, w = transposed_csr(4, s, d, np.ones(3, np.float32))
    off, ver = orc.sample(orc.LT, index, src, w, 0, 0, 100)
    for i, st in enumerate(sets_of(off, ver)):
        r = orc.sample_root(0, i, 4)
        assert st == list(range(r, 4))


def test_lt_stops_on_revisit(orc):
    s = np.array([0, 1])
    d = np.array([1, 0])
    index, src, w = transposed_csr(2, s, d, np.ones(2, np.float32))
    off, ver = orc.sample(orc.LT, index, src, w, 0, 0, 50)
    assert all(st == [0, 1] for st in sets_of(off, ver))


def test_lt_step_distribution_follows_weights(orc):
    # vertex 0 has in-neighbours 1, 2, 3 with weights .2 .3 .4; they have none
    s = np.array([1, 2, 3])
    d = np.array([0, 0, 0])
    wt = np.array([0.2, 0.3, 0.4], np.float32)
    index, src, w = transposed_csr(4, s, d, wt)
    count = 100000
    off, ver = orc.sample(orc.LT, index, src, w, 4, 0, count)
    sets = sets_of(off, ver)
    roots = [orc.sample_root(4, i, 4) for i in range(count)]
    outcome = {(): 0, (1,): 0, (2,): 0, (3,): 0}
    tot = 0
    for r, st in zip(roots, sets):
        if r != 0:
            continue
        tot += 1
        outcome[tuple(x for x in st if x != 0)] += 1
    expect = {(): 0.1, (1,): 0.2, (2,): 0.3, (3,): 0.4}
    for key, p in expect.items():
        assert abs(outcome[key] / tot - p) < 5 * np.sqrt(p * (1 - p) / tot)
