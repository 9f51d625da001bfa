"""Pins the oracle's forward Monte Carlo spread (orc_spread), the restatement
of include/ripples/diffusion_simulation.h:73-178 under the replica stream
contract (replica r = sample stream r; element 1 + f = IC coin of forward
edge f, element 1 + v = LT threshold of v)."""
import itertools
import math

import numpy as np
import pytest

from graphs import karate_csr, random_graph, transposed_csr
from oracle import oracle as O

IC, LT = 0, 1


def forward(index, src, w):
    """Forward CSR from the transposed one, rows in transposed-position order."""
    n = len(index) - 1
    rows = [[] for _ in range(n)]
    for v in range(n):
        for i in range(int(index[v]), int(index[v + 1])):
            rows[int(src[i])].append((v, float(w[i]), i))
    f = 0
    out = []
    for u in range(n):
        r = []
        for v, wt, i in rows[u]:
            r.append((f, v, wt, i))
            f += 1
        out.append(r)
    return out


def reference_ic(index, src, w, seeds, seed, r):
    """diffusion_simulation.h:73-110, with the coin of forward edge f taken
    from element 1 + f of replica r's stream."""
    fw = forward(index, src, w)
    n = len(index) - 1
    visited = [False] * n
    queue = []
    for s in seeds:
        queue.append(s)
        visited[s] = True
    itr, level_end, level = 0, len(queue), 0
    while itr < len(queue):
        v = queue[itr]
        for f, u, wt, _ in fw[v]:
            if not visited[u] and np.float32(O.u01(O.sample_element(seed, r, 1 + f))) <= np.float32(wt):
                visited[u] = True
                queue.append(u)
        itr += 1
        if itr == level_end:
            level_end = len(queue)
            level += 1
    return sum(visited), level


def reference_lt(index, src, w, seeds, seed, r):
    """diffusion_simulation.h:125-178: candidates are the inactive
    out-neighbours of all active vertices; totals in transposed CSR order."""
    fw = forward(index, src, w)
    n = len(index) - 1
    th = [np.float32(O.u01(O.sample_element(seed, r, 1 + v))) for v in range(n)]
    active = set(seeds)
    level = 0
    while True:
        tobe = set()
        for v in sorted(active):
            for _, u, _, _ in fw[v]:
                if u not in active:
                    tobe.add(u)
        new = set()
        for v in sorted(tobe):
            total = np.float32(0)
            for i in range(int(index[v]), int(index[v + 1])):
                if int(src[i]) in active:
                    total = np.float32(total + np.float32(w[i]))
            if total >= th[v]:
                new.add(v)
        active |= new
        level += 1
        if not new:
            break
    return len(active), level


def test_ic_matches_reference_loop():
    n, s, d, w = random_graph(40, 160, 3, 0.0, 0.6)
    index, src, wt = transposed_csr(n, s, d, w)
    seeds = [0, 5, 5, 17]
    act, lev = O.spread(IC, index, src, wt, seeds, 4, 10, 25)
    for i in range(25):
        assert (act[i], lev[i]) == reference_ic(index, src, wt, seeds, 4, 10 + i)


def test_lt_matches_reference_loop():
    n, s, d, w = random_graph(40, 200, 5)
    from graphs import lt_normalise
    w = lt_normalise(n, s, d, w, 0.95)
    index, src, wt = transposed_csr(n, s, d, w)
    seeds = [1, 2, 30]
    act, lev = O.spread(LT, index, src, wt, seeds, 7, 3, 25)
    for i in range(25):
        assert (act[i], lev[i]) == reference_lt(index, src, wt, seeds, 7, 3 + i)


def bfs_layers(index, src, n, seeds):
    fw = forward(index, src, np.ones(len(src)))
    seen = set(seeds)
    layer = sorted(set(seeds))
    layers = 0
    while layer:
        layers += 1
        nxt = []
        for u in layer:
            for _, v, _, _ in fw[u]:
                if v not in seen:
                    seen.add(v)
                    nxt.append(v)
        layer = nxt
    return len(seen), layers


def test_ic_unit_and_zero_weights():
    index, src, _ = karate_csr()
    n = len(index) - 1
    ones = np.ones(len(src), np.float32)
    zeros = np.zeros(len(src), np.float32)
    seeds = [1, 3]
    act, lev = O.spread(IC, index, src, ones, seeds, 0, 0, 50)
    want = bfs_layers(index, src, n, seeds)
    assert np.all(act == want[0]) and np.all(lev == want[1])
    act, lev = O.spread(IC, index, src, zeros, seeds + [3], 0, 0, 50)
    assert np.all(act == 2) and np.all(lev == 1)


def test_ic_expected_spread_by_enumeration():
    # 4 vertices, 6 edges: E[spread] over all 2^6 live-edge worlds
    s = np.array([0, 0, 1, 2, 1, 3])
    d = np.array([1, 2, 3, 3, 2, 0])
    p = np.array([0.5, 0.25, 0.75, 0.5, 0.125, 0.9], np.float32)
    index, src, w = transposed_csr(4, s, d, p)
    pf = [float(np.floor(np.float64(x) * 2 ** 24) / 2 ** 24) for x in p]  # P(u <= w)
    exp = 0.0
    for live in itertools.product([0, 1], repeat=6):
        prob = np.prod([pf[i] if live[i] else 1 - pf[i] for i in range(6)])
        reach = {0}
        changed = True
        while changed:
            changed = False
            for i in range(6):
                if live[i] and s[i] in reach and d[i] not in reach:
                    reach.add(int(d[i]))
                    changed = True
        exp += prob * len(reach)
    R = 40000
    act, _ = O.spread(IC, index, src, w, [0], 11, 0, R)
    sd = act.astype(np.float64).std()
    assert abs(act.mean() - exp) < 4 * sd / math.sqrt(R)


def test_lt_unit_path_and_star():
    k = 12
    index, src, w = transposed_csr(k + 1, np.arange(k), np.arange(1, k + 1), np.ones(k, np.float32))
    act, lev = O.spread(LT, index, src, w, [0], 2, 0, 20)
    assert np.all(act == k + 1) and np.all(lev == k + 1)
    # star: centre 0 -> 1..d, single in-edge of weight p each: Binomial(d, P(u <= p))
    dd, p = 400, 0.3
    index, src, w = transposed_csr(dd + 1, np.zeros(dd, np.int64), np.arange(1, dd + 1),
                                   np.full(dd, p, np.float32))
    R = 2000
    act, lev = O.spread(LT, index, src, w, [0], 9, 0, R)
    q = np.floor(np.float64(np.float32(p)) * 2 ** 24) / 2 ** 24
    mean = 1 + dd * q
    assert abs(act.mean() - mean) < 4 * math.sqrt(dd * q * (1 - q) / R)
    assert set(lev.tolist()) <= {1, 2}


def test_replica_ids_are_stable():
    n, s, d, w = random_graph(60, 300, 9, 0.0, 0.4)
    index, src, wt = transposed_csr(n, s, d, w)
    a1, l1 = O.spread(IC, index, src, wt, [3, 4], 5, 0, 40, threads=1)
    a2, l2 = O.spread(IC, index, src, wt, [3, 4], 5, 17, 23, threads=4)
    assert np.array_equal(a1[17:], a2) and np.array_equal(l1[17:], l2)
