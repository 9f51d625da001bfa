"""Small graphs for the tests, built with numpy independently of the library."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def transposed_csr(n, src, dst, w):
    """Rows = targets, entries = sources sorted ascending (GraphBwd layout)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.float32)
    order = np.lexsort((src, dst))
    src, dst, w = src[order], dst[order], w[order]
    index = np.zeros(n + 1, np.uint64)
    np.add.at(index, dst + 1, 1)
    index = np.cumsum(index).astype(np.uint64)
    return index, src.astype(np.uint32), w


def karate(weight=None):
    """Karate club (test-data/karate.tsv): 78 directed edges u<v, w=0.5."""
    e = np.loadtxt(os.path.join(GOLDEN, "karate.tsv"))
    s = e[:, 0].astype(np.int64)
    d = e[:, 1].astype(np.int64)
    w = e[:, 2].astype(np.float32) if weight is None else np.full(len(s), weight, np.float32)
    n = int(max(s.max(), d.max())) + 1
    return n, s, d, w


def karate_csr(weight=None):
    n, s, d, w = karate(weight)
    return transposed_csr(n, s, d, w)


def random_graph(n, m, seed, wlo=0.0, whi=1.0, self_loops=False):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    if not self_loops:
        keep = s != d
        s, d = s[keep], d[keep]
    key = np.unique(s * n + d)
    s, d = key // n, key % n
    w = rng.uniform(wlo, whi, len(s)).astype(np.float32)
    return n, s, d, w


def lt_normalise(n, s, d, w, total=0.9):
    """Per-target normalisation of in-weights to `total` (float32)."""
    w = np.asarray(w, np.float32).copy()
    sums = np.zeros(n, np.float32)
    np.add.at(sums, d, w)
    w = (w / sums[d]) * np.float32(total)
    return w.astype(np.float32)


def sets_of(offsets, vertices):
    return [vertices[offsets[i]:offsets[i + 1]].tolist() for i in range(len(offsets) - 1)]
