"""rip_pool_coverage_noalloc (no device allocation, fixed library scratch)
against rip_pool_coverage and the oracle's FindCoverage, through the C ABI:
two calls in a row on each pool, the second on a smaller input, so stale
scratch contents of the first call would show; plain pools and pools with
column sets; the size limits."""
import numpy as np
import pytest

import ripples_amd as R

pytestmark = pytest.mark.gpu


def rmat_csr(scale, scheme, scale_w, seed=0):
    s, d = R.rmat_edges(scale)
    return R.csr_from_edges(1 << scale, s, d, scheme, scale_w, seed)


def check_pair(p, n, num_sets, seeds, orc, off, ver):
    got = p.coverage_noalloc(seeds, num_sets)
    want = p.coverage(seeds, num_sets)
    assert got == want, (num_sets, len(seeds))
    assert got == orc.coverage(n, off[:num_sets + 1], ver[:int(off[num_sets])], seeds)
    return got


def test_noalloc_coverage_plain_pool(orc):
    index, edges, w = rmat_csr(12, R.W_CASCADE, 1.0)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(3)
    p.generate_rrr_sets(6000)
    off, ver = p.export()
    rng = np.random.default_rng(4)
    big = rng.choice(n, 300, replace=False).astype(np.uint32)
    small = big[:3][::-1].copy()
    a = check_pair(p, n, 6000, big, orc, off, ver)
    b = check_pair(p, n, 1000, small, orc, off, ver)
    assert a > 0 and b > 0
    # no seeds, no sets, seed 0 and n - 1
    assert p.coverage_noalloc([], 6000) == 0
    assert p.coverage_noalloc(big, 0) == 0
    check_pair(p, n, 6000, np.array([0, n - 1], np.uint32), orc, off, ver)
    p.close()
    g.close()


def test_noalloc_coverage_column_pool(orc):
    index, edges, w = rmat_csr(16, R.W_UNIFORM, 0.1, 3)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(11)
    p.generate_rrr_sets(2500, tune={"colors": 64})
    assert p.columns()[0] > 0
    off, ver = p.export()
    rng = np.random.default_rng(9)
    big = rng.choice(n, 500, replace=False).astype(np.uint32)
    check_pair(p, n, 2500, big, orc, off, ver)
    # hot vertices only hit column sets through their columns
    check_pair(p, n, 777, np.array([int(np.bincount(ver, minlength=n).argmax())], np.uint32),
               orc, off, ver)
    check_pair(p, n, 65, big[:2].copy(), orc, off, ver)
    p.close()
    g.close()


def test_noalloc_coverage_limits():
    index, edges, w = rmat_csr(10, R.W_CASCADE, 1.0)
    n = len(index) - 1
    g = R.Graph(index, edges, w, model=R.IC)
    p = g.pool(0)
    p.generate_rrr_sets(100)
    with pytest.raises(R.RipplesError, match="65536 seeds"):
        p.coverage_noalloc(np.zeros(65537, np.uint32))
    with pytest.raises(R.RipplesError, match="out of range"):
        p.coverage_noalloc(np.array([n], np.uint32))
    with pytest.raises(R.RipplesError, match="beyond the pool"):
        p.coverage_noalloc(np.array([1], np.uint32), 101)
    # the pool still answers after the rejected calls
    assert p.coverage_noalloc(np.arange(n, dtype=np.uint32)) == 100
    p.close()
    g.close()
