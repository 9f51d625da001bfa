"""The column store of large sets against the oracle.  Samples that reach a
giant component (R-MAT with U[0,0.1] weights) are drawn 64 at a time by colour
groups; groups with many members among the hot vertices keep those as bits of
a column.  Exported sets, sizes, occurrence counts, coverage, seed selection
on every prefix and full IMM runs must equal the oracle's, with sets in
columns, in lists and appended from the host mixed in one pool."""
import numpy as np
import pytest

import ripples_amd as R

pytestmark = pytest.mark.gpu


def rmat_csr(scale, scheme, scale_w, seed=0):
    s, d = R.rmat_edges(scale)
    return R.csr_from_edges(1 << scale, s, d, scheme, scale_w, seed)


def assert_same_sets(got, exp):
    assert np.array_equal(got[0], exp[0]), "set sizes differ"
    assert np.array_equal(got[1], exp[1]), "set contents differ"


@pytest.fixture(scope="module")
def giant():
    index, edges, w = rmat_csr(16, R.W_UNIFORM, 0.1, 3)
    g = R.Graph(index, edges, w, model=R.IC)
    yield index, edges, w, g
    g.close()


def test_column_sets_match_oracle(giant, orc):
    index, edges, w, g = giant
    n = len(index) - 1
    p = g.pool(11)
    for c in (1500, 7, 2493):
        p.generate_rrr_sets(c)
    groups, hot, nbytes = p.columns()
    assert groups > 0 and hot == n // 4 and nbytes >= groups * hot * 8
    exp = orc.sample(R.IC, index, edges, w, 11, 0, 4000)
    assert_same_sets(p.export(), exp)
    assert p.size() == (4000, int(exp[0][-1]))
    assert_same_sets(p.export(1234, 1000), orc.sample(R.IC, index, edges, w, 11, 1234, 1000))
    off, ver = exp
    assert np.array_equal(p.count_occurrences(), orc.count(n, off, ver))
    assert np.array_equal(p.count_occurrences(1777), orc.count(n, off[:1778], ver[:int(off[1777])]))
    for k, num in ((1, 4000), (25, 4000), (60, 2222), (5, 33)):
        seeds, cov = p.find_most_influential_set(k, num)
        es, ec = orc.select(n, off[:num + 1], ver[:int(off[num])], k)
        assert seeds.tolist() == es.tolist() and cov == ec, (k, num)
    rng = np.random.default_rng(1)
    for seeds in ([], [0], rng.choice(n, 5, replace=False), rng.choice(n, 400, replace=False)):
        for num in (4000, 999):
            want = orc.coverage(n, off[:num + 1], ver[:int(off[num])], seeds)
            assert p.coverage(seeds, num) == want
    p.close()


def test_columns_with_host_sets_and_growth(giant, orc):
    """Host-appended lists between sampled calls; the selector's captured
    rounds replayed as the column store grows."""
    index, edges, w, g = giant
    n = len(index) - 1
    p = g.pool(5)
    # 64-colour groups: small calls would otherwise use a few colours per
    # group, too few members for a column
    full = {"colors": 64}
    p.generate_rrr_sets(1000, tune=full)
    extra_off = np.array([0, 3, 3, 6], np.uint64)
    extra_ver = np.array([1, 2, 3, 100, 200, 60000], np.uint32)
    p.append(extra_off, extra_ver)
    p.generate_rrr_sets(2000, tune=full)
    a = orc.sample(R.IC, index, edges, w, 5, 0, 1000)
    b = orc.sample(R.IC, index, edges, w, 5, 1003, 2000)
    off = np.concatenate([a[0], a[0][-1] + extra_off[1:], a[0][-1] + extra_off[-1] + b[0][1:]])
    ver = np.concatenate([a[1], extra_ver, b[1]])
    assert p.columns()[0] > 0
    assert_same_sets(p.export(), (off, ver))
    for num in (3003, 1002, 500):
        seeds, cov = p.find_most_influential_set(30, num)
        es, ec = orc.select(n, off[:num + 1], ver[:int(off[num])], 30)
        assert seeds.tolist() == es.tolist() and cov == ec, num
    p.reset()
    assert p.columns()[0] == 0
    p.generate_rrr_sets(600, tune=full)
    assert_same_sets(p.export(), orc.sample(R.IC, index, edges, w, 5, 0, 600))
    p.close()


@pytest.mark.parametrize("k,eps", [(50, 0.5), (20, 0.3)])
def test_imm_on_column_pools_matches_oracle(giant, orc, k, eps):
    index, edges, w, g = giant
    seeds, rec = R.imm(g, k, eps, seed=2)
    es, erec = orc.imm(R.IC, index, edges, w, k, eps, 2)
    assert seeds.tolist() == es.tolist()
    r = erec.rounds
    assert rec["Theta"] == erec.theta and rec["Rounds"] == r
    assert rec["ThetaPrime"] == list(erec.theta_prime[:r])
    assert rec["F"] == list(erec.f[:r]) and rec["FinalF"] == erec.final_f
