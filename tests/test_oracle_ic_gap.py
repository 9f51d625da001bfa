"""Pins the oracle's IC gap sampling: the fixed-point -log2 against the math
library, its strict monotonicity, the geometric law of the gaps it yields
(exact counts over all 2^32 uniforms, by bisection), and the per-edge live
frequencies and their independence on rows spanning many blocks.  The
reference's rule is an independent coin per in-edge with u <= w
(add_rrrset.h:128-134); the gap form must give each edge that probability,
independently of the others."""
import math
import os
import re

import numpy as np
import pytest

from graphs import transposed_csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ONE = 1 << 58


def neg_log2_exact(u):
    return -math.log2(u / 2.0 ** 32)


def test_neg_log2_against_math(orc):
    rng = np.random.default_rng(0)
    us = [1, 2, 3, 255, 256, 257, (1 << 32) - 1, 1 << 32] + \
        [1 << k for k in range(33)] + rng.integers(1, 1 << 32, 100000).tolist()
    worst = 0.0
    for u in us:
        got = orc.ic_neg_log2(int(u)) / ONE
        worst = max(worst, abs(got - neg_log2_exact(int(u))))
    assert worst < 2.0 ** -40
    # exact at powers of two: the table's entry 0 and an empty series
    for k in range(33):
        assert orc.ic_neg_log2(1 << k) == (32 - k) << 58


def test_neg_log2_strictly_decreasing(orc):
    rng = np.random.default_rng(1)
    us = rng.integers(1, (1 << 32) - 1, 50000).tolist()
    # every table boundary (top 8 mantissa bits change) for every exponent
    for e in range(32):
        for i in range(256):
            b = ((256 + i) << e) >> 8
            us += [b - 1, b, b + 1]
    for u in us:
        if 1 <= u < (1 << 32):
            assert orc.ic_neg_log2(u) > orc.ic_neg_log2(u + 1), u


def test_tables_identical_and_exact():
    def read(path):
        txt = open(path).read()
        return re.findall(r"0x[0-9a-f]{16}ull", txt)
    prod = read(os.path.join(ROOT, "ripples_amd", "csrc", "ic_gap_tables.inc"))
    orac = read(os.path.join(ROOT, "oracle", "ic_gap_tables.inc"))
    assert prod == orac and len(prod) == 513
    vals = [int(x[:-3], 16) for x in prod]
    log2c, recip, inv = vals[:256], vals[256:512], vals[512]
    for i in range(256):
        # log2(1 + i/256) to 2^-64 (double math checks the leading 50 bits)
        assert abs(log2c[i] / 2.0 ** 64 - math.log2(1 + i / 256)) < 2.0 ** -50
        assert recip[i] == (1 << 71) // (256 + i)
    assert abs(inv / 2.0 ** 63 - 1 / math.log(2)) < 2.0 ** -50


def count_u_at_least(orc, need):
    """#{U in [1, 2^32] : NL(U) >= need} by bisection (NL strictly decreasing)."""
    if orc.ic_neg_log2(1) < need:
        return 0
    lo, hi = 1, 1 << 32
    if orc.ic_neg_log2(hi) >= need:
        return hi
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if orc.ic_neg_log2(mid) >= need:
            lo = mid
        else:
            hi = mid
    return lo


@pytest.mark.parametrize("T", [1, 17, 1000, 65536, 1677721, 8388608, 16000000, 16777215])
def test_gap_law_is_geometric(orc, T):
    """P(gap >= j) over all 2^32 uniforms equals (1 - T/2^24)^j, up to the
    log's error (below 2^-41 in q and in each NL value, so j 2^-41 + 2^-41
    in the exponent) and the 2^-32 grid of the uniforms."""
    q = orc.ic_gap_q(T)
    p = T / 2.0 ** 24
    assert abs(q / ONE - (-math.log2(1 - p))) < 2.0 ** -41
    for j in (1, 2, 3, 7, 16, 100, 1000, 1 << 16, 1 << 20):
        tail = count_u_at_least(orc, j * q) / 2.0 ** 32
        bound = (j + 1) * 2.0 ** -41 + 2.0 ** -31
        assert abs(tail - (1 - p) ** j) < bound, (T, j)


def test_gap_special_cases(orc):
    assert orc.ic_gap_q(1 << 24) == 0          # w >= 1: every edge a candidate
    assert orc.ic_gap(123456789 << 32, 0) == 0
    assert orc.ic_threshold(0.0) == 0 and orc.ic_threshold(-1.0) == 0
    assert orc.ic_threshold(float("nan")) == 0
    assert orc.ic_threshold(1.0) == 1 << 24 and orc.ic_threshold(3.0) == 1 << 24
    assert orc.ic_threshold(0.5) == 1 << 23
    # blocks hold about two candidates, 16 to 2^20 edges
    assert orc.ic_block_log2(1 << 24) == 4 and orc.ic_block_log2(1) == 20
    for T in (3, 999, 1 << 20, 5000000):
        b = orc.ic_block_log2(T)
        assert (b == 20 or (T << b) >= 1 << 25) and (b == 4 or (T << (b - 1)) < 1 << 25)


def star(weights):
    """Leaves 1..L each with one edge into vertex 0, weights as given."""
    L = len(weights)
    s = np.arange(1, L + 1)
    d = np.zeros(L, np.int64)
    return transposed_csr(L + 1, s, d, np.asarray(weights, np.float32))


@pytest.mark.parametrize("kind", ["uniform", "mixed", "tiny"])
def test_live_frequencies_and_independence(orc, kind):
    """Rows of 60 in-edges span several blocks (16 edges each at thresholds
    above 2^21); each edge of the hub is live with its weight's probability,
    pairs of edges (same block and different blocks) independently."""
    rng = np.random.default_rng(3)
    if kind == "uniform":
        w = np.full(60, 0.45)
    elif kind == "mixed":
        w = rng.uniform(0.02, 0.9, 60)
        w[5] = 1.0
        w[17] = 0.0
    else:
        w = rng.uniform(0.001, 0.02, 60)    # 512-edge blocks: one block
    index, src, ww = star(w)
    count = 1200000
    off, ver = orc.sample(orc.IC, index, src, ww, 5, 0, count, threads=os.cpu_count())
    n = len(index) - 1
    roots = np.array([orc.sample_root(5, i, n) for i in range(count)])
    hub = np.nonzero(roots == 0)[0]
    assert len(hub) > 15000
    member = np.zeros((len(hub), n), bool)
    for r, i in enumerate(hub):
        member[r, ver[off[i]:off[i + 1]]] = True
    freq = member[:, 1:].mean(axis=0)
    p = np.array([orc.ic_threshold(float(np.float32(x))) / 2.0 ** 24 for x in w])
    sig = np.sqrt(p * (1 - p) / len(hub)) + 1e-9
    assert np.all(np.abs(freq - p) <= 5 * sig + 1e-7), (kind, np.max(np.abs(freq - p) / sig))
    for a, b in ((0, 1), (3, 4), (10, 40), (15, 16), (2, 59)):
        joint = (member[:, 1 + a] & member[:, 1 + b]).mean()
        pj = p[a] * p[b]
        assert abs(joint - pj) <= 5 * math.sqrt(pj * (1 - pj) / len(hub)) + 1e-7, (a, b)
