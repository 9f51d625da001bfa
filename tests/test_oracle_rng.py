"""Pins the oracle's lcg64 restatement with algebra computed independently in
Python integers (TRNG 4.25 lcg64: r <- a r + b mod 2^64, leapfrog split)."""
import pytest

M64 = 1 << 64
A = 18145460002477866997
B = 1
S = (1 << 32) - 1


def py_output(seed, t):
    """x_t = a^(t+1) seed + b (a^(t+1) - 1)/(a - 1) mod 2^64, exact."""
    n = t + 1
    an = pow(A, n, M64 * (A - 1))
    geo = (an - 1) // (A - 1)
    return (pow(A, n, M64) * seed + B * geo) % M64


def test_master_output_matches_integer_closed_form(orc):
    for seed in (0, 1, 12345, M64 - 1):
        for t in (0, 1, 2, 7, 1000, 123456789, (1 << 40) + 3, M64 - 2):
            assert orc.master_output(seed, t) == py_output(seed, t)


def test_stepping_equals_jump(orc):
    g = orc.lcg64(99)
    xs = [orc.lcg_next(g) for _ in range(50)]
    for t, x in enumerate(xs):
        assert x == py_output(99, t)
    g2 = orc.lcg64(99)
    orc.lcg_jump(g2, 17)
    assert orc.lcg_next(g2) == xs[17]


def test_jump_is_additive(orc):
    for a, b in ((3, 5), (1000, 1), (1 << 33, 12345)):
        g1 = orc.lcg64(7)
        orc.lcg_jump(g1, a)
        orc.lcg_jump(g1, b)
        g2 = orc.lcg64(7)
        orc.lcg_jump(g2, a + b)
        assert orc.lcg_next(g1) == orc.lcg_next(g2)


@pytest.mark.parametrize("s,i", [(2, 0), (2, 1), (3, 2), (7, 4), (S, 0), (S, 12345)])
def test_split_is_leapfrog(orc, s, i):
    g = orc.lcg64(2024)
    assert orc.lcg_split(g, s, i) == 0
    for j in range(5):
        assert orc.lcg_next(g) == py_output(2024, i + j * s)


def test_split_rejects_bad_arguments(orc):
    g = orc.lcg64(0)
    assert orc.lcg_split(g, 0, 0) != 0
    assert orc.lcg_split(g, 4, 4) != 0
    assert orc.lcg_split(g, 1, 0) == 0  # identity split


def test_sample_stream_is_split_of_split(orc):
    """Element j of sample s = sampling.split(S, s) element j where
    sampling = master.split(2, 1) -- built with the split API."""
    for seed in (0, 5):
        for s in (0, 1, 63, 64, 100000):
            g = orc.lcg64(seed)
            orc.lcg_split(g, 2, 1)
            orc.lcg_split(g, S, s)
            for j in range(4):
                x = orc.lcg_next(g)
                assert x == orc.sample_element(seed, s, j)
                assert x == py_output(seed, 1 + 2 * (s + j * S))


def test_far_sample_elements(orc):
    # element index j*S wraps beyond 2^64 in the exponent; M^(2^64) = identity
    for (s, j) in ((3, 1 << 31), (12345, (1 << 32) + 7)):
        assert orc.sample_element(11, s, j) == py_output(11, (1 + 2 * (s + j * S)) % M64)


def test_u01_range_and_quantisation(orc):
    assert orc.u01(0) == 2.0 ** -24
    assert orc.u01(M64 - 1) == 1.0
    assert orc.u01(1 << 40) == 2.0 ** -23
    prev = 0.0
    for k in range(0, 1 << 24, 1 << 18):
        u = orc.u01(k << 40)
        assert u == (k + 1) * 2.0 ** -24
        assert u > prev
        prev = u


def test_uniform_index(orc):
    assert orc.uniform_index(0, 10) == 0
    assert orc.uniform_index(M64 - 1, 10) == 9
    for n in (1, 2, 34, 65536, 4194304):
        for x in (0, 1 << 63, M64 - 1, 0x9E3779B97F4A7C15):
            assert orc.uniform_index(x, n) == ((x >> 32) * n) >> 32
            assert orc.uniform_index(x, n) < n
