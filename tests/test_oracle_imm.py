"""Pins the oracle's IMM closed forms (imm.h:77-128, utility.h:84-87) and the
driver's round structure (imm.h:145-260, 768-860)."""
import math

import numpy as np
import pytest

from graphs import karate_csr, random_graph, transposed_csr


@pytest.mark.parametrize("n,k", [(10, 3), (34, 5), (100, 50), (200, 1), (60, 0)])
def test_log_binomial_matches_exact(orc, n, k):
    assert orc.log_binomial(n, k) == pytest.approx(math.log(math.comb(n, k)), rel=1e-12, abs=1e-9)


def py_theta_prime(x, epsp, l, k, n):
    k = min(k, n // 2)
    lb = math.lgamma(n + 1) - math.lgamma(n - k + 1) - math.lgamma(k + 1)
    return int((2 + 2.0 / 3.0 * epsp) * (l * math.log(n) + lb + math.log(math.log2(n)))
               * 2.0 ** x / (epsp * epsp))


def py_theta(eps, l, k, LB, n):
    if LB == 0:
        return 0
    k = min(k, n // 2)
    lb = math.lgamma(n + 1) - math.lgamma(n - k + 1) - math.lgamma(k + 1)
    e1 = 1 - 1 / math.e
    alpha = math.sqrt(l * math.log(n) + math.log(2))
    beta = math.sqrt(e1 * (lb + l * math.log(n) + math.log(2)))
    return int(2 * n * (e1 * alpha + beta) ** 2 * eps ** -2 / LB)


@pytest.mark.parametrize("n,k,eps", [(35, 5, 0.5), (65536, 50, 0.5), (1 << 20, 100, 0.13),
                                     (1 << 22, 200, 0.1), (7, 10, 0.3)])
def test_theta_formulas(orc, n, k, eps):
    l = 1 + 1 / math.log2(n)
    epsp = math.sqrt(2) * eps
    for x in range(1, 6):
        got = orc.theta_prime(x, epsp, l, k, n)
        assert abs(got - py_theta_prime(x, epsp, l, k, n)) <= 1
        # doubling in x
        nxt = orc.theta_prime(x + 1, epsp, l, k, n)
        assert abs(nxt - 2 * got) <= 2
    for LB in (1.0, 17.5, n / 3):
        got = orc.theta(eps, l, k, LB, n)
        exp = py_theta(eps, l, k, LB, n)
        assert abs(got - exp) <= max(1, exp * 1e-10)
    assert orc.theta(eps, l, k, 0.0, n) == 0


def test_theta_matches_paper_constants(orc):
    # IMM (Tang et al. 2015) lambda* for n=1000, k=10, eps=0.5, l=1, LB=n
    n, k, eps = 1000, 10, 0.5
    logc = math.log(math.comb(n, k))
    alpha = math.sqrt(math.log(n) + math.log(2))
    beta = math.sqrt((1 - 1 / math.e) * (logc + math.log(n) + math.log(2)))
    lam = 2 * n * ((1 - 1 / math.e) * alpha + beta) ** 2 / eps ** 2
    assert orc.theta(eps, 1.0, k, float(n), n) == pytest.approx(lam / n, abs=1)


@pytest.mark.parametrize("model", [0, 1])
def test_imm_driver_round_structure(orc, model):
    if model == 0:
        index, src, w = karate_csr(0.3)
    else:
        n, s, d, w = random_graph(400, 3000, 5)
        from graphs import lt_normalise
        w = lt_normalise(n, s, d, w)
        index, src, w = transposed_csr(n, s, d, w)
    n = len(index) - 1
    k, eps, seed = 4, 0.5, 3
    seeds, rec = orc.imm(model, index, src, w, k, eps, seed)
    l = 1 + 1 / math.log2(n)
    epsp = math.sqrt(2) * eps
    assert rec.rounds >= 1
    for r in range(rec.rounds):
        x = r + 1
        assert rec.theta_prime[r] == orc.theta_prime(x, epsp, l, k, n)
        # f of the round = coverage of the greedy seeds on all sets so far
        off, ver = orc.sample(model, index, src, w, seed, 0, rec.theta_prime[r])
        _, cov = orc.select(n, off, ver, k)
        assert rec.f[r] == cov / rec.theta_prime[r]
        stop = rec.f[r] >= 2.0 ** -x
        assert stop == (r == rec.rounds - 1) or (r == rec.rounds - 1 and x + 1 >= math.log2(n))
    LB = n * rec.f[rec.rounds - 1] / (1 + epsp)
    assert rec.lb == pytest.approx(LB, rel=1e-15)
    assert rec.theta == orc.theta(eps, l, k, rec.lb, n)
    assert rec.num_sets == max(rec.theta, rec.theta_prime[rec.rounds - 1])
    off, ver = orc.sample(model, index, src, w, seed, 0, rec.theta)
    exp, cov = orc.select(n, off, ver, k)
    assert seeds.tolist() == exp.tolist()
    assert rec.final_f == cov / rec.theta
