"""Pins the oracle's OPIM-C (opimc.h:26-241): the theta and bound closed forms
against the paper's equations evaluated in double precision, FindCoverage
against brute force, the R2 sample-id mapping against a direct jump, and the
round structure against sampling + selection + coverage run step by step."""
import math

import numpy as np
import pytest

import ripples_amd as R
from graphs import karate_csr, lt_normalise, random_graph, sets_of, transposed_csr
from oracle import oracle as O

A = 1 - 1 / math.e


def log_binomial(n, k):
    return math.lgamma(n + 1) - math.lgamma(n - k + 1) - math.lgamma(k + 1)


def helper64(n, k, delta):
    # Tang et al., OPIM-C: 2 * ((1-1/e) sqrt(ln 6/d) + sqrt((1-1/e)(ln C(n,k) + ln 6/d)))^2
    a = math.log(6 / delta)
    return 2 * (A * math.sqrt(a) + math.sqrt(A * (log_binomial(n, k) + a))) ** 2


@pytest.mark.parametrize("n,k,eps,delta", [(34, 3, 0.3, 0.05), (1000, 10, 0.1, 0.05),
                                           (1 << 20, 100, 0.13, 0.01), (5000, 50, 0.5, 0.2)])
def test_theta_zero_and_max(n, k, eps, delta):
    h = helper64(n, k, np.float32(delta))
    assert abs(O.opimc_theta_helper(n, eps, k, delta) - h) <= 1e-5 * h
    t0 = O.opimc_theta_zero(n, eps, k, delta)
    assert abs(t0 - math.ceil(h)) <= 1
    tmax = O.opimc_theta_max(n, eps, k, delta)
    want = n * h / (np.float32(eps) ** 2 * k)
    assert abs(tmax - want) <= 1e-5 * want + 1


@pytest.mark.parametrize("cov,delta,n,theta", [(0, 0.01, 100, 50), (7, 0.003, 1000, 300),
                                               (5000, 0.001, 1 << 20, 20000),
                                               (123456, 1e-4, 1 << 22, 200000)])
def test_bounds_equations(cov, delta, n, theta):
    d = float(np.float32(delta))
    a = math.log(1 / d)
    lower = ((math.sqrt(cov + 2 * a / 9) - math.sqrt(a / 2)) ** 2 - a / 18) * n / theta
    upper = (math.sqrt(cov / A + a / 2) + math.sqrt(a / 2)) ** 2 * n / theta
    lo, up = O.opimc_lower_bound(cov, delta, n, theta), O.opimc_upper_bound(cov, delta, n, theta)
    assert abs(lo - lower) <= 2e-5 * abs(upper) + 1e-6
    assert abs(up - upper) <= 2e-5 * upper
    # the estimate lies between the bounds; the gap shrinks as coverage grows
    assert lo <= cov * n / theta <= up / A + 1e-3


def test_coverage_brute_force():
    rng = np.random.default_rng(4)
    n = 50
    sets = [rng.choice(n, size=rng.integers(0, 8), replace=False) for _ in range(300)]
    off = np.concatenate([[0], np.cumsum([len(s) for s in sets])]).astype(np.uint64)
    ver = np.concatenate(sets).astype(np.uint32)
    for seeds in ([], [3], [3, 17, 49], list(range(n))):
        want = sum(1 for s in sets if set(s.tolist()) & set(seeds))
        assert O.coverage(n, off, ver, seeds) == want


def test_r2_seed_draws_ids_from_two_to_the_31():
    index, src, w = karate_csr(0.3)
    for seed in (0, 9):
        r2 = R.opimc_r2_seed(seed)
        a = O.sample(R.IC, index, src, w, r2, 0, 40)
        b = O.sample(R.IC, index, src, w, seed, 1 << 31, 40)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert O.sample_element(r2, 5, 3) == O.sample_element(seed, (1 << 31) + 5, 3)


def replay(model, index, src, w, k, eps, delta, seed):
    """OPIMC's loop from sampling, selection and coverage, one step at a time."""
    n = len(index) - 1
    theta = O.opimc_theta_zero(n, eps, k, delta)
    tmax = O.opimc_theta_max(n, eps, k, delta)
    dlt = np.float32(delta) / np.float32(3 * math.ceil(math.log2(tmax // theta)))
    rounds = []
    while True:
        off1, ver1 = O.sample(model, index, src, w, seed, 0, theta)
        off2, ver2 = O.sample(model, index, src, w, seed, 1 << 31, theta)
        seeds, cov = O.select(n, off1, ver1, k)
        c1 = int(cov / theta * theta)
        r2 = [set(s) for s in sets_of(off2, ver2)]
        c2 = sum(1 for s in r2 if s & set(seeds.tolist()))
        alpha = np.float32(O.opimc_lower_bound(c2, dlt, n, theta)) / np.float32(
            O.opimc_upper_bound(c1, dlt, n, theta))
        rounds.append((theta, c1, c2))
        if float(alpha) >= float(np.float32(A)) - eps:
            return seeds, rounds
        theta *= 2


@pytest.mark.parametrize("model", [R.IC, R.LT])
def test_opimc_round_structure(model):
    n, s, d, w = random_graph(120, 700, 3, 0.0, 0.3)
    if model == R.LT:
        w = lt_normalise(n, s, d, w, 0.9)
    index, src, wt = transposed_csr(n, s, d, w)
    seeds, rec = O.opimc(model, index, src, wt, 5, 0.3, 0.1, 2)
    want_seeds, rounds = replay(model, index, src, wt, 5, 0.3, 0.1, 2)
    assert seeds.tolist() == want_seeds.tolist()
    assert rec.rounds == len(rounds)
    for i, (theta, c1, c2) in enumerate(rounds):
        assert (rec.coverage1[i], rec.coverage2[i]) == (c1, c2)
        assert rec.rrr_sets_generated[i] == 2 * (theta if i == 0 else theta // 2)
    assert rec.theta == rounds[-1][0] and rec.theta_zero == rounds[0][0]
    assert rec.alpha[rec.rounds - 1] >= np.float32(A) - 0.3
    assert all(rec.alpha[i] < np.float32(A) - 0.3 for i in range(rec.rounds - 1))
