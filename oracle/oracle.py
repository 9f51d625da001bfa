"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  The product never calls it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

IC, LT = 0, 1
ALGO_BFS, ALGO_BATCHED64, ALGO_DEFINITION = 0, 1, 2

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class ImmRecord(C.Structure):
    _fields_ = [
        ("theta", C.c_uint64),
        ("num_sets", C.c_uint64),
        ("rounds", C.c_uint32),
        ("theta_prime", C.c_int64 * 64),
        ("f", C.c_double * 64),
        ("lb", C.c_double),
        ("final_f", C.c_double),
    ]


class Lcg64(C.Structure):
    _fields_ = [("a", C.c_uint64), ("b", C.c_uint64), ("r", C.c_uint64)]


class OpimcRecord(C.Structure):
    _fields_ = [("theta_zero", C.c_uint64), ("theta_max", C.c_uint64), ("theta", C.c_uint64),
                ("rounds", C.c_uint32), ("num_seeds", C.c_uint32),
                ("rrr_sets_generated", C.c_uint64 * 64), ("coverage1", C.c_uint64 * 64),
                ("coverage2", C.c_uint64 * 64), ("lower", C.c_float * 64),
                ("upper", C.c_float * 64), ("alpha", C.c_float * 64)]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run `make -C oracle`")
    L = C.CDLL(path)
    L.orc_lcg64_seed.argtypes = [C.POINTER(Lcg64), C.c_uint64]
    L.orc_lcg64_next.argtypes = [C.POINTER(Lcg64)]
    L.orc_lcg64_next.restype = C.c_uint64
    L.orc_lcg64_jump.argtypes = [C.POINTER(Lcg64), C.c_uint64]
    L.orc_lcg64_split.argtypes = [C.POINTER(Lcg64), C.c_uint64, C.c_uint64]
    L.orc_lcg64_split.restype = C.c_int
    L.orc_master_output.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_master_output.restype = C.c_uint64
    L.orc_u01.argtypes = [C.c_uint64]
    L.orc_u01.restype = C.c_float
    L.orc_uniform_index.argtypes = [C.c_uint64, C.c_uint32]
    L.orc_uniform_index.restype = C.c_uint32
    L.orc_sample_element.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.orc_sample_element.restype = C.c_uint64
    L.orc_sample_root.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
    L.orc_sample_root.restype = C.c_uint32
    L.orc_ic_neg_log2.argtypes = [C.c_uint64]
    L.orc_ic_neg_log2.restype = C.c_uint64
    L.orc_ic_threshold.argtypes = [C.c_float]
    L.orc_ic_threshold.restype = C.c_uint32
    L.orc_ic_gap_q.argtypes = [C.c_uint32]
    L.orc_ic_gap_q.restype = C.c_uint64
    L.orc_ic_block_log2.argtypes = [C.c_uint32]
    L.orc_ic_block_log2.restype = C.c_uint32
    L.orc_ic_gap.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_ic_gap.restype = C.c_uint64
    L.orc_sample.argtypes = [C.c_int, C.c_int, C.c_uint32, _u64p, _u32p, _f32p,
                             C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
    L.orc_sample.restype = C.c_void_p
    L.orc_sets_count.argtypes = [C.c_void_p]
    L.orc_sets_count.restype = C.c_uint64
    L.orc_sets_total.argtypes = [C.c_void_p]
    L.orc_sets_total.restype = C.c_uint64
    L.orc_sets_export.argtypes = [C.c_void_p, _u64p, _u32p]
    L.orc_sets_free.argtypes = [C.c_void_p]
    L.orc_select.argtypes = [C.c_uint32, C.c_uint64, _u64p, _u32p, C.c_uint32,
                             _u32p, C.POINTER(C.c_uint64)]
    L.orc_select.restype = C.c_uint32
    L.orc_count.argtypes = [C.c_uint32, C.c_uint64, _u64p, _u32p, _u32p]
    L.orc_log_binomial.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_log_binomial.restype = C.c_double
    L.orc_theta_prime.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_uint64,
                                  C.c_uint64]
    L.orc_theta_prime.restype = C.c_int64
    L.orc_theta.argtypes = [C.c_double, C.c_double, C.c_uint64, C.c_double,
                            C.c_uint64]
    L.orc_theta.restype = C.c_uint64
    L.orc_imm.argtypes = [C.c_int, C.c_uint32, _u64p, _u32p, _f32p, C.c_uint32,
                          C.c_double, C.c_uint64, C.c_int, _u32p,
                          C.POINTER(ImmRecord)]
    L.orc_imm.restype = C.c_uint32
    L.orc_spread.argtypes = [C.c_int, C.c_uint32, _u64p, _u32p, _f32p, _u32p, C.c_uint32,
                             C.c_uint64, C.c_uint64, C.c_uint64, _u64p, _u32p, C.c_int]
    L.orc_coverage.argtypes = [C.c_uint32, C.c_uint64, _u64p, _u32p, _u32p, C.c_uint32]
    L.orc_coverage.restype = C.c_uint64
    L.orc_opimc_theta_helper.argtypes = [C.c_uint64, C.c_float, C.c_uint64, C.c_float]
    L.orc_opimc_theta_helper.restype = C.c_float
    for f in ("orc_opimc_theta_max", "orc_opimc_theta_zero"):
        getattr(L, f).argtypes = [C.c_uint64, C.c_float, C.c_uint64, C.c_float]
        getattr(L, f).restype = C.c_uint64
    for f in ("orc_opimc_lower_bound", "orc_opimc_upper_bound"):
        getattr(L, f).argtypes = [C.c_uint64, C.c_float, C.c_uint64, C.c_uint64]
        getattr(L, f).restype = C.c_float
    L.orc_opimc.argtypes = [C.c_int, C.c_uint32, _u64p, _u32p, _f32p, C.c_uint32, C.c_double,
                            C.c_double, C.c_uint64, C.c_int, _u32p, C.POINTER(OpimcRecord)]
    L.orc_opimc.restype = C.c_uint32
    _LIB = L
    return L


def _csr(index, src, w):
    return (np.ascontiguousarray(index, dtype=np.uint64),
            np.ascontiguousarray(src, dtype=np.uint32),
            np.ascontiguousarray(w, dtype=np.float32))


def sample(model, index, src, w, seed, first, count, algo=ALGO_BFS, threads=None):
    """RRR sets of samples [first, first+count): (offsets u64[count+1], vertices u32)."""
    index, src, w = _csr(index, src, w)
    n = len(index) - 1
    threads = threads or min(8, os.cpu_count() or 1)
    h = lib().orc_sample(model, algo, n, index, src, w, seed, first, count, threads)
    if not h:
        raise ValueError("oracle sample: bad arguments")
    try:
        cnt = lib().orc_sets_count(h)
        tot = lib().orc_sets_total(h)
        off = np.zeros(cnt + 1, np.uint64)
        ver = np.zeros(max(tot, 1), np.uint32)
        lib().orc_sets_export(h, off, ver)
    finally:
        lib().orc_sets_free(h)
    return off, ver[:tot]


def select(n, offsets, vertices, k):
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    vertices = np.ascontiguousarray(vertices, dtype=np.uint32)
    if len(vertices) == 0:
        vertices = np.zeros(1, np.uint32)
    seeds = np.zeros(max(k, 1), np.uint32)
    cov = C.c_uint64(0)
    got = lib().orc_select(n, len(offsets) - 1, offsets, vertices, k, seeds, C.byref(cov))
    return seeds[:got], cov.value


def count(n, offsets, vertices):
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    vertices = np.ascontiguousarray(vertices, dtype=np.uint32)
    if len(vertices) == 0:
        vertices = np.zeros(1, np.uint32)
    out = np.zeros(n, np.uint32)
    lib().orc_count(n, len(offsets) - 1, offsets, vertices, out)
    return out


def spread(model, index, src, w, seeds, seed, first, count, threads=None):
    """Forward Monte Carlo spread of `seeds`, replicas [first, first+count):
    (activated u64[count], levels u32[count])."""
    index, src, w = _csr(index, src, w)
    n = len(index) - 1
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    if len(src) == 0:
        src, w = np.zeros(1, np.uint32), np.zeros(1, np.float32)
    act = np.zeros(max(count, 1), np.uint64)
    lev = np.zeros(max(count, 1), np.uint32)
    threads = threads or min(8, os.cpu_count() or 1)
    lib().orc_spread(model, n, index, src, w, seeds if len(seeds) else np.zeros(1, np.uint32),
                     len(seeds), seed, first, count, act, lev, threads)
    return act[:count], lev[:count]


def imm(model, index, src, w, k, eps, seed, threads=None):
    index, src, w = _csr(index, src, w)
    n = len(index) - 1
    threads = threads or min(8, os.cpu_count() or 1)
    seeds = np.zeros(max(k, 1), np.uint32)
    rec = ImmRecord()
    got = lib().orc_imm(model, n, index, src, w, k, eps, seed, threads, seeds, C.byref(rec))
    return seeds[:got], rec


def coverage(n, offsets, vertices, seeds):
    """Sets of the pool containing any of `seeds` (opimc.h:100-118)."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    vertices = np.ascontiguousarray(vertices, np.uint32)
    seeds = np.ascontiguousarray(seeds, np.uint32)
    return int(lib().orc_coverage(n, len(offsets) - 1, offsets,
                                  vertices if len(vertices) else np.zeros(1, np.uint32),
                                  seeds if len(seeds) else np.zeros(1, np.uint32), len(seeds)))


def opimc_theta_helper(n, eps, k, delta):
    return lib().orc_opimc_theta_helper(n, eps, k, delta)


def opimc_theta_max(n, eps, k, delta):
    return int(lib().orc_opimc_theta_max(n, eps, k, delta))


def opimc_theta_zero(n, eps, k, delta):
    return int(lib().orc_opimc_theta_zero(n, eps, k, delta))


def opimc_lower_bound(cov, delta, n, theta):
    return lib().orc_opimc_lower_bound(cov, delta, n, theta)


def opimc_upper_bound(cov, delta, n, theta):
    return lib().orc_opimc_upper_bound(cov, delta, n, theta)


def opimc(model, index, src, w, k, eps, delta, seed, threads=None):
    index, src, w = _csr(index, src, w)
    n = len(index) - 1
    threads = threads or min(8, os.cpu_count() or 1)
    seeds = np.zeros(max(k, 1), np.uint32)
    rec = OpimcRecord()
    got = lib().orc_opimc(model, n, index, src, w, k, eps, delta, seed, threads, seeds,
                          C.byref(rec))
    return seeds[:got], rec


def lcg64(seed):
    g = Lcg64()
    lib().orc_lcg64_seed(C.byref(g), seed)
    return g


def lcg_next(g):
    return lib().orc_lcg64_next(C.byref(g))


def lcg_jump(g, steps):
    lib().orc_lcg64_jump(C.byref(g), steps)


def lcg_split(g, s, i):
    return lib().orc_lcg64_split(C.byref(g), s, i)


def master_output(seed, t):
    return lib().orc_master_output(seed, t)


def u01(x):
    return lib().orc_u01(x)


def uniform_index(x, n):
    return lib().orc_uniform_index(x, n)


def sample_element(seed, s, j):
    return lib().orc_sample_element(seed, s, j)


def ic_neg_log2(u):
    return lib().orc_ic_neg_log2(u)


def ic_threshold(w):
    return lib().orc_ic_threshold(w)


def ic_gap_q(T):
    return lib().orc_ic_gap_q(T)


def ic_block_log2(T):
    return lib().orc_ic_block_log2(T)


def ic_gap(x, q):
    return lib().orc_ic_gap(x, q)


def sample_root(seed, s, n):
    return lib().orc_sample_root(seed, s, n)


def log_binomial(n, k):
    return lib().orc_log_binomial(n, k)


def theta_prime(x, eps_prime, l, k, n):
    return lib().orc_theta_prime(x, eps_prime, l, k, n)


def theta(eps, l, k, lb, n):
    return lib().orc_theta(eps, l, k, lb, n)
