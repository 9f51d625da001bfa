"""ripples_amd: MI355X-native IMM hot path (ctypes binding of libripples_amd.so).

The compute path is the HIP library behind the C ABI declared in
``include/ripples_amd.h``; this module only marshals numpy arrays across it.
There is no CPU fallback: if the library is missing, importing the binding
raises, and every compute call needs a HIP device.

Names follow the reference's IMM interface (include/ripples/imm.h,
generate_rrr_sets.h, find_most_influential.h).
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "IC", "LT", "Tune", "ImmRecord", "Graph", "Pool", "RipplesError", "lib",
    "device_count", "rmat_edges", "csr_from_edges", "imm",
    "W_UNIFORM", "W_CASCADE", "W_LT_NORM", "W_CONST", "Server", "SERVER_STATS", "opimc", "OpimcRecord",
]

IC, LT = 0, 1
W_UNIFORM, W_CASCADE, W_LT_NORM, W_CONST = 0, 1, 2, 3
DIST_UNIFORM, DIST_NORMAL, DIST_CONST = 0, 1, 2

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libripples_amd.so")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class RipplesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class Tune(C.Structure):
    _fields_ = [("colors", C.c_uint32), ("block", C.c_uint32),
                ("slots", C.c_uint32), ("queue_cap", C.c_uint32),
                ("heavy_degree", C.c_uint32), ("sampler", C.c_uint32),
                ("out_chunk", C.c_uint32), ("frontier_cap", C.c_uint32),
                ("wave_stage", C.c_uint32)]


class LoadOptions(C.Structure):
    _fields_ = [("model", C.c_int), ("weighted", C.c_int), ("undirected", C.c_int),
                ("renumber", C.c_int), ("distribution", C.c_int), ("mean", C.c_float),
                ("variance", C.c_float), ("scale_factor", C.c_float), ("seed", C.c_uint64)]


class ImmRecord(C.Structure):
    _fields_ = [
        ("theta", C.c_uint64), ("num_sets", C.c_uint64), ("rrr_elements", C.c_uint64),
        ("rounds", C.c_uint32), ("num_seeds", C.c_uint32),
        ("lb", C.c_double), ("final_f", C.c_double),
        ("total_ms", C.c_double), ("estimation_ms", C.c_double),
        ("generate_ms", C.c_double), ("select_ms", C.c_double),
        ("theta_prime", C.c_int64 * 64), ("f", C.c_double * 64),
        ("round_generate_ms", C.c_double * 64), ("round_select_ms", C.c_double * 64),
        ("pool_bytes", C.c_uint64),
    ]

    def as_dict(self):
        r = self.rounds
        return {
            "Theta": self.theta, "NumSets": self.num_sets,
            "RRRElements": self.rrr_elements, "Rounds": r,
            "NumSeeds": self.num_seeds, "LB": self.lb, "FinalF": self.final_f,
            "Total": self.total_ms, "ThetaEstimation": self.estimation_ms,
            "GenerateRRRSets": self.generate_ms,
            "FindMostInfluentialSet": self.select_ms,
            "ThetaPrime": list(self.theta_prime[:min(r, 64)]),
            "F": list(self.f[:min(r, 64)]),
            "ThetaEstimationGenerateRRR": list(self.round_generate_ms[:min(r, 64)]),
            "ThetaEstimationMostInfluential": list(self.round_select_ms[:min(r, 64)]),
            "PoolBytes": self.pool_bytes,
        }


class OpimcRecord(C.Structure):
    """rip_opimc_record (OPIMCExecutionRecord plus per-round bounds)."""
    _fields_ = [("theta_zero", C.c_uint64), ("theta_max", C.c_uint64), ("theta", C.c_uint64),
                ("rrr_elements", C.c_uint64), ("rounds", C.c_uint32),
                ("num_seeds", C.c_uint32), ("total_ms", C.c_double),
                ("rrr_sets_generated", C.c_uint64 * 64), ("generate_ms", C.c_double * 64),
                ("select_ms", C.c_double * 64), ("coverage1", C.c_uint64 * 64),
                ("coverage2", C.c_uint64 * 64), ("lower", C.c_float * 64),
                ("upper", C.c_float * 64), ("alpha", C.c_float * 64)]

    def as_dict(self):
        r = min(self.rounds, 64)
        return {
            "ThetaZero": self.theta_zero, "ThetaMax": self.theta_max, "Theta": self.theta,
            "Rounds": self.rounds, "NumSeeds": self.num_seeds, "Total": self.total_ms,
            "RRRElements": self.rrr_elements,
            "RRRSetsGenerated": list(self.rrr_sets_generated[:r]),
            "GenerateRRRSets": list(self.generate_ms[:r]),
            "FindMostInfluentialSet": list(self.select_ms[:r]),
            "Coverage1": list(self.coverage1[:r]), "Coverage2": list(self.coverage2[:r]),
            "Lower": list(self.lower[:r]), "Upper": list(self.upper[:r]),
            "Alpha": list(self.alpha[:r]),
        }


_LIB = None


def lib():
    """Load libripples_amd.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libripples_amd.so is not built (%s); run "
                           "`python -m ripples_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.rip_last_error.restype = C.c_char_p
    L.rip_version.restype = C.c_int
    L.rip_device_count.argtypes = [C.POINTER(C.c_int)]
    L.rip_graph_create.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint64, _u64p,
                                   _u32p, _f32p, C.POINTER(vp)]
    L.rip_graph_destroy.argtypes = [vp]
    L.rip_graph_info.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_int)]
    L.rip_pool_create.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.rip_pool_destroy.argtypes = [vp]
    L.rip_pool_sample.argtypes = [vp, C.c_uint64, C.POINTER(Tune)]
    L.rip_pool_reset.argtypes = [vp]
    L.rip_pool_append_host.argtypes = [vp, C.c_uint64, _u64p, vp]
    L.rip_pool_size.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.rip_pool_export.argtypes = [vp, C.c_uint64, C.c_uint64, _u64p, vp, C.c_uint64]
    L.rip_pool_select.argtypes = [vp, C.c_uint64, C.c_uint32, _u32p,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
    L.rip_pool_count.argtypes = [vp, C.c_uint64, _u32p]
    L.rip_pool_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_double),
                                          C.POINTER(C.c_uint32)]
    L.rip_pool_last_call_ms.argtypes = [vp, C.POINTER(C.c_double)]
    L.rip_autotune.argtypes = [vp, C.c_uint64, C.c_int, C.POINTER(Tune), C.POINTER(C.c_double),
                               C.POINTER(C.c_uint32)]
    L.rip_pool_last_grid.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32), _u64p,
                                     C.c_uint32]
    L.rip_pool_last_wave.argtypes = [vp, C.POINTER(C.c_double), _u64p, C.c_uint32]
    L.rip_pool_device_bytes.argtypes = [vp, _u64p]
    L.rip_pool_columns.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint64)]
    L.rip_pool_last_kernel_stats.argtypes = [vp, _u64p, C.c_uint32]
    L.rip_pool_last_stats.argtypes = [vp, _u64p, C.c_uint32]
    L.rip_imm.argtypes = [vp, C.c_uint32, C.c_double, C.c_uint64, C.POINTER(Tune),
                          _u32p, C.POINTER(ImmRecord)]
    L.rip_rmat_edges.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                 C.c_double, C.c_uint64, C.POINTER(C.c_uint64), vp, vp,
                                 C.c_uint64]
    L.rip_csr_from_edges.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_int,
                                     C.c_float, C.c_uint64, _u64p, _u32p, _f32p]
    L.rip_spread.argtypes = [vp, _u32p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64,
                             _u64p, _u32p]
    hp = C.POINTER(C.c_void_p)
    L.rip_load_edge_list.argtypes = [C.c_char_p, C.POINTER(LoadOptions), hp]
    L.rip_load_binary.argtypes = [C.c_char_p, C.c_float, hp]
    L.rip_dump_binary.argtypes = [vp, C.c_char_p]
    L.rip_host_graph_info.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
    L.rip_host_graph_arrays.argtypes = [vp, C.POINTER(C.POINTER(C.c_uint64)),
                                        C.POINTER(C.POINTER(C.c_uint32)),
                                        C.POINTER(C.POINTER(C.c_float)),
                                        C.POINTER(C.POINTER(C.c_uint32))]
    L.rip_host_graph_destroy.argtypes = [vp]
    L.rip_pool_coverage.argtypes = [vp, C.c_uint64, _u32p, C.c_uint32, _u64p]
    L.rip_pool_coverage_noalloc.argtypes = [vp, C.c_uint64, _u32p, C.c_uint32, _u64p]
    L.rip_opimc.argtypes = [vp, C.c_uint32, C.c_double, C.c_double, C.c_uint64,
                            C.POINTER(Tune), _u32p, C.POINTER(OpimcRecord)]
    L.rip_opimc_r2_seed.argtypes = [C.c_uint64]
    L.rip_opimc_r2_seed.restype = C.c_uint64
    L.rip_server_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(vp)]
    L.rip_server_add_graph.argtypes = [vp, vp, C.POINTER(C.c_uint32)]
    L.rip_server_submit.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64,
                                    C.POINTER(C.c_uint64)]
    L.rip_server_poll.argtypes = [vp, C.c_uint64, C.c_double, C.POINTER(C.c_int), _u32p,
                                  C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(ImmRecord),
                                  C.POINTER(C.c_double)]
    L.rip_server_stats.argtypes = [vp, _u64p, C.c_uint32]
    L.rip_server_destroy.argtypes = [vp]
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise RipplesError(rc, lib().rip_last_error().decode(errors="replace"))


def _tune(tune):
    if tune is None:
        return None
    if isinstance(tune, Tune):
        return C.byref(tune)
    return C.byref(Tune(**tune))


def device_count():
    n = C.c_int(0)
    _check(lib().rip_device_count(C.byref(n)))
    return n.value


def rmat_edges(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=None):
    """Directed R-MAT edge list (deduplicated, no self loops, sorted)."""
    seed = scale if seed is None else seed
    cap = edge_factor << scale
    src = np.empty(max(cap, 1), np.uint32)
    dst = np.empty(max(cap, 1), np.uint32)
    m = C.c_uint64(0)
    _check(lib().rip_rmat_edges(scale, edge_factor, a, b, c, seed, C.byref(m),
                                src.ctypes.data_as(C.c_void_p),
                                dst.ctypes.data_as(C.c_void_p), cap))
    return src[:m.value].copy(), dst[:m.value].copy()


def csr_from_edges(n, src, dst, scheme=W_UNIFORM, scale=1.0, seed=0):
    """Transposed CSR (in-neighbour rows) with weights: (index, edges, weights)."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    m = len(src)
    index = np.zeros(n + 1, np.uint64)
    edges = np.zeros(max(m, 1), np.uint32)
    weights = np.zeros(max(m, 1), np.float32)
    _check(lib().rip_csr_from_edges(n, m, src if m else np.zeros(1, np.uint32),
                                    dst if m else np.zeros(1, np.uint32), scheme,
                                    scale, seed, index, edges, weights))
    return index, edges[:m], weights[:m]


def _host_graph_arrays(h):
    n, m = C.c_uint32(0), C.c_uint64(0)
    _check(lib().rip_host_graph_info(h, C.byref(n), C.byref(m)))
    pi, pe = C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint32)()
    pw, pr = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)()
    _check(lib().rip_host_graph_arrays(h, C.byref(pi), C.byref(pe), C.byref(pw), C.byref(pr)))
    n, m = n.value, m.value
    cp = lambda ptr, k, dt: np.ctypeslib.as_array(ptr, (k,)).astype(dt) if k else np.zeros(0, dt)
    return cp(pi, n + 1, np.uint64), cp(pe, m, np.uint32), cp(pw, m, np.float32), cp(pr, n, np.uint32)


class HostGraph:
    """Host transposed CSR read from a file: index, edges, weights and the
    reverse map (internal -> original ids), as numpy copies."""

    def __init__(self, handle):
        self._h = handle
        self.index, self.edges, self.weights, self.reverse_map = _host_graph_arrays(handle)

    @property
    def num_nodes(self):
        return len(self.index) - 1

    def dump_binary(self, path):
        """Reference binary CSR, backward direction (graph.h:700-771)."""
        _check(lib().rip_dump_binary(self._h, os.fsencode(path)))

    def close(self):
        if self._h:
            lib().rip_host_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_edge_list(path, model=IC, weighted=False, undirected=False, renumber=True,
                   distribution=DIST_UNIFORM, mean=0.5, variance=1.0, scale_factor=1.0,
                   seed=0):
    """Edge-list file -> HostGraph, with the reference loader's weight and
    renumbering semantics (include/ripples/loaders.h:138-418,
    graph.h:506-599)."""
    opts = LoadOptions(model, int(weighted), int(undirected), int(renumber), distribution,
                       mean, variance, scale_factor, seed)
    h = C.c_void_p()
    _check(lib().rip_load_edge_list(os.fsencode(path), C.byref(opts), C.byref(h)))
    return HostGraph(h)


def load_binary(path, scale_factor=1.0):
    """Reference binary CSR (backward direction) -> HostGraph
    (include/ripples/graph.h:854-995)."""
    h = C.c_void_p()
    _check(lib().rip_load_binary(os.fsencode(path), scale_factor, C.byref(h)))
    return HostGraph(h)


class Graph:
    """Device transposed CSR graph (GraphBwd on the GPU)."""

    def __init__(self, index, edges, weights, model=IC, device=0):
        index = np.ascontiguousarray(index, np.uint64)
        edges = np.ascontiguousarray(edges, np.uint32)
        weights = np.ascontiguousarray(weights, np.float32)
        n = len(index) - 1
        m = len(edges)
        h = C.c_void_p()
        _check(lib().rip_graph_create(device, model, n, m, index,
                                      edges if m else np.zeros(1, np.uint32),
                                      weights if m else np.zeros(1, np.float32),
                                      C.byref(h)))
        self._h = h
        self.n, self.m, self.model = n, m, model

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().rip_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spread(self, seeds, replicas, seed=0, first=0):
        """Monte Carlo spread of `seeds` (diffusion_simulation.h:73-178):
        (activated u64[replicas], levels u32[replicas])."""
        seeds = np.ascontiguousarray(seeds, np.uint32)
        act = np.zeros(max(replicas, 1), np.uint64)
        lev = np.zeros(max(replicas, 1), np.uint32)
        _check(lib().rip_spread(self._h, seeds if len(seeds) else np.zeros(1, np.uint32),
                                len(seeds), seed, first, replicas, act, lev))
        return act[:replicas], lev[:replicas]

    def autotune(self, probe_sets=0, force=False):
        """rip_autotune: the fastest sampler setting on a probe batch (0: sized
        to about 20 ms per setting); the graph uses it from then on.  Returns
        (setting dict, ms per probe batch, settings tried)."""
        t = Tune()
        ms, tried = C.c_double(0), C.c_uint32(0)
        _check(lib().rip_autotune(self._h, probe_sets, 1 if force else 0, C.byref(t),
                                  C.byref(ms), C.byref(tried)))
        return {f: getattr(t, f) for f, _ in Tune._fields_ if getattr(t, f)}, ms.value, tried.value

    def pool(self, seed=0):
        return Pool(self, seed)


class Pool:
    """HBM-resident RRR pool of one (graph, seed) sampling stream."""

    def __init__(self, graph, seed=0):
        self.graph = graph
        h = C.c_void_p()
        _check(lib().rip_pool_create(graph.handle, seed, C.byref(h)))
        self._h = h
        self.seed = seed

    def close(self):
        if self._h:
            lib().rip_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate_rrr_sets(self, count, tune=None):
        """Append the RRR sets of the next `count` sample ids."""
        _check(lib().rip_pool_sample(self._h, count, _tune(tune)))

    sample = generate_rrr_sets

    def append(self, offsets, vertices):
        """Append host-side sorted sets given in CSR form."""
        offsets = np.ascontiguousarray(offsets, np.uint64)
        vertices = np.ascontiguousarray(vertices, np.uint32)
        _check(lib().rip_pool_append_host(self._h, len(offsets) - 1, offsets,
                                          vertices.ctypes.data_as(C.c_void_p)))

    def reset(self):
        _check(lib().rip_pool_reset(self._h))

    def size(self):
        s, e = C.c_uint64(0), C.c_uint64(0)
        _check(lib().rip_pool_size(self._h, C.byref(s), C.byref(e)))
        return s.value, e.value

    def export(self, first=0, count=None):
        """Sets [first, first+count) as (offsets u64[count+1], vertices u32)."""
        if count is None:
            count = self.size()[0] - first
        off = np.zeros(count + 1, np.uint64)
        _check(lib().rip_pool_export(self._h, first, count, off, None, 0))
        total = int(off[-1])
        ver = np.zeros(max(total, 1), np.uint32)
        _check(lib().rip_pool_export(self._h, first, count, off,
                                     ver.ctypes.data_as(C.c_void_p), len(ver)))
        return off, ver[:total]

    def find_most_influential_set(self, k, num_sets=None):
        """Greedy seeds over sets [0, num_sets): (seeds, covered sets)."""
        if num_sets is None:
            num_sets = self.size()[0]
        seeds = np.zeros(max(k, 1), np.uint32)
        got, cov = C.c_uint32(0), C.c_uint64(0)
        _check(lib().rip_pool_select(self._h, num_sets, k, seeds, C.byref(got),
                                     C.byref(cov)))
        return seeds[:got.value], cov.value

    select = find_most_influential_set

    def count_occurrences(self, num_sets=None):
        if num_sets is None:
            num_sets = self.size()[0]
        out = np.zeros(self.graph.n, np.uint32)
        _check(lib().rip_pool_count(self._h, num_sets, out))
        return out

    def coverage(self, seeds, num_sets=None):
        """Sets among the first num_sets containing any seed (FindCoverage)."""
        seeds = np.ascontiguousarray(seeds, np.uint32)
        ns = self.size()[0] if num_sets is None else num_sets
        out = np.zeros(1, np.uint64)
        _check(lib().rip_pool_coverage(self._h, ns, seeds if len(seeds) else
                                       np.zeros(1, np.uint32), len(seeds), out))
        return int(out[0])

    def coverage_noalloc(self, seeds, num_sets=None):
        """coverage() through rip_pool_coverage_noalloc: the library's fixed
        device array instead of allocated scratch."""
        seeds = np.ascontiguousarray(seeds, np.uint32)
        ns = self.size()[0] if num_sets is None else num_sets
        out = np.zeros(1, np.uint64)
        _check(lib().rip_pool_coverage_noalloc(self._h, ns, seeds if len(seeds) else
                                               np.zeros(1, np.uint32), len(seeds), out))
        return int(out[0])

    def last_kernel_ms(self):
        ms, launches = C.c_double(0), C.c_uint32(0)
        _check(lib().rip_pool_last_kernel_ms(self._h, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def last_grid(self):
        """(summed ms of the grid sampler's level kernels, levels, its work
        counters) for the last sampling call."""
        ms, lv = C.c_double(0), C.c_uint32(0)
        st = np.zeros(5, np.uint64)
        _check(lib().rip_pool_last_grid(self._h, C.byref(ms), C.byref(lv), st, 5))
        keys = ("algo_bytes", "edges", "coins", "expansions", "levels")
        return ms.value, lv.value, {k: int(v) for k, v in zip(keys, st)}

    def last_wave(self):
        """(ms of the wave stage's kernel, its work counters) for the last
        sampling call; 0 and zeros when the stage did not run."""
        ms = C.c_double(0)
        st = np.zeros(5, np.uint64)
        _check(lib().rip_pool_last_wave(self._h, C.byref(ms), st, 5))
        keys = ("algo_bytes", "edges", "coins", "expansions", "levels")
        return ms.value, {k: int(v) for k, v in zip(keys, st)}

    def last_call_ms(self):
        """Device time of the whole last sampling call (every kernel)."""
        ms = C.c_double(0)
        _check(lib().rip_pool_last_call_ms(self._h, C.byref(ms)))
        return ms.value

    STAT_NAMES = ("algo_bytes", "edges", "coins", "expansions", "levels")

    def last_stats(self):
        """Work counters of the last sampling call's main kernel."""
        out = np.zeros(len(self.STAT_NAMES), np.uint64)
        _check(lib().rip_pool_last_stats(self._h, out, len(out)))
        return {k: int(v) for k, v in zip(self.STAT_NAMES, out)}

    def bytes(self):
        """Device bytes the pool holds (rip_pool_device_bytes)."""
        out = np.zeros(1, np.uint64)
        _check(lib().rip_pool_device_bytes(self._h, out))
        return int(out[0])

    def columns(self):
        """(column groups in use, hot vertices per column, column bytes allocated)."""
        g, h, b = C.c_uint32(0), C.c_uint32(0), C.c_uint64(0)
        _check(lib().rip_pool_columns(self._h, C.byref(g), C.byref(h), C.byref(b)))
        return g.value, h.value, b.value

    def last_kernel_stats(self):
        """Work counters of the last sampling call's main kernel alone."""
        out = np.zeros(len(self.STAT_NAMES), np.uint64)
        _check(lib().rip_pool_last_kernel_stats(self._h, out, len(out)))
        return {k: int(v) for k, v in zip(self.STAT_NAMES, out)}

    def last_algo_bytes(self):
        return self.last_stats()["algo_bytes"]


def imm(graph, k, epsilon, seed=0, tune=None):
    """IMM (imm.h:768-860): returns (seeds, record dict)."""
    seeds = np.zeros(max(k, 1), np.uint32)
    rec = ImmRecord()
    _check(lib().rip_imm(graph.handle, k, epsilon, seed, _tune(tune), seeds,
                         C.byref(rec)))
    return seeds[:rec.num_seeds], rec.as_dict()


SERVER_STATS = ["launches", "sets_drawn", "selections", "pool_bytes", "pools", "reused",
                "evicted", "recycled"]


class Server:
    """IMM queries against resident graphs with continuous batching.  A query
    (graph, k, epsilon, seed) returns the seeds of imm(graph, k, epsilon, seed)."""

    def __init__(self, slots=8, workers=4, budget_bytes=0):
        h = C.c_void_p()
        _check(lib().rip_server_create(slots, workers, budget_bytes, C.byref(h)))
        self._h = h
        self._graphs = []

    def add_graph(self, graph):
        gid = C.c_uint32()
        _check(lib().rip_server_add_graph(self._h, graph.handle, C.byref(gid)))
        self._graphs.append(graph)  # the graph must outlive the server
        return gid.value

    def submit(self, graph_id, k, epsilon, seed=0):
        qid = C.c_uint64()
        _check(lib().rip_server_submit(self._h, graph_id, k, epsilon, seed, C.byref(qid)))
        return qid.value

    def poll(self, query_id, wait_ms=-1.0, cap=None):
        """(seeds, record dict, latency ms) once the query is done, else None.
        wait_ms < 0 waits without limit."""
        cap = cap or 1 << 16
        seeds = np.zeros(cap, np.uint32)
        done, ns, lat = C.c_int(), C.c_uint32(), C.c_double()
        rec = ImmRecord()
        wait = 1e12 if wait_ms < 0 else wait_ms
        _check(lib().rip_server_poll(self._h, query_id, wait, C.byref(done), seeds, cap,
                                     C.byref(ns), C.byref(rec), C.byref(lat)))
        if not done.value:
            return None
        return seeds[:ns.value].copy(), rec.as_dict(), lat.value

    def stats(self):
        out = np.zeros(len(SERVER_STATS), np.uint64)
        _check(lib().rip_server_stats(self._h, out, len(out)))
        return dict(zip(SERVER_STATS, (int(v) for v in out)))

    def close(self):
        if self._h:
            lib().rip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def opimc(graph, k, epsilon, delta=0.05, seed=0, tune=None):
    """OPIM-C (opimc.h:150-241): returns (seeds, record dict)."""
    seeds = np.zeros(max(k, 1), np.uint32)
    rec = OpimcRecord()
    _check(lib().rip_opimc(graph.handle, k, epsilon, delta, seed, _tune(tune), seeds,
                           C.byref(rec)))
    return seeds[:rec.num_seeds], rec.as_dict()


def opimc_r2_seed(seed):
    """Seed whose pool draws sample ids 2^31, 2^31 + 1, ... of `seed`'s stream."""
    return int(lib().rip_opimc_r2_seed(seed))
