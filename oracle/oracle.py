"""ctypes binding of the CPU restatement in msj_oracle.c.

TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  Loaded by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; the product
library never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsj_oracle.so")

SEQ_NVVM7, SEQ_ALLFMA, SEQ_NOFMA = 0, 1, 2
MAXL = 16


class Params(C.Structure):
    _fields_ = [
        ("maxrp", C.c_uint32),
        ("rp_per_layer", C.c_uint32),
        ("numidim", C.c_uint32),
        ("seed", C.c_uint32),
        ("sample_per", C.c_double),
        ("min_calcs_per_node", C.c_uint64),
        ("seq", C.c_uint32),
        ("join", C.c_uint32),
        ("sample_every", C.c_uint32),
        ("nthreads", C.c_int32),
        ("index", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("d", C.c_uint32),
        ("num_nodes", C.c_uint32),
        ("num_layers", C.c_uint32),
        ("num_candidates", C.c_uint32),
        ("pad0", C.c_uint32),
        ("num_calcs", C.c_uint64),
        ("evals_done", C.c_uint64),
        ("total_pairs", C.c_uint64),
        ("counts", C.POINTER(C.c_uint32)),
        ("offsets", C.POINTER(C.c_uint64)),
        ("nbrs", C.POINTER(C.c_uint32)),
        ("node_order", C.POINTER(C.c_uint32)),
        ("node_start", C.POINTER(C.c_uint32)),
        ("adj_off", C.POINTER(C.c_uint32)),
        ("adj", C.POINTER(C.c_uint32)),
        ("node_keys", C.POINTER(C.c_uint32)),
        ("chosen", C.c_uint32 * MAXL),
        ("sse", C.c_double * MAXL),
        ("nodes_per_layer", C.c_uint32 * MAXL),
        ("calcs_per_layer", C.c_uint64 * MAXL),
        ("flips_allfma", C.c_uint64),
        ("flips_nofma", C.c_uint64),
        ("t_construct", C.c_double),
        ("t_join", C.c_double),
        ("t_output", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        L.or_default_params.argtypes = [C.POINTER(Params)]
        L.or_glibc_rand.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_int32)]
        L.or_dim_order.argtypes = [dp, C.c_uint32, C.c_uint32, up, dp]
        L.or_rp_layers.argtypes = [dp, C.c_uint32, C.c_uint32, C.POINTER(Params), C.c_uint32, dp]
        L.or_keys.argtypes = [dp, C.c_uint32, C.c_uint32, C.c_double, dp, C.c_uint32, C.c_uint32, up]
        L.or_seq_sum.argtypes = [dp, dp, C.c_uint32, C.c_uint32]
        L.or_seq_sum.restype = C.c_double
        L.or_self_join.argtypes = [dp, C.c_uint32, C.c_uint32, C.c_double, C.POINTER(Params),
                                   C.POINTER(Result)]
        L.or_brute_force.argtypes = [dp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32,
                                     C.c_int32, C.POINTER(Result)]
        L.or_result_free.argtypes = [C.POINTER(Result)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def default_params(**kw):
    p = Params()
    lib().or_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def glibc_rand(seed, count):
    out = np.zeros(count, dtype=np.int32)
    lib().or_glibc_rand(seed, count, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def dim_order(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n, d = A.shape
    order = np.zeros(d, dtype=np.uint32)
    sigma = np.zeros(d, dtype=np.float64)
    lib().or_dim_order(_dp(A), n, d, _up(order), _dp(sigma))
    return order, sigma


def rp_layers(data, layers=6, **kw):
    data = np.ascontiguousarray(data, dtype=np.float64)
    n, d = data.shape
    p = default_params(**kw)
    out = np.zeros((layers, p.rp_per_layer, d), dtype=np.float64)
    lib().or_rp_layers(_dp(data), n, d, C.byref(p), layers, _dp(out))
    return out


def keys(data, eps, rp, naxis):
    data = np.ascontiguousarray(data, dtype=np.float64)
    rp = np.ascontiguousarray(rp, dtype=np.float64)
    n, d = data.shape
    nrp = rp.shape[0]
    out = np.zeros((nrp + naxis, n), dtype=np.uint32)
    lib().or_keys(_dp(data), n, d, float(eps), _dp(rp), nrp, naxis, _up(out))
    return out


def seq_sum(p, q, seq=SEQ_NVVM7):
    p = np.ascontiguousarray(p, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    return lib().or_seq_sum(_dp(p), _dp(q), p.shape[0], seq)


class JoinResult:
    """Python copy of an or_result (the C buffers are freed on construction)."""

    def __init__(self, r):
        n = r.n
        self.n = n
        self.num_nodes = r.num_nodes
        self.num_layers = r.num_layers
        self.num_candidates = r.num_candidates
        self.num_calcs = r.num_calcs
        self.evals_done = r.evals_done
        self.total_pairs = r.total_pairs
        R = r.total_pairs

        def arr(ptr, cnt, dt):
            if not ptr or cnt == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(cnt,)).copy()

        self.counts = arr(r.counts, n, np.uint32)
        self.offsets = arr(r.offsets, n + 1, np.uint64)
        self.nbrs = arr(r.nbrs, R, np.uint32)
        self.node_order = arr(r.node_order, n, np.uint32)
        self.node_start = arr(r.node_start, r.num_nodes + 1 if r.node_start else 0, np.uint32)
        self.adj_off = arr(r.adj_off, r.num_nodes + 1 if r.adj_off else 0, np.uint32)
        na = int(self.adj_off[-1]) if len(self.adj_off) else 0
        self.adj = arr(r.adj, na, np.uint32)
        L = r.num_layers
        self.node_keys = arr(r.node_keys, r.num_nodes * L if r.node_keys else 0,
                             np.uint32).reshape(-1, L) if L else np.zeros((0, 0), np.uint32)
        self.chosen = list(r.chosen)[:L]
        self.sse = list(r.sse)[:L]
        self.nodes_per_layer = list(r.nodes_per_layer)[:L]
        self.calcs_per_layer = list(r.calcs_per_layer)[:L]
        self.flips_allfma = r.flips_allfma
        self.flips_nofma = r.flips_nofma
        self.t_construct = r.t_construct
        self.t_join = r.t_join
        self.t_output = r.t_output

    def pairs(self):
        """(row, neighbour) pairs as an (R, 2) uint32 array, lexicographically sorted."""
        rows = np.repeat(np.arange(self.n, dtype=np.uint32), self.counts.astype(np.int64))
        return np.stack([rows, self.nbrs], axis=1)


def self_join(data, eps, **kw):
    data = np.ascontiguousarray(data, dtype=np.float64)
    n, d = data.shape
    p = default_params(**kw)
    r = Result()
    rc = lib().or_self_join(_dp(data), n, d, float(eps), C.byref(p), C.byref(r))
    if rc != 0:
        raise ValueError("or_self_join rejected its arguments (rc=%d)" % rc)
    try:
        return JoinResult(r)
    finally:
        lib().or_result_free(C.byref(r))


def brute_force(data, eps, seq=SEQ_NVVM7, nthreads=0):
    data = np.ascontiguousarray(data, dtype=np.float64)
    n, d = data.shape
    r = Result()
    rc = lib().or_brute_force(_dp(data), n, d, float(eps), seq, nthreads, C.byref(r))
    if rc != 0:
        raise ValueError("or_brute_force rejected its arguments (rc=%d)" % rc)
    try:
        return JoinResult(r)
    finally:
        lib().or_result_free(C.byref(r))
