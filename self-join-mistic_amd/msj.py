"""ctypes binding of libmsj.so (include/msj.h).

Mirrors the reference's join entry point nodeLauncher5 (reference
src/launcher.cu:3093, include/launcher.cuh:77-82) and its host helper stddev
(src/utils.cu:22-53).  Every call goes to the HIP library; there is no CPU
fallback: if libmsj.so is missing or fails to load, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsj.so")

MAX_LAYERS = 16
OK, E_INVAL, E_NOMEM, E_HIP, E_OVERFLOW, E_UNSUPPORTED, E_NODEV = 0, -1, -2, -3, -4, -5, -6
STORAGE_F64, STORAGE_F32 = 0, 1
FORM_PLAIN, FORM_MFMA_BF16, FORM_MFMA_SPLIT = 0, 1, 2
OUT_HOST, OUT_DEVICE = 0, 1
INDEX_NODENET, INDEX_BRUTE, INDEX_COSS = 0, 1, 2

# Every symbol the C headers under include/ declare for libmsj.so.
EXPORTS = (
    "msj_version", "msj_strerror", "msj_last_error", "msj_default_params",
    "msj_self_join", "msj_result_free", "msj_ctx_create", "msj_ctx_destroy",
    "msj_ctx_load", "msj_ctx_join", "msj_dim_order", "msj_batch_create", "msj_batch_destroy",
    "msj_batch_load", "msj_batch_run", "msj_join_batch",
    "msj_probe_index", "msj_probe_key_math", "msj_probe_repeat_add", "msj_ctx_load_unordered",
    "msj_ctx_set_params",
    "msj_self_join_noalloc",
)


class Params(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("device", C.c_int32),
        ("maxrp", C.c_uint32),
        ("rp_per_layer", C.c_uint32),
        ("numidim", C.c_uint32),
        ("seed", C.c_uint32),
        ("sample_per", C.c_double),
        ("min_calcs_per_node", C.c_uint64),
        ("form", C.c_uint32),
        ("storage", C.c_uint32),
        ("output", C.c_uint32),
        ("verbose", C.c_uint32),
        ("count_work", C.c_uint32),
        ("index", C.c_uint32),
        ("page_pairs", C.c_uint64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_uint32),
        ("num_layers", C.c_uint32),
        ("num_candidates", C.c_uint32),
        ("num_pages", C.c_uint32),
        ("num_calcs", C.c_uint64),
        ("total_pairs", C.c_uint64),
        ("chosen", C.c_uint32 * MAX_LAYERS),
        ("nodes_per_layer", C.c_uint32 * MAX_LAYERS),
        ("calcs_per_layer", C.c_uint64 * MAX_LAYERS),
        ("sse", C.c_double * MAX_LAYERS),
        ("ref_num_batches", C.c_uint32),
        ("pad0", C.c_uint32),
        ("ref_var_calc", C.c_double),
        ("ms_index", C.c_double),
        ("ms_join", C.c_double),
        ("ms_join_kernel", C.c_double),
        ("ms_output", C.c_double),
        ("ms_d2h", C.c_double),
        ("ms_total", C.c_double),
        ("join_launches", C.c_uint64),
        ("dims_evaluated", C.c_uint64),
        ("band_pairs", C.c_uint64),
        ("max_score_err", C.c_double),
        ("mean_score_err", C.c_double),
        ("filter_checked", C.c_uint64),
        ("filter_violations", C.c_uint64),
        ("form_used", C.c_uint32),
        ("split_share_ppm", C.c_uint32),
        ("mfma_flops", C.c_double),
        ("ms_main_kernel", C.c_double),
        ("tile_rows", C.c_uint64),
        ("filter_checked_bf16", C.c_uint64),
    ]


class Job(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("n", C.c_uint32),
        ("d", C.c_uint32),
        ("eps", C.c_double),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("njobs", C.c_uint32),
        ("workers", C.c_uint32),
        ("grid_blocks", C.c_uint32),
        ("classes", C.c_uint32),
        ("class_ds", C.c_uint32 * 8),
        ("class_epi_log2", C.c_uint32 * 8),
        ("work_items", C.c_uint64),
        ("total_pairs", C.c_uint64),
        ("num_calcs", C.c_uint64),
        ("ms_index", C.c_double),
        ("ms_join", C.c_double),
        ("ms_output", C.c_double),
        ("ms_tune", C.c_double),
        ("ms_total", C.c_double),
        ("graph_replays", C.c_uint32),
        ("solo_jobs", C.c_uint32),
        ("class_slab", C.c_uint32 * 8),
    ]


class Result(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("pad0", C.c_uint32),
        ("total_pairs", C.c_uint64),
        ("counts", C.POINTER(C.c_uint32)),
        ("offsets", C.POINTER(C.c_uint64)),
        ("nbrs", C.POINTER(C.c_uint32)),
        ("node_order", C.POINTER(C.c_uint32)),
        ("on_device", C.c_uint32),
        ("pad1", C.c_uint32),
        ("stats", Stats),
    ]


class IndexProbe(C.Structure):
    _fields_ = [
        ("layer", C.c_uint32),
        ("ncand", C.c_uint32),
        ("num_nodes", C.c_uint32),
        ("num_layers", C.c_uint32),
        ("keys", C.POINTER(C.c_uint32)),
        ("node_keys", C.POINTER(C.c_uint32)),
        ("node_start", C.POINTER(C.c_uint32)),
        ("node_order", C.POINTER(C.c_uint32)),
        ("adj_off", C.POINTER(C.c_uint32)),
        ("adj", C.POINTER(C.c_uint32)),
        ("adj_cap", C.c_uint64),
        ("num_edges", C.c_uint64),
        ("num_calcs", C.c_uint64),
        ("chosen", C.c_uint32 * MAX_LAYERS),
        ("nodes_per_layer", C.c_uint32 * MAX_LAYERS),
        ("sse", C.c_double * MAX_LAYERS),
    ]


class MsjError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("msj error %d (%s): %s" % (code, _strerror(code), msg))
        self.code = code


_lib = None


def lib():
    """Loads libmsj.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmsj.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.msj_version.restype = C.c_char_p
        L.msj_strerror.restype = C.c_char_p
        L.msj_strerror.argtypes = [C.c_int]
        L.msj_last_error.restype = C.c_char_p
        L.msj_default_params.argtypes = [C.POINTER(Params)]
        L.msj_self_join.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double,
                                    C.POINTER(Params), C.POINTER(Result)]
        L.msj_result_free.argtypes = [C.POINTER(Result)]
        L.msj_ctx_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.msj_ctx_destroy.argtypes = [C.c_void_p]
        L.msj_ctx_load.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.msj_ctx_join.argtypes = [C.c_void_p, C.c_double, C.POINTER(Result)]
        L.msj_ctx_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
        L.msj_self_join_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double,
                                            C.POINTER(Params), C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_uint64, C.c_void_p,
                                            C.POINTER(Stats)]
        L.msj_ctx_load_unordered.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_uint32,
                                             C.c_uint32, C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_double)]
        L.msj_dim_order.argtypes = [C.POINTER(C.c_double), C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.msj_batch_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.msj_batch_destroy.argtypes = [C.c_void_p]
        L.msj_batch_load.argtypes = [C.c_void_p, C.POINTER(Job), C.c_uint32]
        L.msj_batch_run.argtypes = [C.c_void_p, C.POINTER(Result), C.POINTER(BatchStats)]
        L.msj_join_batch.argtypes = [C.POINTER(Job), C.c_uint32, C.POINTER(Params),
                                     C.POINTER(Result)]
        L.msj_probe_index.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double,
                                      C.POINTER(Params), C.POINTER(IndexProbe)]
        L.msj_probe_repeat_add.argtypes = [C.c_double, C.c_double, C.c_uint64]
        L.msj_probe_repeat_add.restype = C.c_double
        L.msj_probe_key_math.argtypes = [C.POINTER(C.c_double), C.c_uint64, C.c_double,
                                         C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int]
        _lib = L
    return _lib


def _strerror(code):
    try:
        return lib().msj_strerror(code).decode()
    except Exception:  # noqa: BLE001 - message only
        return "?"


def _check(rc):
    if rc != OK:
        raise MsjError(rc, lib().msj_last_error().decode())


def version():
    return lib().msj_version().decode()


def default_params(**kw):
    p = Params()
    lib().msj_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError("msj_params has no field %r" % k)
        setattr(p, k, v)
    return p


def dim_order(data):
    """utils.cu:22-53: dims by decreasing population sigma (ties: higher index first)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    n, d = data.shape
    order = np.zeros(d, dtype=np.uint32)
    sigma = np.zeros(d, dtype=np.float64)
    _check(lib().msj_dim_order(data.ctypes.data_as(C.POINTER(C.c_double)), n, d,
                               order.ctypes.data_as(C.POINTER(C.c_uint32)),
                               sigma.ctypes.data_as(C.POINTER(C.c_double))))
    return order, sigma


class JoinResult:
    """Host copy of an msj_result.  `stats` is a dict of msj_stats."""

    def __init__(self, r, copy_arrays=True):
        self.n = r.n
        self.total_pairs = r.total_pairs
        st = r.stats
        L = st.num_layers
        self.stats = {
            "num_nodes": st.num_nodes, "num_layers": L, "num_candidates": st.num_candidates,
            "num_pages": st.num_pages, "num_calcs": st.num_calcs, "total_pairs": st.total_pairs,
            "chosen": list(st.chosen)[:L], "nodes_per_layer": list(st.nodes_per_layer)[:L],
            "calcs_per_layer": list(st.calcs_per_layer)[:L], "sse": list(st.sse)[:L],
            "ref_num_batches": st.ref_num_batches, "ref_var_calc": st.ref_var_calc,
            "ms_index": st.ms_index, "ms_join": st.ms_join, "ms_join_kernel": st.ms_join_kernel,
            "ms_output": st.ms_output, "ms_d2h": st.ms_d2h, "ms_total": st.ms_total,
            "join_launches": st.join_launches, "dims_evaluated": st.dims_evaluated,
            "band_pairs": st.band_pairs,
            "max_score_err": st.max_score_err, "mean_score_err": st.mean_score_err,
            "filter_checked": st.filter_checked, "filter_violations": st.filter_violations,
            "form_used": st.form_used, "mfma_flops": st.mfma_flops,
            "split_share_ppm": st.split_share_ppm,
            "ms_main_kernel": st.ms_main_kernel, "tile_rows": st.tile_rows,
            "filter_checked_bf16": st.filter_checked_bf16,
        }
        self.on_device = bool(r.on_device)
        if copy_arrays and not r.on_device:
            n, R = r.n, r.total_pairs

            def arr(ptr, cnt, dt):
                if not ptr or cnt == 0:
                    return np.zeros(cnt if ptr else 0, dtype=dt)
                return np.ctypeslib.as_array(ptr, shape=(cnt,)).copy()

            self.counts = arr(r.counts, n, np.uint32)
            self.offsets = arr(r.offsets, n + 1, np.uint64)
            self.nbrs = arr(r.nbrs, R, np.uint32)
            self.node_order = arr(r.node_order, n, np.uint32)

    def pairs(self):
        rows = np.repeat(np.arange(self.n, dtype=np.uint32), self.counts.astype(np.int64))
        return np.stack([rows, self.nbrs], axis=1)


def _as_input(data, storage):
    dt = np.float32 if storage == STORAGE_F32 else np.float64
    data = np.ascontiguousarray(data, dtype=dt)
    if data.ndim != 2:
        raise ValueError("data must be 2-D (n, d)")
    return data


def self_join(data, eps, **kw):
    """One-shot join: the nodeLauncher5 replacement (launcher.cu:3093-3532)."""
    p = default_params(**kw)
    data = _as_input(data, p.storage)
    n, d = data.shape
    r = Result()
    rc = lib().msj_self_join(data.ctypes.data_as(C.c_void_p), n, d, float(eps), C.byref(p),
                             C.byref(r))
    try:
        _check(rc)
        return JoinResult(r)
    finally:
        lib().msj_result_free(C.byref(r))


def probe_index(data, eps, layer=0, adjacency=False, **kw):
    """msj_probe_index: the GPU index build's candidate keys at `layer`
    ([ncand, n], row order), the final nodes' key tuples ([M, layers]), node
    boundaries, node order, numCalcs and, with adjacency=True (a second
    build sized by the first), the neighbour-node CSR.  Runs no join."""
    p = default_params(**kw)
    data = _as_input(data, p.storage)
    n, d = data.shape
    axes = min(p.maxrp, d) if p.index == INDEX_COSS else min(p.numidim, d)
    ncand = p.rp_per_layer + axes
    keys = np.zeros(ncand * n, dtype=np.uint32)
    nkeys = np.zeros(n * p.maxrp, dtype=np.uint32)
    nstart = np.zeros(n + 1, dtype=np.uint32)
    order = np.zeros(n, dtype=np.uint32)
    aoff = np.zeros(n + 1, dtype=np.uint32)
    up = C.POINTER(C.c_uint32)

    def build(adj):
        pr = IndexProbe()
        pr.layer = layer
        pr.keys = keys.ctypes.data_as(up)
        pr.node_keys = nkeys.ctypes.data_as(up)
        pr.node_start = nstart.ctypes.data_as(up)
        pr.node_order = order.ctypes.data_as(up)
        pr.adj_off = aoff.ctypes.data_as(up)
        if adj is not None:
            pr.adj = adj.ctypes.data_as(up)
            pr.adj_cap = adj.shape[0]
        _check(lib().msj_probe_index(data.ctypes.data_as(C.c_void_p), n, d, float(eps),
                                     C.byref(p), C.byref(pr)))
        return pr

    pr = build(None)
    adj = None
    if adjacency:
        adj = np.zeros(max(pr.num_edges, 1), dtype=np.uint32)
        pr = build(adj)
        adj = adj[:pr.num_edges]
    M, L = pr.num_nodes, pr.num_layers
    return {"keys": keys[:pr.ncand * n].reshape(pr.ncand, n),
            "node_keys": nkeys[:M * L].reshape(M, L), "node_start": nstart[:M + 1],
            "node_order": order, "num_nodes": M, "num_layers": L,
            "adj_off": aoff[:M + 1], "adj": adj, "num_edges": pr.num_edges,
            "num_calcs": pr.num_calcs, "chosen": list(pr.chosen)[:L],
            "nodes_per_layer": list(pr.nodes_per_layer)[:L], "sse": list(pr.sse)[:L]}


def probe_repeat_add(s, c, g):
    """msj_probe_repeat_add: s after g steps of s = RN(s + c) (host code)."""
    return lib().msj_probe_repeat_add(float(s), float(c), int(g))


def probe_key_math(D, eps, device=0):
    """msj_probe_key_math: (sqrt(D)/eps, key) computed on the GPU."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    q = np.zeros(D.shape[0], dtype=np.float64)
    k = np.zeros(D.shape[0], dtype=np.uint32)
    dp = C.POINTER(C.c_double)
    _check(lib().msj_probe_key_math(D.ctypes.data_as(dp), D.shape[0], float(eps),
                                    q.ctypes.data_as(dp),
                                    k.ctypes.data_as(C.POINTER(C.c_uint32)), device))
    return q, k


class Context:
    """Data resident in HBM across joins (msj_ctx_*)."""

    def __init__(self, **kw):
        self.params = default_params(**kw)
        h = C.c_void_p()
        _check(lib().msj_ctx_create(C.byref(self.params), C.byref(h)))
        self.h = h
        self.n = self.d = 0

    def load(self, data):
        data = _as_input(data, self.params.storage)
        n, d = data.shape
        _check(lib().msj_ctx_load(self.h, data.ctypes.data_as(C.c_void_p), n, d))
        self.n, self.d = n, d

    def set_params(self, **kw):
        """msj_ctx_set_params: change knobs (form, index, ...) for later joins."""
        p = Params.from_buffer_copy(self.params)
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError("msj_params has no field %r" % k)
            setattr(p, k, v)
        _check(lib().msj_ctx_set_params(self.h, C.byref(p)))
        self.params = p

    def load_unordered(self, rows):
        """msj_ctx_load_unordered: FP64 rows in file order; the dimension
        order (utils.cu:22-53) is computed and applied on the GPU.  Returns
        (order, sigma)."""
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        if rows.ndim != 2:
            raise ValueError("rows must be 2-D (n, d)")
        n, d = rows.shape
        order = np.zeros(d, dtype=np.uint32)
        sigma = np.zeros(d, dtype=np.float64)
        _check(lib().msj_ctx_load_unordered(self.h, rows.ctypes.data_as(C.POINTER(C.c_double)),
                                            n, d, order.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            sigma.ctypes.data_as(C.POINTER(C.c_double))))
        self.n, self.d = n, d
        return order, sigma

    def join(self, eps, copy_arrays=True):
        r = Result()
        rc = lib().msj_ctx_join(self.h, float(eps), C.byref(r))
        try:
            _check(rc)
            return JoinResult(r, copy_arrays=copy_arrays)
        finally:
            lib().msj_result_free(C.byref(r))

    def close(self):
        if self.h:
            lib().msj_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def _jobs(datas, eps, storage):
    arrs = [_as_input(x, storage) for x in datas]
    jobs = (Job * max(1, len(arrs)))()
    for k, (x, e) in enumerate(zip(arrs, eps)):
        jobs[k].data = x.ctypes.data
        jobs[k].n, jobs[k].d = x.shape
        jobs[k].eps = float(e)
    return arrs, jobs


def join_batch(datas, eps, **kw):
    """Batching form, one shot: a list of (n_k, d_k) arrays and their eps."""
    p = default_params(**kw)
    arrs, jobs = _jobs(datas, eps, p.storage)
    outs = (Result * max(1, len(arrs)))()
    rc = lib().msj_join_batch(jobs, len(arrs), C.byref(p), outs)
    try:
        _check(rc)
        return [JoinResult(outs[k]) for k in range(len(arrs))]
    finally:
        for k in range(len(arrs)):
            lib().msj_result_free(C.byref(outs[k]))


class Batch:
    """Batching form with the jobs resident in HBM across runs (msj_batch_*)."""

    def __init__(self, **kw):
        self.params = default_params(**kw)
        h = C.c_void_p()
        _check(lib().msj_batch_create(C.byref(self.params), C.byref(h)))
        self.h = h
        self.njobs = 0

    def load(self, datas, eps):
        arrs, jobs = _jobs(datas, eps, self.params.storage)
        _check(lib().msj_batch_load(self.h, jobs, len(arrs)))
        self.njobs = len(arrs)

    def run(self, copy_arrays=True):
        outs = (Result * max(1, self.njobs))()
        st = BatchStats()
        rc = lib().msj_batch_run(self.h, outs, C.byref(st))
        try:
            _check(rc)
            res = [JoinResult(outs[k], copy_arrays=copy_arrays) for k in range(self.njobs)]
        finally:
            for k in range(self.njobs):
                lib().msj_result_free(C.byref(outs[k]))
        stats = {f[0]: getattr(st, f[0]) for f in BatchStats._fields_}
        stats["class_ds"] = list(st.class_ds)[:st.classes]
        stats["class_epi_log2"] = list(st.class_epi_log2)[:st.classes]
        stats["class_slab"] = list(st.class_slab)[:st.classes]
        return res, stats

    def close(self):
        if self.h:
            lib().msj_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
