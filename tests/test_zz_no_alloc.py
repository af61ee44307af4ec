"""msj_self_join_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory) against msj_self_join, the entry point it
mirrors: identical counts, offsets, neighbour lists, node order and node-net
statistics, on two calls in a row with the second input smaller, in the
plain and the split-BF16 forms.  The test allocates the device inputs and
outputs itself (through the HIP runtime libmsj.so links), as a caller of the
entry point does."""
import ctypes as C

import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu

HIP_MEMCPY_DEFAULT = 4


class Hip:
    """The caller's side: device buffers from the same HIP runtime."""

    def __init__(self, msj_mod):
        msj_mod.lib()  # loads libamdhip64.so.7 as its dependency
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 1)) == 0
        self.live.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.L.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes,
                                HIP_MEMCPY_DEFAULT) == 0
        return p

    def get(self, p, count, dtype):
        out = np.zeros(count, dtype=dtype)
        if count:
            assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes,
                                    HIP_MEMCPY_DEFAULT) == 0
        return out

    def free_all(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _noalloc(msj_mod, hip, x, eps, nbrs_cap, **kw):
    n, d = x.shape
    p = msj_mod.default_params(**kw)
    rows = hip.put(x)
    counts = hip.alloc(4 * n)
    offsets = hip.alloc(8 * (n + 1))
    nbrs = hip.alloc(4 * nbrs_cap)
    order = hip.alloc(4 * n)
    st = msj_mod.Stats()
    rc = msj_mod.lib().msj_self_join_noalloc(rows, n, d, float(eps), C.byref(p), counts, offsets,
                                             nbrs, nbrs_cap, order, C.byref(st))
    if rc != msj_mod.OK:
        return rc, st, None
    R = st.total_pairs
    res = {"counts": hip.get(counts, n, np.uint32), "offsets": hip.get(offsets, n + 1, np.uint64),
           "nbrs": hip.get(nbrs, R, np.uint32), "node_order": hip.get(order, n, np.uint32)}
    return rc, st, res


def _same(res, st, ref):
    assert st.total_pairs == ref.total_pairs
    assert st.num_nodes == ref.stats["num_nodes"]
    assert st.num_calcs == ref.stats["num_calcs"]
    assert list(st.chosen)[:st.num_layers] == ref.stats["chosen"]
    assert np.array_equal(res["counts"], ref.counts)
    assert np.array_equal(res["offsets"], ref.offsets)
    assert np.array_equal(res["nbrs"], ref.nbrs)
    assert np.array_equal(res["node_order"], ref.node_order)


@pytest.mark.parametrize("form,cases", [
    (0, [((20000, 18, 91), 0.12), ((6000, 18, 92), 0.15)]),
    (2, [((6000, 64, 93), 0.9), ((2500, 64, 94), 0.9)]),
])
def test_noalloc_equals_self_join_twice(form, cases, msj_mod):
    hip = Hip(msj_mod)
    try:
        for (n, d, seed), eps in cases:
            x = datagen.mixture(n, d, seed) if d < 32 else datagen.correlated(n, d, seed)
            order, _ = msj_mod.dim_order(x)
            x = np.ascontiguousarray(datagen.reorder_dims(x, order))
            print("\n[case] noalloc form=%d n=%d d=%d eps=%r" % (form, n, d, eps), flush=True)
            ref = msj_mod.self_join(x, eps, form=form)
            rc, st, res = _noalloc(msj_mod, hip, x, eps, int(ref.total_pairs) + 10, form=form)
            assert rc == msj_mod.OK, msj_mod.lib().msj_last_error()
            _same(res, st, ref)
            assert st.form_used == ref.stats["form_used"]
            hip.free_all()
    finally:
        hip.free_all()


def test_noalloc_capacity_and_size_errors(msj_mod):
    hip = Hip(msj_mod)
    try:
        x = datagen.mixture(3000, 8, 95)
        order, _ = msj_mod.dim_order(x)
        x = np.ascontiguousarray(datagen.reorder_dims(x, order))
        print("\n[case] noalloc nbrs_cap too small", flush=True)
        ref = msj_mod.self_join(x, 0.08)
        rc, st, _ = _noalloc(msj_mod, hip, x, 0.08, 10)
        assert rc == msj_mod.E_OVERFLOW
        assert st.total_pairs == ref.total_pairs  # the room a retry needs
        # the retry with that room succeeds and matches
        rc, st, res = _noalloc(msj_mod, hip, x, 0.08, int(st.total_pairs))
        assert rc == msj_mod.OK
        _same(res, st, ref)
        # rows whose copies alone exceed the scratch: refused before any work
        p = msj_mod.default_params()
        big = msj_mod.lib().msj_self_join_noalloc(C.c_void_p(8), 4_000_000, 8, 0.1, C.byref(p),
                                                  C.c_void_p(8), C.c_void_p(8), C.c_void_p(8),
                                                  1, C.c_void_p(8), None)
        assert big == msj_mod.E_OVERFLOW
        assert b"scratch" in msj_mod.lib().msj_last_error()
    finally:
        hip.free_all()
