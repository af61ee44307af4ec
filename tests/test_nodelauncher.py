"""The reference's entry point nodeLauncher5 and its result type struct
neighborTable (reference src/launcher.cu:3093-3532, 4-50, 3382-3395;
include/launcher.cuh:77-82; include/utils.cuh:33-41), served by
libmsj_nodelauncher.so over the C ABI.

CPU: the library exports the reference's C++ symbol with its exact
signature, and a translation unit written against the reference's call
(main.cu:171-178) compiles and links.  GPU: the tables equal the oracle's
result mapped to node-order positions: entry 0 the placeholder
{p, p, pointArray}, entry 1 the neighbours of p as node-order positions.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "self-join-mistic_amd")
LIB = os.path.join(PKG, "libmsj_nodelauncher.so")
HEADER = os.path.join(ROOT, "include", "msj_nodelauncher.h")

# Itanium C++ mangling of nodeLauncher5(double*, unsigned, unsigned, unsigned,
# unsigned*, double): the symbol the reference's main.o links against
MANGLED = "_Z13nodeLauncher5PdjjjPjd"


def _declared_c():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(msj_[a-z_0-9]+)\s*\(", txt)))


def test_exports(tmp_path):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (\S+)", out))
    assert MANGLED in exported
    for name in _declared_c():
        assert name in exported, name


def test_reference_call_site_compiles(tmp_path):
    """A caller written as the reference's main calls it (main.cu:171-178)."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include <type_traits>
#include "msj_nodelauncher.h"
static_assert(std::is_same<decltype(&nodeLauncher5),
              struct neighborTable *(*)(double *, unsigned int, unsigned int, unsigned int,
                                        unsigned int *, double)>::value, "signature");
static_assert(std::is_same<decltype(neighborTable::cntNDataArrays), unsigned int>::value, "");
static_assert(std::is_same<decltype(neighborTable::vectdataPtr),
                           std::vector<unsigned int *>>::value, "");
int main(int argc, char **argv)
{
    if (argc < 0) {  // link check only
        double *d = nullptr; unsigned int *pa = nullptr;
        struct neighborTable *t = nodeLauncher5(d, 2, 0, 7, pa, 0.01);
        msj_neighbor_table_free(t, 0);
    }
    return 0;
}
''')
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", PKG, "-lmsj_nodelauncher", "-lmsj",
                    "-Wl,-rpath," + PKG], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def _lib():
    L = C.CDLL(LIB)
    up = C.POINTER(C.c_uint32)
    L.msj_nodelauncher5.restype = C.c_void_p
    L.msj_nodelauncher5.argtypes = [C.POINTER(C.c_double), C.c_uint, C.c_uint, up, C.c_double,
                                    C.POINTER(C.c_int)]
    L.msj_neighbor_table_flatten.restype = C.c_uint64
    L.msj_neighbor_table_flatten.argtypes = [C.c_void_p, C.c_uint, up, up]
    L.msj_neighbor_table_entries.restype = C.c_uint
    L.msj_neighbor_table_entries.argtypes = [C.c_void_p, C.c_uint]
    L.msj_neighbor_table_entry.restype = up
    L.msj_neighbor_table_entry.argtypes = [C.c_void_p, C.c_uint, C.c_uint, up, up]
    L.msj_neighbor_table_free.argtypes = [C.c_void_p, C.c_uint]
    return L


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,eps", [
    ("uniform 20000x2", lambda: datagen.uniform(20000, 2, 11), 0.01),
    ("mixture 20000x18", lambda: datagen.mixture(20000, 18, 13), 0.12),
    ("correlated 5000x90", lambda: datagen.correlated(5000, 90, 22), 1.14),
])
def test_tables_equal_oracle_ranges(name, make, eps, msj_mod, oracle_mod):
    x = make()
    order, _ = msj_mod.dim_order(x)
    x = np.ascontiguousarray(datagen.reorder_dims(x, order))
    n, d = x.shape
    print("\n[case] nodeLauncher5 %s eps=%r" % (name, eps), flush=True)
    L = _lib()
    pa = np.zeros(n, dtype=np.uint32)
    rc = C.c_int(0)
    t = L.msj_nodelauncher5(x.ctypes.data_as(C.POINTER(C.c_double)), d, n,
                            pa.ctypes.data_as(C.POINTER(C.c_uint32)), eps, C.byref(rc))
    assert t, rc.value
    try:
        ref = oracle_mod.self_join(x, eps)
        assert np.array_equal(pa, ref.node_order)  # pointArray as buildNodeNet leaves it
        counts = np.zeros(n, dtype=np.uint32)
        ids = np.zeros(int(ref.total_pairs), dtype=np.uint32)
        R = L.msj_neighbor_table_flatten(t, n, counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         ids.ctypes.data_as(C.POINTER(C.c_uint32)))
        assert R == ref.total_pairs
        pos = np.empty(n, dtype=np.int64)
        pos[ref.node_order] = np.arange(n)
        assert np.array_equal(counts, ref.counts[ref.node_order])
        off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
        lo, hi = C.c_uint(0), C.c_uint(0)
        for p in range(0, n, max(1, n // 2000)):
            row = ref.node_order[p]
            want = np.sort(pos[ref.nbrs[ref.offsets[row]:ref.offsets[row + 1]]])
            assert np.array_equal(ids[off[p]:off[p + 1]], want), p
            # entry 0: the placeholder {p, p, pointArray} (launcher.cu:3382-3395)
            assert L.msj_neighbor_table_entries(t, p) == (2 if counts[p] else 1)
            ptr = L.msj_neighbor_table_entry(t, p, 0, C.byref(lo), C.byref(hi))
            assert lo.value == p and hi.value == p and ptr[p] == row
            ptr = L.msj_neighbor_table_entry(t, p, 1, C.byref(lo), C.byref(hi))
            assert hi.value - lo.value + 1 == counts[p]
        # every position: neighbour multiset through the flattened ids
        full = [np.sort(pos[ref.nbrs[ref.offsets[r]:ref.offsets[r + 1]]])
                for r in ref.node_order]
        assert np.array_equal(ids, np.concatenate(full) if full else ids)
    finally:
        L.msj_neighbor_table_free(t, n)
