"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/msj.h declares, the Python mirror's struct layouts equal the C
compiler's, defaults equal the params.cuh knobs, and argument errors come back
as status codes (never aborts)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# C headers of libmsj.so (msj_nodelauncher.h belongs to the reference-side shim)
HEADERS = [os.path.join(ROOT, "include", h) for h in ("msj.h", "msj_probe.h")]
PARAMS_H = os.path.join(ROOT, "include", "msj_params.h")


def _declared():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(msj_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_exports_every_declared_symbol(msj_mod):
    names = _declared()
    assert names == sorted(msj_mod.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", msj_mod.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (msj_\w+)", out))
    for n in names:
        assert n in exported, n
        assert hasattr(msj_mod.lib(), n)


def test_struct_layout_matches_c(tmp_path, msj_mod):
    src = tmp_path / "layout.c"
    fields = {
        "msj_params": [f[0] for f in msj_mod.Params._fields_],
        "msj_stats": [f[0] for f in msj_mod.Stats._fields_],
        "msj_result": [f[0] for f in msj_mod.Result._fields_],
        "msj_job": [f[0] for f in msj_mod.Job._fields_],
        "msj_batch_stats": [f[0] for f in msj_mod.BatchStats._fields_],
        "msj_index_probe": [f[0] for f in msj_mod.IndexProbe._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "msj.h"',
             '#include "msj_probe.h"', "int main(void){"]
    for st, fl in fields.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (st, st))
        for f in fl:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (st, f, st, f))
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run(
        [str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    py = {"msj_params": msj_mod.Params, "msj_stats": msj_mod.Stats, "msj_result": msj_mod.Result,
          "msj_job": msj_mod.Job, "msj_batch_stats": msj_mod.BatchStats,
          "msj_index_probe": msj_mod.IndexProbe}
    for st, cls in py.items():
        assert int(got[st]) == C.sizeof(cls), st
        for f in fields[st]:
            assert int(got["%s.%s" % (st, f)]) == getattr(cls, f).offset, (st, f)


def test_defaults_are_the_reference_knobs(msj_mod):
    defs = dict(re.findall(r"#define\s+(\w+)\s+([0-9.]+)", open(PARAMS_H).read()))
    p = msj_mod.default_params()
    assert p.struct_size == C.sizeof(msj_mod.Params)
    assert p.maxrp == int(defs["MAXRP"]) == 6
    assert p.rp_per_layer == int(defs["RPPERLAYER"]) == 32
    assert p.numidim == int(defs["NUMIDIM"]) == 6
    assert p.sample_per == float(defs["SAMPLE_PER"]) == 0.01
    assert p.min_calcs_per_node == 100000
    assert p.seed == 1 and p.form == msj_mod.FORM_PLAIN and p.storage == msj_mod.STORAGE_F64


def test_strerror_and_version(msj_mod):
    assert "gfx950" in msj_mod.version()
    for code in (0, -1, -2, -3, -4, -5, -6):
        assert msj_mod.lib().msj_strerror(code)
    assert msj_mod.lib().msj_strerror(-99) == b"unknown status"


def test_dim_order_matches_oracle(msj_mod, oracle_mod):
    rng = np.random.default_rng(8)
    a = rng.random((20000, 11)) * rng.random(11)
    a[:, 7] = a[:, 2]
    o1, s1 = msj_mod.dim_order(a)
    o2, s2 = oracle_mod.dim_order(a)
    assert np.array_equal(o1, o2)
    assert np.array_equal(s1, s2)  # same sequential sums, bit for bit


@pytest.mark.parametrize("field,value,code", [
    ("struct_size", 8, -1),
    ("rp_per_layer", 33, -1),
    ("maxrp", 0, -1),
    ("maxrp", 17, -1),
    ("sample_per", 1.5, -1),
    ("storage", 7, -1),
    ("form", 5, -1),
])
def test_bad_params_are_rejected(field, value, code, msj_mod):
    x = np.random.default_rng(1).random((50, 2))
    with pytest.raises(msj_mod.MsjError) as e:
        msj_mod.self_join(x, 0.1, **{field: value})
    assert e.value.code == code


def test_bad_output_mode_rejected_by_context(msj_mod):
    with pytest.raises(msj_mod.MsjError) as e:
        msj_mod.Context(output=9)
    assert e.value.code == msj_mod.E_INVAL


def test_null_and_bad_arguments(msj_mod):
    L = msj_mod.lib()
    p = msj_mod.default_params()
    r = msj_mod.Result()
    assert L.msj_ctx_create(C.byref(p), None) == msj_mod.E_INVAL
    assert L.msj_ctx_load(None, None, 1, 1) == msj_mod.E_INVAL
    assert L.msj_ctx_join(None, 0.1, C.byref(r)) == msj_mod.E_INVAL
    assert L.msj_dim_order(None, 1, 1, None, None) == msj_mod.E_INVAL
    L.msj_result_free(None)
    L.msj_ctx_destroy(None)


def test_driver_usage_errors():
    exe = os.path.join(ROOT, "self-join-mistic_amd", "msj_main")
    assert subprocess.run([exe], capture_output=True).returncode == 2
    r = subprocess.run([exe, "/nonexistent.bin", "2", "7", "0.01"], capture_output=True)
    assert r.returncode == 1
