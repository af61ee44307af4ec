"""The command-line tools without a GPU: input checks of msj_main (reference
src/main.cu:25-50), the CSV converter (toCSV.cpp:51-85) and argument errors
of the C ABI that must come back before any device is touched."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "self-join-mistic_amd")
MAIN = os.path.join(PKG, "msj_main")
VERIFY = os.path.join(PKG, "msj_verify")
CSV2BIN = os.path.join(PKG, "msj_csv2bin")


def run(*args, **kw):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120,
                          **kw)


# ---------------------------------------------------------------- msj_main


def test_dry_run_reports_points_and_dims(tmp_path):
    x = np.random.default_rng(1).random((1000, 3))
    f = tmp_path / "x.bin"
    x.tofile(f)
    r = run(MAIN, f, 3, 7, 0.05, "--dry-run")
    assert r.returncode == 0, r.stderr
    assert "Number points: 1000 " in r.stdout
    assert "Number Dimensions: 3 " in r.stdout
    assert "Number Reference Points: 7 " in r.stdout
    assert "Distance Threshold: 0.050000 " in r.stdout
    assert r.stderr.split() == ["0.050000", "1000"]  # the record's first fields (main.cu:84)


def test_trailing_partial_row(tmp_path):
    """main.cu:50 drops a trailing partial row (N = size/8/dim); msj_main does
    the same with a warning, and refuses it under --strict."""
    x = np.random.default_rng(2).random(3001)  # 1000 rows of 3 and one extra value
    f = tmp_path / "x.bin"
    x.tofile(f)
    r = run(MAIN, f, 3, 7, 0.05, "--dry-run")
    assert r.returncode == 0
    assert "Number points: 1000 " in r.stdout
    assert "Warning: 8 trailing bytes" in r.stdout
    r = run(MAIN, f, 3, 7, 0.05, "--dry-run", "--strict")
    assert r.returncode == 1
    assert "8 trailing bytes" in r.stderr
    # a file shorter than one row: zero points
    g = tmp_path / "short.bin"
    np.zeros(2).tofile(g)
    r = run(MAIN, g, 3, 7, 0.05, "--dry-run")
    assert r.returncode == 0 and "Number points: 0 " in r.stdout


def test_point_ids_beyond_32_bits_are_refused(tmp_path):
    """2^32 points of one dimension (a sparse 32 GiB file): the reference's
    unsigned numPoints would wrap (main.cu:50); msj_main stops before mapping."""
    f = tmp_path / "huge.bin"
    with open(f, "wb") as h:
        h.truncate(8 * (1 << 32))
    r = run(MAIN, f, 1, 7, 0.05, "--dry-run")
    assert r.returncode == 1
    assert "exceed the 32-bit point ids" in r.stderr


def test_values_beyond_2_pow_40_are_refused(tmp_path):
    f = tmp_path / "wide.bin"
    d = 1 << 20
    with open(f, "wb") as h:  # 2^20 rows of 2^20 doubles: 8 TiB, sparse
        try:
            h.truncate(8 * d * d)
        except OSError:
            pytest.skip("file system refuses an 8 TiB sparse file")
    r = run(MAIN, f, d, 7, 0.05, "--dry-run")
    assert r.returncode == 1
    assert "exceed 2^40 values" in r.stderr


@pytest.mark.parametrize("args", [
    [],                       # too few arguments
    ["x.bin", "0", "7", "0.1"],   # d = 0
    ["x.bin", "3", "7", "-1"],    # eps <= 0
    ["x.bin", "3", "7", "nan"],
    ["x.bin", "3", "7", "abc"],
    ["x.bin", "3x", "7", "0.1"],
    ["x.bin", "3", "7", "0.1", "--form", "fp8"],
])
def test_usage_errors(args, tmp_path):
    np.zeros(6).tofile(tmp_path / "x.bin")
    r = run(MAIN, *args, cwd=tmp_path)
    assert r.returncode == 2


def test_missing_file():
    r = run(MAIN, "/nonexistent/x.bin", 2, 7, 0.01)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_verify_usage():
    assert run(VERIFY).returncode == 2
    assert run(VERIFY, "x.bin", "0", "0.1").returncode == 2


# ------------------------------------------------------------- msj_csv2bin


def test_csv2bin_reads_as_the_reference(tmp_path):
    """Numbers separated by an optional comma and white space (the fscanf
    "%lf, " / "%lf " loop of toCSV.cpp:61-64), written as raw FP64."""
    vals = np.random.default_rng(3).random((50, 4)) * 1e3 - 500
    lines = []
    for k, row in enumerate(vals):
        sep = [", ", ",", " ", ",\t"][k % 4]
        lines.append(sep.join(repr(float(v)) for v in row) + ("," if k % 3 == 0 else ""))
    src = tmp_path / "t.csv"
    src.write_text("\n".join(lines) + "\n")
    out = tmp_path / "t.bin"
    r = run(CSV2BIN, src, out)
    assert r.returncode == 0, r.stderr
    assert "size of data: 200" in r.stdout
    assert np.array_equal(np.fromfile(out, dtype="<f8"), vals.ravel())
    r = run(CSV2BIN, src, out, "--dim", 4)
    assert r.returncode == 0


def test_csv2bin_stops_at_text(tmp_path):
    src = tmp_path / "t.csv"
    src.write_text("1.5, 2.5\n3.5, x, 4.5\n")
    out = tmp_path / "t.bin"
    r = run(CSV2BIN, src, out)
    assert r.returncode == 0
    assert np.array_equal(np.fromfile(out, dtype="<f8"), [1.5, 2.5, 3.5])
    r = run(CSV2BIN, src, out, "--dim", 2)
    assert r.returncode == 3
    # a comma must follow the number directly: "%lf, " matches no white
    # space before it, and the next "%lf" then meets the comma
    src.write_text("1.5\t, 2.5\n")
    assert run(CSV2BIN, src, out).returncode == 0
    assert np.array_equal(np.fromfile(out, dtype="<f8"), [1.5])


def test_csv2bin_ragged_rows(tmp_path):
    src = tmp_path / "t.csv"
    src.write_text("1,2,3\n4,5\n6,7,8\n")
    r = run(CSV2BIN, src, tmp_path / "o.bin", "--dim", 3)
    assert r.returncode == 3 and "line 2" in r.stderr
    assert run(CSV2BIN, src, tmp_path / "o.bin").returncode == 0  # the reference keeps all 8


def test_csv2bin_then_main_dry_run(tmp_path):
    src = tmp_path / "t.csv"
    src.write_text("\n".join("%d.25, %d.5" % (i, i) for i in range(300)) + "\n")
    out = tmp_path / "t.bin"
    assert run(CSV2BIN, src, out, "--dim", 2).returncode == 0
    r = run(MAIN, out, 2, 7, 0.1, "--dry-run")
    assert r.returncode == 0 and "Number points: 300 " in r.stdout


def test_csv2bin_usage_and_missing(tmp_path):
    assert run(CSV2BIN).returncode == 2
    assert run(CSV2BIN, "a", "b", "--dim", "0").returncode == 2
    r = run(CSV2BIN, tmp_path / "none.csv", tmp_path / "o.bin")
    assert r.returncode == 1 and "No such File" in r.stdout


# ------------------------------------------------- C ABI argument checks


def test_self_join_argument_errors_need_no_device(msj_mod):
    """Oversized n*d, d = 0, a NULL pointer and a bad eps are refused before
    any device is touched (this container has none)."""
    L = msj_mod.lib()
    p = msj_mod.default_params()
    r = msj_mod.Result()
    big_n, big_d = 1 << 30, 1 << 10  # n*d = 2^40
    assert L.msj_self_join(C.c_void_p(1), big_n, big_d, 0.1, C.byref(p), C.byref(r)) == \
        msj_mod.E_OVERFLOW
    assert L.msj_self_join(C.c_void_p(1), 10, 0, 0.1, C.byref(p), C.byref(r)) == msj_mod.E_INVAL
    assert L.msj_self_join(None, 10, 2, 0.1, C.byref(p), C.byref(r)) == msj_mod.E_INVAL
    x = np.zeros((4, 2))
    for eps in (0.0, -1.0, float("nan"), float("inf")):
        assert L.msj_self_join(x.ctypes.data_as(C.c_void_p), 4, 2, eps, C.byref(p),
                               C.byref(r)) == msj_mod.E_INVAL
    assert L.msj_self_join(x.ctypes.data_as(C.c_void_p), 4, 2, 0.1, C.byref(p), None) == \
        msj_mod.E_INVAL
    q = msj_mod.default_params(index=7)
    assert L.msj_self_join(x.ctypes.data_as(C.c_void_p), 4, 2, 0.1, C.byref(q),
                           C.byref(r)) == msj_mod.E_INVAL
    assert "index" in msj_mod.lib().msj_last_error().decode()
