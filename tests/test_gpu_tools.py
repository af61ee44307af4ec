"""End-to-end runs of the tools on the GPU, checked against the oracle.

  * msj_main (reference src/main.cu:20-486): a raw FP64 file in, the
    reference's stdout lines and whitespace stderr record (main.cu:84,182,444;
    nodes.cu:28,329; launcher.cu:3119-3120,3233,3513-3515) and the --out
    arrays + manifest;
  * msj_verify (nodes.cu:749-885 nodeForce): 0 missing pairs on a correct
    result, and exactly the pairs planted (on the CPU) into a saved result;
  * result paging with a small page size (several pinned pages);
  * the device dimension order of msj_ctx_load_unordered (utils.cu:22-53);
  * the brute-force candidate generator, and the tree/COSS one through the
    driver (--index coss).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "self-join-mistic_amd")
MAIN = os.path.join(PKG, "msj_main")
VERIFY = os.path.join(PKG, "msj_verify")


def run(*args):
    print("[case] %s" % " ".join(str(a) for a in args), flush=True)
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=300)


def _raw(tmp_path, x, name="x.bin"):
    f = tmp_path / name
    np.ascontiguousarray(x, dtype="<f8").tofile(f)
    return f


def _ordered(x, msj_mod):
    order, _ = msj_mod.dim_order(x)
    return datagen.reorder_dims(x, order), order


@pytest.mark.parametrize("name,make,eps,extra", [
    ("mixture 20000x8", lambda: datagen.mixture(20000, 8, 81), 0.06, []),
    ("uniform 30000x2", lambda: datagen.uniform(30000, 2, 82), 0.01, []),
    ("correlated 4000x64 split", lambda: datagen.correlated(4000, 64, 83), 0.9,
     ["--form", "split"]),
    ("mixture 8000x18 f32", lambda: datagen.mixture(8000, 18, 84), 0.15, ["--f32"]),
])
def test_main_end_to_end(name, make, eps, extra, tmp_path, msj_mod, oracle_mod):
    x = make()
    # columns deliberately out of sigma order: the driver reorders them
    x = np.ascontiguousarray(x[:, ::-1])
    f = _raw(tmp_path, x)
    prefix = tmp_path / "res"
    r = run(MAIN, f, x.shape[1], 7, repr(eps), "--out", prefix, *extra)
    assert r.returncode == 0, r.stderr
    xo, order = _ordered(x, msj_mod)
    if "--f32" in extra:
        xo = xo.astype(np.float32).astype(np.float64)
    ref = oracle_mod.self_join(xo, eps)
    n, d = x.shape
    out = r.stdout
    assert "Number points: %d " % n in out
    assert "Number Dimensions: %d " % d in out
    assert "Final graph has %d nodes, %d calcs total" % (ref.num_nodes, ref.num_calcs) in out
    assert "Total results Set Size: %d " % ref.total_pairs in out
    for key in ("Time to read in file:", "Node Construct time:", "Batchs:", "Kernel Time:",
                "Node Laucnher time:", "Total Time:"):
        assert key in out, key
    for l in range(ref.num_layers):
        assert ("SubGraph 0 Layer %d Selecting RP %d with Nodes: %d and calcs: %d and stdev"
                % (l, ref.chosen[l], ref.nodes_per_layer[l], ref.calcs_per_layer[l])) in out
    # the stderr record: eps N 32 numSplits numNodes numCalcs constructT
    # numBatches varCalc kernelT totals launcherT totalT
    rec = r.stderr.split()
    assert len(rec) == 13, r.stderr
    assert rec[0] == "%f" % eps and int(rec[1]) == n and rec[2] == "32"
    assert int(rec[3]) == ref.num_layers and int(rec[4]) == ref.num_nodes
    assert int(rec[5]) == ref.num_calcs and int(rec[10]) == ref.total_pairs
    for k in (6, 8, 9, 11, 12):
        float(rec[k])
    assert int(rec[7]) >= 1
    # --out arrays and the manifest
    man = json.loads((tmp_path / "res.manifest.json").read_text())
    assert man["n"] == n and man["d"] == d and man["total_pairs"] == ref.total_pairs
    assert man["dim_order"] == [int(v) for v in order]
    arrays = {}
    for key, spec in man["files"].items():
        arrays[key] = np.fromfile(tmp_path / spec["path"], dtype=spec["dtype"])
        assert arrays[key].size == spec["count"], key
    assert np.array_equal(arrays["counts"], ref.counts)
    assert np.array_equal(arrays["offsets"], ref.offsets)
    assert np.array_equal(arrays["nbrs"], ref.nbrs)
    assert np.array_equal(arrays["node_order"], ref.node_order)


def test_main_tree_index(tmp_path, msj_mod, oracle_mod):
    """msj_main --index coss: the tree/COSS bins (tree.cu) as the candidate
    generator; the --out arrays equal the oracle's tree join."""
    x = datagen.mixture(12000, 6, 86)
    f = _raw(tmp_path, x)
    prefix = tmp_path / "coss"
    eps = 0.07
    r = run(MAIN, f, 6, 7, repr(eps), "--index", "coss", "--out", prefix)
    assert r.returncode == 0, r.stderr
    xo, _ = _ordered(x, msj_mod)
    ref = oracle_mod.self_join(xo, eps, index=2)
    assert "Total results Set Size: %d " % ref.total_pairs in r.stdout
    assert "Final graph has %d nodes, %d calcs total" % (ref.num_nodes, ref.num_calcs) in r.stdout
    assert np.array_equal(np.fromfile(str(prefix) + ".counts.u32", dtype="<u4"), ref.counts)
    assert np.array_equal(np.fromfile(str(prefix) + ".nbrs.u32", dtype="<u4"), ref.nbrs)
    assert np.array_equal(np.fromfile(str(prefix) + ".node_order.u32", dtype="<u4"),
                          ref.node_order)


def test_verify_clean_and_planted(tmp_path):
    x = datagen.mixture(6000, 8, 85)
    f = _raw(tmp_path, x)
    eps = 0.07
    r = run(VERIFY, f, 8, repr(eps))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Missing 0 pairs:" in r.stdout and "Extra 0 pairs:" in r.stdout
    # a saved result with one pair taken out and one pair put in (on the CPU)
    prefix = tmp_path / "saved"
    assert run(MAIN, f, 8, 7, repr(eps), "--out", prefix).returncode == 0
    counts = np.fromfile(str(prefix) + ".counts.u32", dtype="<u4")
    offs = np.fromfile(str(prefix) + ".offsets.u64", dtype="<u8")
    nbrs = np.fromfile(str(prefix) + ".nbrs.u32", dtype="<u4")
    row = int(np.argmax(counts))
    lo, hi = int(offs[row]), int(offs[row + 1])
    victim = int(nbrs[lo + 1])
    nb_row = nbrs[lo:hi]
    free = int(np.setdiff1d(np.arange(x.shape[0]), nb_row)[0])
    new_row = np.sort(np.concatenate([np.delete(nb_row, 1), [free]]).astype(np.uint32))
    nbrs2 = np.concatenate([nbrs[:lo], new_row, nbrs[hi:]]).astype("<u4")
    nbrs2.tofile(str(prefix) + ".nbrs.u32")
    r = run(VERIFY, f, 8, repr(eps), "--against", prefix)
    assert r.returncode == 3
    assert "Missing 1 pairs:\n(%d,%d)," % (row, victim) in r.stdout
    assert "Extra 1 pairs:\n(%d,%d)," % (row, free) in r.stdout


def test_paging_many_pages(msj_mod, oracle_mod):
    x, _ = _ordered(datagen.mixture(20000, 8, 86), msj_mod)
    eps = 0.08
    print("[case] paging page_pairs=3000", flush=True)
    got = msj_mod.self_join(x, eps, page_pairs=3000)
    ref = oracle_mod.self_join(x, eps)
    assert got.stats["num_pages"] == -(-int(ref.total_pairs) // 3000) > 1
    assert np.array_equal(got.counts, ref.counts)
    assert np.array_equal(got.offsets, ref.offsets)
    assert np.array_equal(got.nbrs, ref.nbrs)
    assert np.array_equal(got.node_order, ref.node_order)
    # one page smaller than the counts array too
    got = msj_mod.self_join(x, eps, page_pairs=777)
    assert np.array_equal(got.nbrs, ref.nbrs) and np.array_equal(got.counts, ref.counts)


@pytest.mark.parametrize("f32", [False, True])
def test_load_unordered_matches_host_order(f32, msj_mod, oracle_mod):
    rng = np.random.default_rng(87)
    x = datagen.mixture(30000, 12, 87)[:, rng.permutation(12)]
    x[:, 5] = x[:, 2]  # a tie in sigma: higher index first
    order_h, sigma_h = msj_mod.dim_order(x)
    storage = msj_mod.STORAGE_F32 if f32 else msj_mod.STORAGE_F64
    print("[case] load_unordered f32=%s" % f32, flush=True)
    with msj_mod.Context(storage=storage) as ctx:
        order_d, sigma_d = ctx.load_unordered(x)
        got = ctx.join(0.1)
    assert np.array_equal(order_d, order_h)
    assert np.array_equal(sigma_d, sigma_h)  # same sequential sums, bit for bit
    xo = datagen.reorder_dims(x, order_h)
    if f32:
        xo = xo.astype(np.float32).astype(np.float64)
    ref = oracle_mod.self_join(xo, 0.1)
    assert np.array_equal(got.counts, ref.counts) and np.array_equal(got.nbrs, ref.nbrs)
    assert np.array_equal(got.node_order, ref.node_order)


@pytest.mark.parametrize("form", [0, 2])
def test_brute_force_generator(form, msj_mod, oracle_mod):
    x, _ = _ordered(datagen.correlated(5000, 40, 88), msj_mod)
    eps = 0.55
    print("[case] brute force form=%d" % form, flush=True)
    got = msj_mod.self_join(x, eps, index=msj_mod.INDEX_BRUTE, form=form)
    bf = oracle_mod.brute_force(x, eps)
    assert got.stats["num_nodes"] == 1
    assert got.stats["num_calcs"] == x.shape[0] ** 2
    assert np.array_equal(got.node_order, np.arange(x.shape[0]))
    assert np.array_equal(got.counts, bf.counts)
    assert np.array_equal(got.nbrs, bf.nbrs)
    net = msj_mod.self_join(x, eps)
    assert np.array_equal(net.nbrs, bf.nbrs)


def test_set_params_switches_generator(msj_mod):
    x, _ = _ordered(datagen.mixture(5000, 8, 89), msj_mod)
    print("[case] set_params", flush=True)
    with msj_mod.Context() as ctx:
        ctx.load(x)
        a = ctx.join(0.08)
        ctx.set_params(index=msj_mod.INDEX_BRUTE)
        b = ctx.join(0.08)
        with pytest.raises(msj_mod.MsjError):
            ctx.set_params(storage=msj_mod.STORAGE_F32)
    assert a.stats["num_nodes"] > 1 and b.stats["num_nodes"] == 1
    assert np.array_equal(a.nbrs, b.nbrs) and np.array_equal(a.counts, b.counts)
