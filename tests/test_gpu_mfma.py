"""GPU parity of the precision forms (BF16 / split-BF16 MFMA filter + FP64
band recheck) against the CPU oracle, through the C ABI.

The bar is the plain form's: identical counts, offsets, neighbour lists,
node order and node-net statistics.  With count_work the library also
verifies, in FP64, the filter's class of every pair of a sample of work
items; a certain-accept or certain-reject class that FP64 contradicts would
be a wrong error bound (filter_violations must be 0).
"""
import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu

FORMS = ["bf16", "split"]


def _form(msj_mod, name):
    return msj_mod.FORM_MFMA_BF16 if name == "bf16" else msj_mod.FORM_MFMA_SPLIT


def _announce(name, x, eps, form):
    print("\n[case] %s form=%s n=%d d=%d eps=%r" % (name, form, x.shape[0], x.shape[1], eps),
          flush=True)


def _assert_same(got, ref):
    assert got.total_pairs == ref.total_pairs
    assert np.array_equal(got.counts, ref.counts)
    assert np.array_equal(got.offsets, ref.offsets)
    assert np.array_equal(got.nbrs, ref.nbrs)
    assert np.array_equal(got.node_order, ref.node_order)
    st = got.stats
    assert st["num_nodes"] == ref.num_nodes
    assert st["num_calcs"] == ref.num_calcs
    assert st["chosen"] == ref.chosen
    assert st["sse"] == ref.sse


def _prep(x, msj_mod):
    order, _ = msj_mod.dim_order(np.asarray(x, dtype=np.float64))
    return datagen.reorder_dims(x, order)


CASES = [
    # name, generator, eps
    ("correlated 6000x90", lambda: datagen.correlated(6000, 90, 22), 1.14),
    ("correlated 20000x90 dense", lambda: datagen.correlated(20000, 90, 23), 0.77),
    ("mixture 6000x57", lambda: datagen.mixture(6000, 57, 21), 0.553),
    ("mixture 20000x18", lambda: datagen.mixture(20000, 18, 13), 0.204),
    ("uniform 4096x2", lambda: datagen.uniform(4096, 2, 11), 0.02),
    ("mixture 3000x33", lambda: datagen.mixture(3000, 33, 24), 0.45),
    # either side of the tile-geometry switch: d = 128 (8 k-steps) runs half
    # tiles, d = 129 (10 k-steps) full 256 x 256 tiles
    ("correlated 4000x128 half tiles", lambda: datagen.correlated(4000, 128, 41), 1.497),
    ("correlated 4000x129 full tiles", lambda: datagen.correlated(4000, 129, 42), 1.495),
]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("name,gen,eps", CASES, ids=[c[0] for c in CASES])
def test_filter_parity(name, gen, eps, form, msj_mod, oracle_mod):
    x = _prep(gen(), msj_mod)
    _announce(name, x, eps, form)
    got = msj_mod.self_join(x, eps, form=_form(msj_mod, form))
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)
    st = got.stats
    assert st["form_used"] == _form(msj_mod, form)
    assert st["mfma_flops"] > 0
    print("[case] band %d pairs, max score err %.3g" % (st["band_pairs"], st["max_score_err"]),
          flush=True)


@pytest.mark.parametrize("form", FORMS)
def test_filter_fp32_384(form, msj_mod, oracle_mod):
    x = _prep(datagen.clusters(3000, 384, seed=31, ncl=40), msj_mod)
    eps = 1.26
    _announce("fp32 clusters 3000x384", x, eps, form)
    got = msj_mod.self_join(x, eps, form=_form(msj_mod, form), storage=msj_mod.STORAGE_F32)
    ref = oracle_mod.self_join(x.astype(np.float64), eps)
    _assert_same(got, ref)


def test_filter_bf16_first_settles_tiles(msj_mod, oracle_mod):
    """Split-BF16 above d = 128: tiles between rows of different clusters are
    settled by their BF16 scores, the others take the cross products; the
    pair set equals the oracle's and the sampled classes hold in FP64."""
    x = _prep(datagen.clusters(12000, 384, seed=33, ncl=60), msj_mod)
    eps = 1.26
    _announce("fp32 clusters 12000x384 BF16 first", x, eps, "split")
    got = msj_mod.self_join(x, eps, form=msj_mod.FORM_MFMA_SPLIT,
                            storage=msj_mod.STORAGE_F32, count_work=1)
    ref = oracle_mod.self_join(x.astype(np.float64), eps)
    _assert_same(got, ref)
    st = got.stats
    assert st["form_used"] == msj_mod.FORM_MFMA_SPLIT
    assert 0 < st["split_share_ppm"] < 1000000  # some settled in BF16, some not
    assert st["filter_violations"] == 0
    # the BF16 first pass's own certain rejects were among the verified pairs
    assert st["filter_checked_bf16"] > 0
    print("[case] cross products on %.1f%% of the pairs" % (st["split_share_ppm"] / 1e4),
          flush=True)


BOUND_CASES = [
    ("correlated 20000x90 half tiles", lambda: datagen.correlated(20000, 90, 25), 0.77),
    # full tiles: split-BF16 runs its BF16 pass first (class-3 checks)
    ("correlated 4000x160 full tiles", lambda: datagen.correlated(4000, 160, 43), 1.6),
]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("name,gen,eps", BOUND_CASES, ids=[c[0] for c in BOUND_CASES])
def test_filter_bound_verified(name, gen, eps, form, msj_mod, oracle_mod):
    """count_work: sampled pairs' filter classes re-decided in FP64."""
    x = _prep(gen(), msj_mod)
    _announce("bound check " + name, x, eps, form)
    got = msj_mod.self_join(x, eps, form=_form(msj_mod, form), count_work=1)
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)
    st = got.stats
    assert st["filter_checked"] > 100000
    assert st["filter_violations"] == 0
    if x.shape[1] > 128 and st["form_used"] == msj_mod.FORM_MFMA_SPLIT:
        assert st["filter_checked_bf16"] > 0
    print("[case] checked %d sampled pairs (%d BF16-first rejects), mean |score err| %.3g" % (
        st["filter_checked"], st["filter_checked_bf16"], st["mean_score_err"]), flush=True)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("d", [40, 160])
def test_filter_extreme_rows(d, form, msj_mod, oracle_mod):
    """Rows far outside the data (|a'|^2 > 2^100) fall in the band; a
    duplicate pair of them is a result.  d = 160 runs full tiles (the
    split form's BF16-first pass meets the NaN scores there)."""
    x = datagen.mixture(3000, d, 26)
    x[17] = 1e60
    x[2900] = 1e60
    x[5, :7] = -3e55
    x = _prep(x, msj_mod)
    eps = 0.5 if d == 40 else 1.0
    _announce("extreme rows", x, eps, form)
    got = msj_mod.self_join(x, eps, form=_form(msj_mod, form), count_work=1)
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)
    assert got.stats["filter_violations"] == 0


@pytest.mark.parametrize("form", FORMS)
def test_filter_offset_data(form, msj_mod, oracle_mod):
    """A large common offset: centring keeps the filter useful and exact."""
    x = datagen.correlated(5000, 64, 27) + 1000.0
    x = _prep(x, msj_mod)
    eps = 1.0
    _announce("offset data", x, eps, form)
    got = msj_mod.self_join(x, eps, form=_form(msj_mod, form))
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)


def test_filter_outside_range_runs_plain(msj_mod, oracle_mod):
    x = _prep(datagen.uniform(3000, 40, 28), msj_mod)
    eps = 1e-20  # eps^2 below the filter's range
    _announce("eps below filter range", x, eps, "split")
    got = msj_mod.self_join(x, eps, form=msj_mod.FORM_MFMA_SPLIT)
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)
    assert got.stats["form_used"] == msj_mod.FORM_PLAIN


def test_filter_context_reuse(msj_mod, oracle_mod):
    x = _prep(datagen.correlated(8000, 90, 29), msj_mod)
    _announce("filter context reuse", x, 0.8, "split")
    with msj_mod.Context(form=msj_mod.FORM_MFMA_SPLIT) as ctx:
        ctx.load(x)
        for eps in (0.8, 0.5, 0.8):
            got = ctx.join(eps)
            ref = oracle_mod.self_join(x, eps)
            _assert_same(got, ref)
        small = x[:1500].copy()
        ctx.load(small)
        got = ctx.join(0.8)
        _assert_same(got, oracle_mod.self_join(small, 0.8))
