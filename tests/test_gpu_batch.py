"""GPU parity of the batching form (msj_batch_* / msj_join_batch): every
job's result equals the oracle's (and msj_self_join's), bit for bit, for
jobs of uneven sizes and row widths in one persistent launch; a second run
replays the captured graph with the tuned tile shapes."""
import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu


def _jobs(msj_mod):
    rng = np.random.default_rng(77)
    specs = [(1500, 2, 0.03), (4000, 8, 0.09), (9000, 18, 0.2), (2500, 57, 0.553),
             (3000, 90, 1.14), (700, 18, 0.25), (1, 8, 0.1), (12000, 8, 0.07),
             (2000, 3, 0.05), (5000, 90, 0.9)]
    datas, eps = [], []
    for k, (n, d, e) in enumerate(specs):
        x = datagen.uniform(n, d, 900 + k) if d <= 3 else (
            datagen.correlated(n, d, 900 + k) if d == 90 else datagen.mixture(n, d, 900 + k))
        order, _ = msj_mod.dim_order(x)
        datas.append(datagen.reorder_dims(x, order))
        eps.append(e * float(rng.uniform(0.9, 1.1)))
    return datas, eps


def _same(a, b):
    assert a.total_pairs == b.total_pairs
    assert np.array_equal(a.counts, b.counts)
    assert np.array_equal(a.offsets, b.offsets)
    assert np.array_equal(a.nbrs, b.nbrs)
    assert np.array_equal(a.node_order, b.node_order)


def test_batch_matches_oracle_and_single_joins(msj_mod, oracle_mod):
    datas, eps = _jobs(msj_mod)
    print("\n[case] batch of %d jobs" % len(datas), flush=True)
    with msj_mod.Batch() as b:
        b.load(datas, eps)
        res, st = b.run()
        print("[case] tuned classes %s epi_log2 %s grid %d" % (
            st["class_ds"], st["class_epi_log2"], st["grid_blocks"]), flush=True)
        assert st["njobs"] == len(datas)
        assert st["total_pairs"] == sum(r.total_pairs for r in res)
        for x, e, r in zip(datas, eps, res):
            ref = oracle_mod.self_join(x, e)
            assert r.total_pairs == ref.total_pairs
            assert np.array_equal(r.counts, ref.counts)
            assert np.array_equal(r.offsets, ref.offsets)
            assert np.array_equal(r.nbrs, ref.nbrs)
            assert np.array_equal(r.node_order, ref.node_order)
            assert r.stats["num_calcs"] == ref.num_calcs
        # second run: graph replay with the tuned shapes
        res2, st2 = b.run()
        assert st2["ms_tune"] == 0.0
        for r1, r2 in zip(res, res2):
            _same(r1, r2)
    for x, e, r in zip(datas[:4], eps[:4], res[:4]):
        _same(r, msj_mod.self_join(x, e))


def test_join_batch_one_shot(msj_mod, oracle_mod):
    datas, eps = _jobs(msj_mod)
    datas, eps = datas[:3], eps[:3]
    print("\n[case] one-shot batch of %d jobs" % len(datas), flush=True)
    res = msj_mod.join_batch(datas, eps)
    for x, e, r in zip(datas, eps, res):
        _same(r, oracle_mod.self_join(x, e))


def test_batch_reload_and_fp32(msj_mod, oracle_mod):
    x1 = datagen.clusters(2000, 384, seed=41, ncl=30)
    x2 = datagen.clusters(1200, 64, seed=42, ncl=10)
    datas = []
    for x in (x1, x2):
        order, _ = msj_mod.dim_order(x.astype(np.float64))
        datas.append(datagen.reorder_dims(x, order))
    eps = [1.26, 0.6]
    print("\n[case] batch fp32 storage", flush=True)
    with msj_mod.Batch(storage=msj_mod.STORAGE_F32) as b:
        b.load(datas, eps)
        res, _ = b.run()
        for x, e, r in zip(datas, eps, res):
            _same(r, oracle_mod.self_join(x.astype(np.float64), e))
        b.load(datas[1:], eps[1:])
        res, st = b.run()
        assert st["njobs"] == 1
        _same(res[0], oracle_mod.self_join(datas[1].astype(np.float64), eps[1]))


def test_batch_invalid_job(msj_mod):
    x = np.random.default_rng(3).random((100, 4))
    with pytest.raises(msj_mod.MsjError) as e:
        msj_mod.join_batch([x], [0.0])
    assert e.value.code == msj_mod.E_INVAL


def test_batch_solo_jobs_overlap_builds(msj_mod):
    """Jobs with enough candidate evaluations to fill the GPU are joined alone
    as soon as their node nets are built (while the next nets are built);
    the rest share the persistent launch.  Every job equals msj_self_join."""
    specs = [(400000, 18, 0.14, 951), (250000, 8, 0.05, 952), (3000, 57, 0.5, 953),
             (5000, 8, 0.09, 954), (800, 18, 0.25, 955)]
    datas, eps = [], []
    for n, d, e, seed in specs:
        x = datagen.mixture(n, d, seed)
        order, _ = msj_mod.dim_order(x)
        datas.append(datagen.reorder_dims(x, order))
        eps.append(e)
    print("\n[case] batch with solo jobs", flush=True)
    with msj_mod.Batch() as b:
        b.load(datas, eps)
        for rnd in range(2):
            res, st = b.run()
            assert st["solo_jobs"] >= 1 and st["solo_jobs"] < len(datas)
            # the tuned slab width of the wide-row class (class_ds 0: every
            # padded width above 32), 0 for the others
            for ds, sl in zip(st["class_ds"], st["class_slab"]):
                assert sl in ((8, 16) if ds == 0 else (0,))
            assert 0 in st["class_ds"]
            for x, e, r in zip(datas, eps, res):
                _same(r, msj_mod.self_join(x, e))


def test_batch_solo_jobs_in_filter_form(msj_mod):
    """With an MFMA form the solo jobs run the filter and its FP64 band
    recheck (as msj_self_join would); pooled jobs stay on the plain
    persistent kernel.  Every job still equals the plain single join."""
    datas, eps = [], []
    for x, e in [(datagen.correlated(100000, 90, 961), 0.77),
                 (datagen.mixture(3000, 57, 962), 0.5),
                 (datagen.mixture(5000, 8, 963), 0.09)]:
        order, _ = msj_mod.dim_order(x)
        datas.append(datagen.reorder_dims(x, order))
        eps.append(e)
    print("\n[case] batch with filter-form solo jobs", flush=True)
    with msj_mod.Batch(form=msj_mod.FORM_MFMA_SPLIT) as b:
        b.load(datas, eps)
        for rnd in range(2):
            res, st = b.run()
            assert st["solo_jobs"] == 1
            assert res[0].stats["form_used"] == msj_mod.FORM_MFMA_SPLIT
            assert res[1].stats["form_used"] == msj_mod.FORM_PLAIN
            for x, e, r in zip(datas, eps, res):
                _same(r, msj_mod.self_join(x, e))
