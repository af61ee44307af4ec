"""GPU parity: libmsj.so (through its C ABI) against the CPU oracle.

Bit-exact bar: counts, CSR offsets, neighbour lists, node order and the
node-net statistics (chosen candidate per layer, node counts, candidate
evaluation counts) must be identical.
"""
import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu


def _announce(name, x, eps):
    print("\n[case] %s n=%d d=%d eps=%r" % (name, x.shape[0], x.shape[1], eps), flush=True)


def _assert_same(got, ref):
    assert got.total_pairs == ref.total_pairs
    assert np.array_equal(got.counts, ref.counts)
    assert np.array_equal(got.offsets, ref.offsets)
    assert np.array_equal(got.nbrs, ref.nbrs)
    assert np.array_equal(got.node_order, ref.node_order)
    st = got.stats
    assert st["num_nodes"] == ref.num_nodes
    assert st["num_calcs"] == ref.num_calcs
    assert st["num_layers"] == ref.num_layers
    assert st["chosen"] == ref.chosen
    assert st["nodes_per_layer"] == ref.nodes_per_layer
    assert st["calcs_per_layer"] == ref.calcs_per_layer
    assert st["sse"] == ref.sse  # bit-identical selection statistic


def _prep(x, msj_mod):
    order, _ = msj_mod.dim_order(x)
    return datagen.reorder_dims(x, order)


@pytest.mark.parametrize("n,d,eps,seed", [
    (4096, 2, 0.02, 11),
    (65536, 2, 0.01, 1001),  # config 1, full size
    (3000, 3, 0.05, 12),
    (20000, 18, 0.204, 13),  # mean 28 neighbours
    (20000, 18, 0.12, 13),
    (8000, 8, 0.0879, 14),   # mean 15 neighbours
])
def test_uniform_and_mixture(n, d, eps, seed, msj_mod, oracle_mod):
    x = datagen.uniform(n, d, seed) if d <= 3 else datagen.mixture(n, d, seed)
    x = _prep(x, msj_mod)
    _announce("plain n=%d d=%d" % (n, d), x, eps)
    got = msj_mod.self_join(x, eps)
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)


@pytest.mark.parametrize("n,d,eps,seed", [
    (6000, 57, 0.553, 21),  # mean 20 neighbours
    (6000, 57, 0.3, 21),
    (5000, 90, 1.14, 22),   # mean 17 neighbours
    (5000, 90, 0.4, 22),
])
def test_high_dim_plain(n, d, eps, seed, msj_mod, oracle_mod):
    x = _prep(datagen.correlated(n, d, seed) if d == 90 else datagen.mixture(n, d, seed), msj_mod)
    _announce("plain high-d", x, eps)
    got = msj_mod.self_join(x, eps)
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)


def test_fp32_storage_384(msj_mod, oracle_mod):
    x = datagen.clusters(3000, 384, seed=31, ncl=40)
    order, _ = msj_mod.dim_order(x.astype(np.float64))
    x = datagen.reorder_dims(x, order)
    eps = 1.26  # mean 20 neighbours
    _announce("fp32 d=384", x, eps)
    got = msj_mod.self_join(x, eps, storage=msj_mod.STORAGE_F32)
    ref = oracle_mod.self_join(x.astype(np.float64), eps)
    _assert_same(got, ref)


@pytest.mark.parametrize("name,x,eps", [
    ("one point", np.array([[0.25, 0.5]]), 0.1),
    ("two far points", np.array([[0.0, 0.0], [1.0, 1.0]]), 0.1),
    ("duplicates", np.full((500, 3), 0.5), 0.01),
    ("d=1", np.random.default_rng(5).random((2000, 1)), 0.003),
    ("under 100 rows (empty RP sample)", np.random.default_rng(6).random((77, 4)), 0.3),
    ("ragged cluster sizes", np.concatenate([np.full((300, 2), 0.1),
                                             np.random.default_rng(7).random((50, 2))]), 0.05),
])
def test_edge_cases(name, x, eps, msj_mod, oracle_mod):
    x = np.ascontiguousarray(x, dtype=np.float64)
    _announce(name, x, eps)
    got = msj_mod.self_join(x, eps)
    ref = oracle_mod.self_join(x, eps)
    _assert_same(got, ref)


def test_context_reuse_two_eps(msj_mod, oracle_mod):
    x = _prep(datagen.mixture(12000, 18, seed=41), msj_mod)
    _announce("context reuse", x, 0.1)
    with msj_mod.Context() as ctx:
        ctx.load(x)
        for eps in (0.1, 0.06, 0.1):
            got = ctx.join(eps)
            ref = oracle_mod.self_join(x, eps)
            _assert_same(got, ref)


@pytest.mark.parametrize("gen,n,d,eps,seed", [
    ("uniform", 65536, 2, 0.01, 1001),
    ("mixture", 20000, 18, 0.12, 13),
    ("correlated", 5000, 90, 1.14, 22),
])
def test_node_sizes_dense_and_sort(gen, n, d, eps, seed, msj_mod, monkeypatch):
    """Candidate node sizes from the dense (parent, key) bins and from the
    composite-key sort (MSJ_NODE_SSE=sort) give the same choices, SSE bits,
    node counts and node order in every layer."""
    x = _prep(getattr(datagen, gen)(n, d, seed), msj_mod)
    _announce("dense vs sort node sizes, " + gen, x, eps)
    a = msj_mod.self_join(x, eps)
    monkeypatch.setenv("MSJ_NODE_SSE", "sort")
    b = msj_mod.self_join(x, eps)
    assert a.total_pairs == b.total_pairs
    assert np.array_equal(a.node_order, b.node_order)
    assert np.array_equal(a.offsets, b.offsets)
    assert np.array_equal(a.nbrs, b.nbrs)
    for k in ("num_nodes", "num_calcs", "num_layers", "chosen", "nodes_per_layer",
              "calcs_per_layer", "sse"):
        assert a.stats[k] == b.stats[k], k


def test_empty_input(msj_mod):
    x = np.zeros((0, 5))
    _announce("empty", x, 0.1)
    got = msj_mod.self_join(x, 0.1)
    assert got.total_pairs == 0
    assert got.offsets.tolist() == [0]


def test_invalid_arguments(msj_mod):
    x = np.random.default_rng(1).random((100, 2))
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(msj_mod.MsjError) as e:
            msj_mod.self_join(x, bad)
        assert e.value.code == msj_mod.E_INVAL
