"""The tree/COSS candidate generator on the GPU (msj_params.index =
MSJ_INDEX_COSS) against the oracle's restatement of buildTree (reference
tree.cu:3-593 with MINSQRS) and its 3^L neighbour probes (tree.cu:886-996).

Per case, bit for bit: every candidate key row of one layer (RP keys from
the oracle's reference points, axis keys ceil((x + 1e-7) / eps) + 10), the
chosen candidate, its statistic's FP64 bits and the non-empty bin count of
every layer, the node order and boundaries, the nodes' bin-number tuples,
every node's neighbour set, numCalcs, and the join's neighbour table (which
must also equal brute force: the 3^L neighbourhoods are complete).
"""
import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu

CASES = [
    ("uniform 6000x2", lambda: datagen.uniform(6000, 2, 71), 0.01),
    ("mixture 6000x5", lambda: datagen.mixture(6000, 5, 72), 0.06),
    ("mixture 5000x18", lambda: datagen.mixture(5000, 18, 73), 0.2),
    ("correlated 3000x40", lambda: datagen.correlated(3000, 40, 74), 0.5),
    ("centred 4000x3 (wrapping axes)", lambda: datagen.uniform(4000, 3, 75) * 2.0 - 1.0, 0.02),
    ("uniform 60000x3", lambda: datagen.uniform(60000, 3, 77), 0.02),
]


def _prep(x, msj_mod):
    order, _ = msj_mod.dim_order(x)
    return datagen.reorder_dims(x, order)


@pytest.mark.parametrize("name,gen,eps", CASES, ids=[c[0] for c in CASES])
def test_coss_index_equals_oracle(name, gen, eps, msj_mod, oracle_mod):
    x = _prep(gen(), msj_mod)
    n, d = x.shape
    print("\n[case] coss index %s n=%d d=%d eps=%g" % (name, n, d, eps), flush=True)
    ref = oracle_mod.self_join(x, eps, join=0, index=2)
    pr = msj_mod.probe_index(x, eps, layer=1, adjacency=True, index=msj_mod.INDEX_COSS)
    assert pr["num_layers"] == ref.num_layers == 6
    assert pr["num_nodes"] == ref.num_nodes
    assert np.array_equal(pr["node_order"], ref.node_order)
    assert np.array_equal(pr["node_start"], ref.node_start)
    assert np.array_equal(pr["node_keys"], ref.node_keys)
    assert pr["num_calcs"] == ref.num_calcs
    for v in range(ref.num_nodes):
        mine = pr["adj"][pr["adj_off"][v]:pr["adj_off"][v + 1]]
        want = np.sort(ref.adj[ref.adj_off[v]:ref.adj_off[v + 1]])
        assert np.array_equal(mine, want), v
    # the keys of layer 1: RP rows from the oracle's points, then the axes
    rps = oracle_mod.rp_layers(x, layers=2)[1]
    naxis = min(6, d)
    want = np.concatenate([
        oracle_mod.keys(x, eps, rps, 0),
        np.stack([_axis_keys(x[:, a], eps) for a in range(naxis)])])
    assert np.array_equal(pr["keys"], want)


def _axis_keys(col, eps):
    q = np.ceil((col + 0.0000001) / eps) + 10.0
    return np.where(np.abs(q) < 2.0 ** 63, q, 0).astype(np.int64).astype(np.uint32)


@pytest.mark.parametrize("name,gen,eps", CASES, ids=[c[0] for c in CASES])
def test_coss_join_equals_oracle_and_brute_force(name, gen, eps, msj_mod, oracle_mod):
    x = _prep(gen(), msj_mod)
    n, d = x.shape
    print("\n[case] coss join %s n=%d d=%d eps=%g" % (name, n, d, eps), flush=True)
    ref = oracle_mod.self_join(x, eps, index=2)
    got = msj_mod.self_join(x, eps, index=msj_mod.INDEX_COSS)
    st = got.stats
    assert st["chosen"] == list(ref.chosen)
    assert [v.hex() for v in st["sse"]] == [v.hex() for v in ref.sse]
    assert st["nodes_per_layer"] == list(ref.nodes_per_layer)
    assert st["num_nodes"] == ref.num_nodes
    assert st["num_calcs"] == ref.num_calcs
    assert np.array_equal(got.node_order, ref.node_order)
    assert np.array_equal(got.counts, ref.counts)
    assert np.array_equal(got.nbrs, ref.nbrs)
    bf = oracle_mod.brute_force(x, eps)
    assert np.array_equal(got.counts, bf.counts) and np.array_equal(got.nbrs, bf.nbrs)


def test_coss_f32_storage(msj_mod, oracle_mod):
    x = _prep(datagen.mixture(8000, 6, 78), msj_mod).astype(np.float32)
    print("\n[case] coss f32 8000x6", flush=True)
    ref = oracle_mod.self_join(x.astype(np.float64), 0.08, index=2)
    got = msj_mod.self_join(x, 0.08, index=msj_mod.INDEX_COSS, storage=msj_mod.STORAGE_F32)
    assert got.stats["chosen"] == list(ref.chosen)
    assert np.array_equal(got.node_order, ref.node_order)
    assert np.array_equal(got.counts, ref.counts) and np.array_equal(got.nbrs, ref.nbrs)


def test_coss_layers_and_candidates(msj_mod, oracle_mod):
    """maxrp = 3 layers and rp_per_layer = 7 (fewer axes than dimensions)."""
    x = _prep(datagen.mixture(5000, 8, 79), msj_mod)
    print("\n[case] coss maxrp=3 rp=7 5000x8", flush=True)
    ref = oracle_mod.self_join(x, 0.15, index=2, maxrp=3, rp_per_layer=7)
    got = msj_mod.self_join(x, 0.15, index=msj_mod.INDEX_COSS, maxrp=3, rp_per_layer=7)
    assert got.stats["num_candidates"] == 7 + 3
    assert got.stats["chosen"] == list(ref.chosen)
    assert [v.hex() for v in got.stats["sse"]] == [v.hex() for v in ref.sse]
    assert np.array_equal(got.counts, ref.counts) and np.array_equal(got.nbrs, ref.nbrs)


def test_coss_no_fitting_candidate_is_an_error(msj_mod, oracle_mod):
    """Every candidate of layer 1 needs 2^32 bins or more: the oracle fails
    the build and so does the library (MSJ_E_OVERFLOW), without a join."""
    x = _prep(datagen.uniform(20000, 2, 76), msj_mod)
    print("\n[case] coss no candidate fits", flush=True)
    with pytest.raises(ValueError):
        oracle_mod.self_join(x, 4e-6, join=0, index=2)
    with pytest.raises(msj_mod.MsjError) as e:
        msj_mod.self_join(x, 4e-6, index=msj_mod.INDEX_COSS)
    assert e.value.code == msj_mod.E_OVERFLOW


def test_coss_statistic_past_2_53(msj_mod, oracle_mod):
    """Two dense clusters (80001 copies each) and eps = 2e-8: every layer has
    two non-empty bins among ~10^8, so the statistic, (#bins - 2) * 80001^2
    plus what it carries, passes 2^53 and the reference's sequential FP64
    sum rounds on every empty bin (millions off the exact sum).  The
    library's statistic, choice and nodes must still match bit for bit."""
    n = 160002
    x = np.zeros((n, 2))
    x[:n // 2] = [0.1, 0.15]
    x[n // 2:] = [0.9, 0.8]
    print("\n[case] coss statistic past 2^53 n=%d" % n, flush=True)
    ref = oracle_mod.self_join(x, 2e-8, join=0, index=2)
    assert max(ref.sse) > 2.0 ** 53
    pr = msj_mod.probe_index(x, 2e-8, index=msj_mod.INDEX_COSS)
    assert pr["chosen"] == list(ref.chosen)
    assert [v.hex() for v in pr["sse"]] == [v.hex() for v in ref.sse]
    assert pr["nodes_per_layer"] == list(ref.nodes_per_layer)
    assert np.array_equal(pr["node_order"], ref.node_order)
    assert pr["num_calcs"] == ref.num_calcs
