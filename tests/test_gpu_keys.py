"""Direct checks of the GPU index build's keys (reference kernel.cu:880-909
binningKernel, nodes.cu:128-135 axis partitions) and node key tuples
(nodes.cuh:9-38 binNumbers), through msj_probe_index / msj_probe_key_math.

  * every candidate key row of every layer equals the oracle's keys from the
    oracle's reference points, bit for bit, on all points;
  * the final nodes' key tuples equal the oracle's (stopped layers included),
    and so do every node's neighbour set, numCalcs, the chosen candidates
    and the bits of their statistic;
  * the device FP64 sqrt and division behind each key equal IEEE (numpy's,
    correctly rounded) on 1e8 inputs, many of them within a few ulps of a
    bin edge.
"""
import numpy as np
import pytest

import datagen
from test_node_net_pins import stop_rule_input

pytestmark = pytest.mark.gpu


def _announce(name, n, d):
    print("\n[case] %s n=%d d=%d" % (name, n, d), flush=True)


@pytest.mark.parametrize("name,n,d,eps", [
    ("uniform 65536x2 (config 1)", 65536, 2, 0.01),
    ("mixture 200000x18 (config 2 shape)", 200000, 18, 0.138),
    ("correlated 20000x90", 20000, 90, 0.5),
])
def test_key_rows_every_layer(name, n, d, eps, msj_mod, oracle_mod):
    if d == 2:
        x = datagen.uniform(n, d, 1001)
    elif d == 18:
        x = datagen.mixture(n, d, 1002)
    else:
        x = datagen.correlated(n, d, 66)
    order, _ = msj_mod.dim_order(x)
    x = datagen.reorder_dims(x, order)
    _announce(name, n, d)
    rps = oracle_mod.rp_layers(x, layers=6)
    naxis = min(6, d)
    for layer in range(6):
        got = msj_mod.probe_index(x, eps, layer=layer)["keys"]
        want = oracle_mod.keys(x, eps, rps[layer], naxis)
        assert got.shape == want.shape
        bad = np.argwhere(got != want)
        assert bad.size == 0, (layer, bad[:5].tolist())


@pytest.mark.parametrize("name,make,eps", [
    ("mixture 40000x8", lambda: datagen.mixture(40000, 8, 64), 0.08),
    ("mixture 30000x18", lambda: datagen.mixture(30000, 18, 65), 0.2),
    ("stop rule at 1e5", lambda: stop_rule_input()[0], 2.0 ** -5),
])
def test_node_key_tuples(name, make, eps, msj_mod, oracle_mod):
    x = make()
    if x.shape[1] > 1:
        order, _ = msj_mod.dim_order(x)
        x = datagen.reorder_dims(x, order)
    _announce(name, *x.shape)
    got = msj_mod.probe_index(x, eps, adjacency=True)
    ref = oracle_mod.self_join(x, eps, join=0)
    assert got["num_nodes"] == ref.num_nodes
    assert np.array_equal(got["node_start"], ref.node_start)
    assert np.array_equal(got["node_order"], ref.node_order)
    assert np.array_equal(got["node_keys"], ref.node_keys)
    # neighbour-node sets (splitNodes + updateNeighbors, nodes.cu:452-633, 719-747)
    # and numCalcs
    assert got["num_calcs"] == ref.num_calcs
    assert got["num_edges"] == len(ref.adj)
    for v in range(ref.num_nodes):
        mine = np.sort(got["adj"][got["adj_off"][v]:got["adj_off"][v + 1]])
        assert np.array_equal(mine, np.sort(ref.adj[ref.adj_off[v]:ref.adj_off[v + 1]])), v
    assert got["chosen"] == list(ref.chosen)
    assert [v.hex() for v in got["sse"]] == [v.hex() for v in ref.sse]


def _host_keys(D, eps):
    q = np.sqrt(D) / eps
    f = np.floor(q)
    with np.errstate(invalid="ignore"):
        key = np.where(f > 0, np.where(f >= 4294967295.0, 4294967295.0, f), 0.0)
    key = np.nan_to_num(key, nan=0.0).astype(np.uint32)
    return q, key


def _edge_inputs(eps, rng, m):
    """Squared distances within a few ulps of (k eps)^2: keys change there."""
    k = rng.integers(0, 1 << 22, size=m).astype(np.float64)
    D = (k * eps) ** 2
    steps = rng.integers(-4, 5, size=m)
    bits = D.view(np.int64) + steps
    out = bits.view(np.float64)
    return np.where(out >= 0, out, 0.0)


def test_device_sqrt_and_division_are_ieee(msj_mod):
    rng = np.random.default_rng(77)
    total = 0
    chunk = 10_000_000
    eps_list = [0.01, 0.7707833306231173, 2.0 ** -5, 1e-3, 0.138, 3.0]
    specials = np.array([0.0, -0.0, 5e-324, 1e-310, 2.2250738585072014e-308, 1e300,
                         np.inf, np.nan, 1.0, 4.0, 0.25, 1.7976931348623157e308])
    for c in range(10):
        eps = eps_list[c % len(eps_list)]
        half = chunk // 2
        a = rng.random(half) * 10.0 ** rng.integers(-12, 12, size=half)
        b = _edge_inputs(eps, rng, chunk - half - specials.size)
        D = np.concatenate([specials, a, b])
        print("[case] key math chunk %d eps=%r m=%d" % (c, eps, D.size), flush=True)
        q, key = msj_mod.probe_key_math(D, eps)
        hq, hkey = _host_keys(D, eps)
        same_q = (q.view(np.uint64) == hq.view(np.uint64)) | (np.isnan(q) & np.isnan(hq))
        assert same_q.all(), D[~same_q][:5]
        assert np.array_equal(key, hkey), D[key != hkey][:5]
        total += D.size
    assert total >= 100_000_000
