"""The counting build of the plain kernels (count_work): dimensions consumed
by the early-exit test, summed over every evaluated unordered candidate
pair, checked against an independent count.  This sum is the algorithmic
work the bench's FP64 roofline uses."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs_dims(x, eps, ref):
    """Exact count: integer-valued data (every partial sum exact in FP64);
    candidate pairs from the oracle's node net, unordered, p <= c inside a
    node; a chunk of 4 dimensions is consumed while the pair is undecided."""
    n, d = x.shape
    xo = x[ref.node_order]
    ns = np.asarray(ref.node_start, dtype=np.int64)
    eps2 = eps * eps
    total = 0
    for X in range(len(ns) - 1):
        a0, a1 = ns[X], ns[X + 1]
        for Y in ref.adj[ref.adj_off[X]:ref.adj_off[X + 1]]:
            if Y < X:
                continue
            b0, b1 = ns[Y], ns[Y + 1]
            diff = xo[a0:a1, None, :] - xo[None, b0:b1, :]
            sq = diff * diff
            alive = np.ones(sq.shape[:2], dtype=bool)
            if Y == X:
                alive &= np.triu(np.ones_like(alive))
            s = np.zeros(sq.shape[:2])
            for l in range(0, d, 4):
                total += int(alive.sum()) * min(4, d - l)
                s = s + sq[:, :, l:l + 4].sum(axis=2)
                alive &= ~(s > eps2)
    return total


@pytest.mark.parametrize("d,scale,eps", [(8, 16, 0.35), (40, 8, 1.1), (18, 16, 0.3)])
def test_dims_count_exact(d, scale, eps, msj_mod, oracle_mod):
    rng = np.random.default_rng(d)
    x = rng.integers(0, scale, size=(1500, d)).astype(np.float64) / scale
    order, _ = msj_mod.dim_order(x)
    x = np.ascontiguousarray(x[:, order])
    print("\n[case] dims count d=%d" % d, flush=True)
    got = msj_mod.self_join(x, eps, count_work=1)
    ref = oracle_mod.self_join(x, eps)
    assert got.total_pairs == ref.total_pairs
    assert got.stats["dims_evaluated"] == _pairs_dims(x, eps, ref)


def test_dims_count_single_chunk(msj_mod):
    """d <= 4: one chunk per pair, so the count is d per unordered pair."""
    x = np.random.default_rng(3).random((5000, 3))
    got = msj_mod.self_join(x, 0.05, count_work=1)
    st = got.stats
    assert st["dims_evaluated"] == 3 * (st["num_calcs"] + x.shape[0]) // 2
