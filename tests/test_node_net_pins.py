"""Independent pins of the oracle's node-net decisions.

Reference: src/nodes.cu:5-356 (buildNodeNet), 358-449 (initNodes),
452-633 (splitNodes), 677-706 (updateNodeCalcs), 719-747 (updateNeighbors).

The node net is rebuilt here layer by layer in numpy, from the oracle's
reference points and keys only (both pinned in test_oracle.py against libc
rand(), numpy and exact arithmetic):

  * every layer's candidate partitions (32 reference points + min(6, d)
    axes) are formed from the previous layer's nodes in the reference's loop
    order (nodes.cu:168-199): node by node, children by ascending key; a node
    whose numCalcs is below MIN_CALCS_PER_NODE is kept whole (nodes.cu:476);
  * the selection statistic sum((size - N/#nodes)^2) is summed sequentially
    in that order (nodes.cu:194-199) and the winner is the first candidate
    with the smallest value (nodes.cu:204-220: strict `<` in the critical
    section; the first arrival among equal values, taken here as the lowest
    index);
  * adjacency is updateNeighbors' rule (children of adjacent parents linked
    when either stopped splitting or their keys differ by at most one, mod
    2^32; siblings when their keys differ by exactly one) and numCalcs =
    |X| * sum over X's neighbours (self included) of |Y|.

The oracle's chosen candidate, the bits of its statistic, node counts,
candidate counts, node order, node keys and adjacency must all equal this
restatement's.  Planted mistakes (argmax instead of argmin, `<=` in the stop
rule) are run through the same comparison and must be caught.
"""
import numpy as np
import pytest

import datagen

U32 = 0xFFFFFFFF


def _seq_sse(sizes, n):
    """nodes.cu:194-199: avg = N*1.0/#nodes; sum of (size - avg)^2 in node
    order, one rounding per operation (np.add.accumulate is sequential)."""
    avg = n * 1.0 / len(sizes)
    dev = sizes.astype(np.float64) - avg
    return float(np.add.accumulate(dev * dev)[-1])


def _children(parent, ckey, n):
    """Runs of (parent, key) in node order: returns (order of the rows, child
    parent, child key, child size).  Rows keep ascending id inside a child
    (thrust's sort_by_key is stable, nodes.cu:372,505)."""
    rows = np.arange(n, dtype=np.int64)
    order = np.lexsort((rows, ckey, parent))
    p, k = parent[order], ckey[order]
    start = np.ones(n, dtype=bool)
    start[1:] = (p[1:] != p[:-1]) | (k[1:] != k[:-1])
    idx = np.flatnonzero(start)
    sizes = np.diff(np.append(idx, n))
    return order, p[idx], k[idx], sizes


def numpy_node_net(x, eps, oracle_mod, min_calcs=100000, layers=6, plant=None):
    """Returns a dict of the per-layer decisions and the final net."""
    n, d = x.shape
    rps = oracle_mod.rp_layers(x, layers=layers)
    naxis = min(6, d)
    node_of = None          # row -> node id (node order)
    keys = None             # [M, l] key tuples
    A = None                # [M, M] adjacency (self included)
    sizes = None
    calcs = None
    out = {"chosen": [], "sse": [], "nodes": [], "calcs": []}
    for layer in range(layers):
        K = oracle_mod.keys(x, eps, rps[layer], naxis).astype(np.int64)  # [ncand, n]
        ncand = K.shape[0]
        if layer == 0:
            parent = np.zeros(n, dtype=np.int64)
            stop_row = np.zeros(n, dtype=bool)
        else:
            parent = node_of
            if plant == "stop_le":
                stop_node = calcs <= min_calcs
            else:
                stop_node = calcs < min_calcs  # nodes.cu:476
            stop_row = stop_node[parent]
        sse = []
        for t in range(ncand):
            ck = np.where(stop_row, U32, K[t])
            _, _, _, sz = _children(parent, ck, n)
            sse.append(_seq_sse(sz, n))
        sse = np.array(sse)
        if plant == "argmax":
            best = int(np.flatnonzero(sse == sse.max())[0])
        else:
            best = int(np.flatnonzero(sse == sse.min())[0])
        ck = np.where(stop_row, U32, K[best])
        order, cpar, ckey, csz = _children(parent, ck, n)
        Mn = len(csz)
        new_node_of = np.empty(n, dtype=np.int64)
        new_node_of[order] = np.repeat(np.arange(Mn), csz)
        if layer == 0:
            diff = np.abs(ckey[:, None] - ckey[None, :])
            An = diff <= 1
            new_keys = ckey[:, None]
            cstop = np.zeros(Mn, dtype=bool)
        else:
            cstop = stop_node[cpar]
            same = cpar[:, None] == cpar[None, :]
            wrap = (ckey[:, None] - ckey[None, :]) % (1 << 32)
            near = (wrap == 0) | (wrap == 1) | (wrap == U32)
            linked_other = A[np.ix_(cpar, cpar)] & (cstop[:, None] | cstop[None, :] | near)
            sib = np.abs(ckey[:, None] - ckey[None, :]) == 1
            An = np.where(same, sib, linked_other)
            np.fill_diagonal(An, True)
            new_keys = np.concatenate([keys[cpar], ckey[:, None]], axis=1)
        sizes = csz.astype(np.int64)
        calcs = sizes * (An.astype(np.int64) @ sizes)
        node_of, keys, A = new_node_of, new_keys, An
        out["chosen"].append(best)
        out["sse"].append(float(sse[best]))
        out["nodes"].append(Mn)
        out["calcs"].append(int(calcs.sum()))
    rows = np.arange(n, dtype=np.int64)
    order = np.lexsort((rows, node_of))
    out["node_order"] = order.astype(np.uint32)
    out["node_start"] = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    out["node_keys"] = keys.astype(np.uint32)
    out["adj"] = A
    out["num_calcs"] = int(calcs.sum())
    return out


def _compare(ref, mine):
    """Mismatching fields between the oracle's result and the restatement's."""
    bad = []
    L = ref.num_layers
    if list(ref.chosen) != mine["chosen"][:L]:
        bad.append("chosen")
    if [float(v).hex() for v in ref.sse] != [float(v).hex() for v in mine["sse"][:L]]:
        bad.append("sse")
    if list(ref.nodes_per_layer) != mine["nodes"][:L]:
        bad.append("nodes_per_layer")
    if [int(v) for v in ref.calcs_per_layer] != mine["calcs"][:L]:
        bad.append("calcs_per_layer")
    if not np.array_equal(ref.node_order, mine["node_order"]):
        bad.append("node_order")
    if not np.array_equal(ref.node_start, mine["node_start"]):
        bad.append("node_start")
    if ref.num_calcs != mine["num_calcs"]:
        bad.append("num_calcs")
    if ref.node_keys.shape != mine["node_keys"].shape or \
            not np.array_equal(ref.node_keys, mine["node_keys"]):
        bad.append("node_keys")
    else:
        M = ref.num_nodes
        A = mine["adj"]
        for v in range(M):
            got = set(ref.adj[ref.adj_off[v]:ref.adj_off[v + 1]].tolist())
            if got != set(np.flatnonzero(A[v]).tolist()):
                bad.append("adjacency")
                break
    return bad


def _prep(x, oracle_mod):
    order, _ = oracle_mod.dim_order(x)
    return datagen.reorder_dims(x, order)


# sizes chosen so nodes keep splitting over several layers and stop at
# different depths (config 1's shape gives 100 strips that never split again:
# re-using the same axis leaves every strip whole and has the smallest SSE)
CASES = [
    ("uniform 65536x2 (config 1)", lambda: datagen.uniform(65536, 2, 1001), 0.01),
    ("uniform d=3", lambda: datagen.uniform(40000, 3, 62), 0.03),
    ("mixture d=5", lambda: datagen.mixture(40000, 5, 63), 0.05),
    ("mixture d=8", lambda: datagen.mixture(40000, 8, 64), 0.08),
    ("mixture d=18", lambda: datagen.mixture(30000, 18, 65), 0.2),
    ("correlated d=40", lambda: datagen.correlated(20000, 40, 66), 0.3),
    ("correlated d=90", lambda: datagen.correlated(20000, 90, 66), 0.5),
]


@pytest.mark.parametrize("name,gen,eps", CASES, ids=[c[0] for c in CASES])
def test_oracle_node_net_equals_numpy_restatement(name, gen, eps, oracle_mod):
    x = _prep(gen(), oracle_mod)
    ref = oracle_mod.self_join(x, eps, join=0)
    mine = numpy_node_net(x, eps, oracle_mod, layers=ref.num_layers)
    assert ref.num_layers == 6
    assert _compare(ref, mine) == [], name


def test_selection_is_the_minimum_not_the_maximum(oracle_mod):
    """A planted argmax (nodes.cu:207 reads `stdDev < lowestStdDev`) must be
    caught by the comparison on every case."""
    for name, gen, eps in CASES[2:5]:
        x = _prep(gen(), oracle_mod)
        ref = oracle_mod.self_join(x, eps, join=0)
        planted = numpy_node_net(x, eps, oracle_mod, plant="argmax")
        assert "chosen" in _compare(ref, planted), name


def stop_rule_input():
    """Four clusters of identical rows on one axis, inside [2, 3] so every
    reference point lies on one side of all of them: every candidate gives the
    same partition {A, B, C, D}.  A-B and C-D are one eps apart (adjacent
    bins), the pairs far apart, so after layer 0
      numCalcs(A) = 200 * (200 + 300) = 100,000   (not below 1e5: splits)
      numCalcs(C) = 123 * (123 + 690) =  99,999   (below 1e5: stops)
    The first rows (the 1% reference-point sample) hold D, the maximum."""
    eps = 2.0 ** -5
    pos = {"A": 2.25, "B": 2.25 + eps, "C": 2.5, "D": 2.5 + eps}
    cnt = {"A": 200, "B": 300, "C": 123, "D": 690}
    rows = [np.full(cnt[k], pos[k]) for k in ("D", "A", "B", "C")]
    return np.concatenate(rows)[:, None].copy(), eps, pos


def test_stop_rule_boundary_at_min_calcs(oracle_mod):
    """nodes.cu:476: a node with numCalcs < MIN_CALCS_PER_NODE (1e5) is copied
    unsplit (key UINT_MAX from then on); numCalcs == 1e5 still splits."""
    x, eps, pos = stop_rule_input()
    ref = oracle_mod.self_join(x, eps, join=0)
    assert ref.nodes_per_layer[0] == 4
    mine = numpy_node_net(x, eps, oracle_mod)
    assert _compare(ref, mine) == []
    # node of each cluster, by its rows' coordinate
    start = ref.node_start.astype(np.int64)
    node_pos = [float(x[ref.node_order[start[v]], 0]) for v in range(ref.num_nodes)]
    size = np.diff(start)
    calcs0 = {}
    for v in range(ref.num_nodes):
        nb = ref.adj[ref.adj_off[v]:ref.adj_off[v + 1]]
        calcs0[node_pos[v]] = int(size[v]) * int(size[nb].sum())
    assert calcs0[pos["A"]] == 100000 and calcs0[pos["C"]] == 99999
    keys = {node_pos[v]: ref.node_keys[v] for v in range(ref.num_nodes)}
    assert np.all(keys[pos["A"]] != U32)        # 100,000: keeps splitting
    assert np.all(keys[pos["C"]][1:] == U32)    # 99,999: stopped after layer 0
    assert keys[pos["C"]][0] != U32
    # a planted `<=` would stop A too: the comparison catches it
    planted = numpy_node_net(x, eps, oracle_mod, plant="stop_le")
    assert "node_keys" in _compare(ref, planted)


@pytest.mark.parametrize("offset", [0, 1])
def test_stop_rule_threshold_parameter(offset, oracle_mod):
    """The same boundary through the min_calcs_per_node knob on generic data:
    with the threshold equal to a node's numCalcs the node splits, one above
    it stops."""
    x = _prep(datagen.mixture(6000, 8, 67), oracle_mod)
    eps = 0.08
    base = oracle_mod.self_join(x, eps, join=0, maxrp=1)
    start = base.node_start.astype(np.int64)
    size = np.diff(start)
    calcs = np.array([int(size[v]) * int(size[base.adj[base.adj_off[v]:base.adj_off[v + 1]]].sum())
                      for v in range(base.num_nodes)])
    big = int(np.argmax(calcs))
    thr = int(calcs[big]) + offset
    ref = oracle_mod.self_join(x, eps, join=0, maxrp=2, min_calcs_per_node=thr)
    mine = numpy_node_net(x, eps, oracle_mod, min_calcs=thr, layers=2)
    assert _compare(ref, mine) == []
    # locate the layer-0 node `big` in the 2-layer net by its layer-0 key
    k0 = base.node_keys[big, 0]
    rows = ref.node_keys[:, 0] == k0
    second = ref.node_keys[rows, 1]
    if offset == 0:
        assert np.all(second != U32)
    else:
        assert np.all(second == U32) and rows.sum() == 1


# ------------------------------------------------------------ tree / COSS


def _x86_u32(v):
    """(unsigned) of a double on x86-64: cvttsd2si to 64 bits, low half;
    out of range -> 0x8000000000000000 (low half 0)."""
    v = np.asarray(v, dtype=np.float64)
    ok = (v > -9.223372036854775808e18) & (v < 9.223372036854775808e18)
    out = np.zeros(v.shape, dtype=np.int64)
    out[ok] = np.trunc(v[ok]).astype(np.int64)
    return (out & 0xFFFFFFFF).astype(np.int64)


def _x86_i32(v):
    """(int) of a double on x86-64: cvttsd2si to 32 bits; out of range ->
    INT_MIN."""
    if not (-2147483649.0 < v < 2147483648.0):
        return -(1 << 31)
    return int(np.trunc(v))


def numpy_coss(x, eps, oracle_mod, layers=6, plant=None):
    """buildTree (tree.cu:3-593, MINSQRS) and the 3^L probes (tree.cu:886-996)
    restated in numpy from the oracle's reference points and RP keys."""
    n, d = x.shape
    rps = oracle_mod.rp_layers(x, layers=layers)
    nrp = rps.shape[1]
    naxis = min(layers, d)
    ncand = nrp + naxis
    sumsq = np.zeros(ncand)
    prank = np.ones(n, dtype=np.int64)
    prev_nonempty = 0
    bnum = np.zeros((n, layers), dtype=np.int64)
    out = {"chosen": [], "sse": [], "nodes": []}
    M32 = 0xFFFFFFFF
    for cl in range(layers):
        K = oracle_mod.keys(x, eps, rps[cl], 0).astype(np.int64)  # RP keys only
        ax = np.stack([_x86_u32(np.ceil((x[:, a] + 0.0000001) / eps) + 10)
                       for a in range(naxis)]) if naxis else np.zeros((0, n), np.int64)
        K = np.concatenate([K, ax])
        sumsq[:nrp] = 0.0
        score = np.full(ncand, np.inf)
        cand = []
        for t in range(ncand):
            k = K[t]
            if t < nrp:
                maxb, minb = int(k.max()), int(k.min())
            else:
                a = t - nrp
                maxbi = _x86_i32(np.ceil((x[0, a] + 0.000001) / eps) + 10) & M32
                minbi = (maxbi - 1) & M32
                maxb, minb = max(maxbi, int(k.max())), min(minbi, int(k.min()))
            skip = (minb - 2) & M32
            nb = (maxb + 2 - skip) & M32
            count = nb if cl == 0 else prev_nonempty * nb
            cand.append((skip, nb, count))
            if count == 0 or count > M32 or (maxb - minb) + 4 > M32:
                continue
            bn = (k - skip) & M32
            addr = bn if cl == 0 else ((prank - 1) * nb) % (1 << 32) + bn
            bins = np.bincount(addr, minlength=count).astype(np.float64)
            nonempty = int((bins > 0).sum())
            avg = float(n // nonempty)
            if t >= nrp and avg >= n / 3.0:
                sumsq[t] = float(n * n)
            e = bins - avg
            acc = np.add.accumulate(np.concatenate([[sumsq[t]], e * e]))
            sumsq[t] = float(acc[-1])
            score[t] = sumsq[t]
        best = int(np.flatnonzero(score == score.min())[0]) if plant != "argmax" else \
            int(np.flatnonzero(score == score[np.isfinite(score)].max())[0])
        skip, nb, count = cand[best]
        k = K[best]
        bn = (k - skip) & M32
        addr = bn if cl == 0 else ((prank - 1) * nb) % (1 << 32) + bn
        uniq = np.unique(addr)
        prank = np.searchsorted(uniq, addr) + 1
        bnum[:, cl] = bn
        prev_nonempty = len(uniq)
        out["chosen"].append(best)
        out["sse"].append(float(score[best]))
        out["nodes"].append(len(uniq))
    order = np.lexsort((np.arange(n), prank))
    tuples = bnum[order]
    starts = np.concatenate([[0], np.flatnonzero(np.any(tuples[1:] != tuples[:-1], axis=1)) + 1])
    keys = tuples[starts]
    sizes = np.diff(np.concatenate([starts, [n]]))
    index = {tuple(t): v for v, t in enumerate(keys.tolist())}
    adj = []
    import itertools
    for v, t in enumerate(keys.tolist()):
        nb_set = set()
        for delta in itertools.product((-1, 0, 1), repeat=layers):
            w = index.get(tuple(a + b for a, b in zip(t, delta)))
            if w is not None:
                nb_set.add(w)
        adj.append(nb_set)
    calcs = sum(int(sizes[v]) * sum(int(sizes[w]) for w in adj[v]) for v in range(len(keys)))
    out.update(node_order=order.astype(np.uint32), node_keys=keys.astype(np.uint32),
               node_start=np.concatenate([starts, [n]]).astype(np.uint32), adj=adj,
               num_calcs=calcs)
    return out


COSS_CASES = [
    ("uniform d=2", lambda: datagen.uniform(6000, 2, 71), 0.01),
    ("mixture d=5", lambda: datagen.mixture(6000, 5, 72), 0.06),
    ("mixture d=18", lambda: datagen.mixture(5000, 18, 73), 0.2),
    ("correlated d=40", lambda: datagen.correlated(3000, 40, 74), 0.5),
    # axis keys below zero wrap the unsigned bin span: those axes are not eligible
    ("centred d=3", lambda: datagen.uniform(4000, 3, 75) * 2.0 - 1.0, 0.02),
]


@pytest.mark.parametrize("name,gen,eps", COSS_CASES, ids=[c[0] for c in COSS_CASES])
def test_oracle_coss_equals_numpy_restatement(name, gen, eps, oracle_mod):
    x = _prep(gen(), oracle_mod)
    ref = oracle_mod.self_join(x, eps, join=0, index=2)
    mine = numpy_coss(x, eps, oracle_mod)
    assert list(ref.chosen) == mine["chosen"]
    assert [v.hex() for v in ref.sse] == [float(v).hex() for v in mine["sse"]]
    assert list(ref.nodes_per_layer) == mine["nodes"]
    assert np.array_equal(ref.node_order, mine["node_order"])
    assert np.array_equal(ref.node_start, mine["node_start"])
    assert np.array_equal(ref.node_keys, mine["node_keys"])
    assert ref.num_calcs == mine["num_calcs"]
    for v in range(ref.num_nodes):
        assert set(ref.adj[ref.adj_off[v]:ref.adj_off[v + 1]].tolist()) == mine["adj"][v]
    # the selection is the smallest statistic: a planted argmax is caught
    planted = numpy_coss(x, eps, oracle_mod, plant="argmax")
    assert planted["chosen"] != list(ref.chosen)


@pytest.mark.parametrize("name,gen,eps", COSS_CASES, ids=[c[0] for c in COSS_CASES])
def test_coss_join_is_complete(name, gen, eps, oracle_mod):
    """Every pair within eps lies in adjacent bins: the join over the tree's
    3^L neighbourhoods equals brute force."""
    x = _prep(gen(), oracle_mod)
    got = oracle_mod.self_join(x, eps, index=2)
    bf = oracle_mod.brute_force(x, eps)
    assert np.array_equal(got.counts, bf.counts)
    assert np.array_equal(got.nbrs, bf.nbrs)


def test_coss_fails_when_no_candidate_fits(oracle_mod):
    """A layer whose every candidate needs 2^32 bins or more (the reference's
    unsigned bin count would wrap) fails the build instead of guessing."""
    x = _prep(datagen.uniform(20000, 2, 76), oracle_mod)
    with pytest.raises(ValueError):
        oracle_mod.self_join(x, 4e-6, join=0, index=2)
