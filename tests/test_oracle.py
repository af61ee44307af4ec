"""Pins of the CPU oracle (oracle/msj_oracle.c) against facts that do not
come from the oracle itself: the C library's own rand(), numpy, exact rational
arithmetic, closed-form lattice counts, brute force, and an independent
characterisation of the node-net adjacency.  The reference ships no tests or
fixtures for this path, so every pin is built here.
"""
import ctypes
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import datagen

# ----------------------------------------------------------------- rand()


@pytest.mark.parametrize("seed", [1, 2, 12345, 0x7FFFFFFF])
def test_glibc_rand_matches_libc(seed, oracle_mod):
    """utils.cu:77 uses rand(); the oracle's TYPE_3 restatement must reproduce
    this platform's glibc stream bit for bit."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    want = np.array([libc.rand() for _ in range(5000)], dtype=np.int32)
    assert np.array_equal(oracle_mod.glibc_rand(seed, 5000), want)


# ------------------------------------------------------------ dim order


def test_dim_order_matches_numpy_and_tie_rule(oracle_mod):
    """utils.cu:22-53: population sigma, descending; equal sigma -> the higher
    column index first (stable ascending sort, then reversed)."""
    rng = np.random.default_rng(3)
    a = rng.random((4000, 6)) * np.array([1.0, 3.0, 0.5, 2.0, 0.1, 1.5])
    a[:, 5] = a[:, 1]  # exact tie between columns 1 and 5
    order, sigma = oracle_mod.dim_order(a)
    ref = np.std(a, axis=0)  # ddof=0
    assert np.allclose(sigma, ref[order], rtol=1e-12, atol=0)
    assert list(order) == [5, 1, 3, 0, 2, 4]


# --------------------------------------------------------- reference points


def _box_coeff(i, j, d, ntest=64):
    """(i*step / (int)2^j) % 2 with x86-64 conversions, written out:
    step = (unsigned)(2^d / ntest) keeps the low 32 bits of the int64 value;
    (int)2^j is INT_MIN (0x80000000 as unsigned) for j >= 31."""
    v = 2.0 ** d / ntest
    step = (int(v) & 0xFFFFFFFF) if v < 2.0 ** 63 else 0
    div = (1 << j) if j <= 30 else 0x80000000
    return ((i * step) & 0xFFFFFFFF) // div % 2


@pytest.mark.parametrize("n,d", [(3000, 2), (5000, 18), (2000, 40)])
def test_reference_points_structure(n, d, oracle_mod):
    """utils.cu:66-137 per layer: 16 of 64 rand()-drawn candidates with the
    largest sigma of sample distances, then the sample's max corner and 15
    corner patterns; each layer consumes 64*d rand() values."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    x = np.random.default_rng(n + d).random((n, d))
    rps = oracle_mod.rp_layers(x, layers=2)
    S = int(n * 0.01)
    sample = x[:S]
    maxc = np.maximum(0.0, sample.max(axis=0))
    for layer in range(2):
        cand = np.array([libc.rand() for _ in range(64 * d)], dtype=np.float64) / 2147483647.0
        cand = cand.reshape(64, d)
        dist = np.sqrt(((sample[:, None, :] - cand[None, :, :]) ** 2).sum(axis=2))
        sig = dist.std(axis=0)
        top = np.argsort(-sig, kind="stable")[:16]
        assert np.allclose(rps[layer, :16], cand[top], rtol=0, atol=0)
        assert np.array_equal(rps[layer, 16], maxc)
        for i in range(1, 16):
            coeff = np.array([_box_coeff(i, j, d) for j in range(d)], dtype=np.float64)
            assert np.array_equal(rps[layer, 16 + i], coeff * maxc)


# ------------------------------------------------------------------- keys


def test_keys_on_dyadic_grid(oracle_mod):
    """kernel.cu:895-905 / nodes.cu:130 on inputs where every FP64 operation is
    exact, so numpy's differently ordered sums must agree bit for bit."""
    rng = np.random.default_rng(9)
    x = rng.integers(0, 1024, size=(500, 5)).astype(np.float64) / 1024.0
    rp = rng.integers(0, 1024, size=(32, 5)).astype(np.float64) / 1024.0
    eps = 1.0 / 64.0
    k = oracle_mod.keys(x, eps, rp, naxis=5)
    D = ((x[None, :, :] - rp[:, None, :]) ** 2).sum(axis=2)
    assert np.array_equal(k[:32], np.floor(np.sqrt(D) / eps).astype(np.uint32))
    ax = np.ceil((x.T + 0.0000001) / eps).astype(np.uint32)
    assert np.array_equal(k[32:], ax)


def test_key_conversions(oracle_mod):
    """Saturating device conversion for RP keys; x86-64 low-32-bit conversion
    for axis keys (cvttsd2si to 64 bits)."""
    eps = 1e-9
    x = np.array([[0.0], [5.0]])
    rp = np.array([[0.0]])
    k = oracle_mod.keys(x, eps, rp, naxis=1)
    assert k[0, 0] == 0 and k[0, 1] == 0xFFFFFFFF  # 5e9 saturates
    q = math.ceil((5.0 + 0.0000001) / eps)
    assert k[1, 1] == (q & 0xFFFFFFFF)


# ------------------------------------------------------ distance sequences


def _rn_fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _emulate(p, q, seq):
    s = 0.0
    d = len(p)
    for l in range(0, d, 4):
        t = p[l] - q[l]
        s = s + t * t if seq == 2 else _rn_fma(t, t, s)
        for j in range(1, 4):
            if l + j < d:
                u = p[l + j] - q[l + j]
                s = _rn_fma(u, u, s) if seq == 1 else s + u * u
    return s


def test_distance_sequences_exact(oracle_mod):
    """kernel.cu:833-865 under the three contraction readings, emulated with
    correctly rounded rational arithmetic; the three must also differ."""
    rng = np.random.default_rng(17)
    differ = set()
    for trial in range(300):
        d = int(rng.integers(1, 23))
        p = rng.random(d)
        q = rng.random(d)
        got = [oracle_mod.seq_sum(p, q, s) for s in range(3)]
        for s in range(3):
            assert got[s] == _emulate(list(p), list(q), s), (trial, s)
        differ.add((got[0] != got[1], got[0] != got[2]))
    assert (True, True) in differ or ((True, False) in differ and (False, True) in differ)


def test_brute_force_decision_vs_exact_distance(oracle_mod):
    rng = np.random.default_rng(23)
    x = rng.random((300, 7))
    eps = 0.35
    b = oracle_mod.brute_force(x, eps)
    got = set(map(tuple, b.pairs().tolist()))
    e2 = Fraction(eps) ** 2
    for i in range(len(x)):
        for j in range(len(x)):
            exact = sum((Fraction(a) - Fraction(c)) ** 2 for a, c in zip(x[i], x[j]))
            if abs(exact - e2) > Fraction(1, 10 ** 12):
                assert ((i, j) in got) == (exact <= e2)


# ------------------------------------------------------ closed-form lattice


@pytest.mark.parametrize("m,d,r", [(40, 2, 3), (12, 3, 2), (7, 4, 2)])
def test_lattice_closed_form(m, d, r, oracle_mod):
    """Points on an m^d grid of spacing 2^-5: every distance is exact, so the
    neighbour count of each point is the number of integer offsets o with
    |o|^2 <= r^2 that stay on the grid."""
    h = 2.0 ** -5
    pts = np.array(list(itertools.product(range(m), repeat=d)), dtype=np.float64)
    x = pts * h
    eps = r * h
    offs = [o for o in itertools.product(range(-r, r + 1), repeat=d)
            if sum(v * v for v in o) <= r * r]
    offs = np.array(offs)
    expect = np.zeros(len(pts), dtype=np.int64)
    for o in offs:
        ok = np.all((pts + o >= 0) & (pts + o < m), axis=1)
        expect += ok
    res = oracle_mod.self_join(x, eps)
    assert np.array_equal(res.counts.astype(np.int64), expect)
    bf = oracle_mod.brute_force(x, eps)
    assert np.array_equal(res.nbrs, bf.nbrs)


def test_uniform_square_mean_neighbours(oracle_mod):
    """Config 1 statistics: mean neighbours in the unit square with boundary
    loss is (N-1)(pi e^2 - 8 e^3/3 + e^4/2)."""
    x = datagen.uniform(65536, 2, 1001)
    eps = 0.01
    r = oracle_mod.self_join(x, eps)
    mean = r.total_pairs / len(x) - 1
    expect = (len(x) - 1) * (math.pi * eps ** 2 - 8 * eps ** 3 / 3 + eps ** 4 / 2)
    assert abs(mean - expect) < 0.15, (mean, expect)


# ------------------------------------------------------------- node net


def _adjacent_by_keys(ka, kb, sa, sb):
    """Independent statement of the node-net adjacency: X ~ Y iff, on every
    layer l below both stop depths, X and Y are siblings whose keys differ by
    exactly one, or have different parents and keys within one (mod 2^32)."""
    L = min(sa, sb)
    for l in range(L):
        same_parent = all(ka[i] == kb[i] for i in range(l))
        a, b = int(ka[l]), int(kb[l])
        if same_parent:
            if a != b and abs(a - b) != 1:
                return False
        elif ((a - b) % (1 << 32)) not in (0, 1, (1 << 32) - 1):
            return False
    return True


def _stop_depth(k):
    for l, v in enumerate(k):
        if v == 0xFFFFFFFF:
            return l
    return len(k)


@pytest.mark.parametrize("n,d,eps,seed", [(6000, 3, 0.03, 1), (5000, 8, 0.1, 2),
                                          (4000, 18, 0.3, 3)])
def test_node_net_invariants(n, d, eps, seed, oracle_mod):
    x = datagen.mixture(n, d, seed) if d > 3 else datagen.uniform(n, d, seed)
    order, _ = oracle_mod.dim_order(x)
    x = datagen.reorder_dims(x, order)
    r = oracle_mod.self_join(x, eps)
    M = r.num_nodes
    ns = r.node_start.astype(np.int64)
    assert ns[0] == 0 and ns[-1] == n and np.all(np.diff(ns) > 0)
    assert np.array_equal(np.sort(r.node_order), np.arange(n))
    for v in range(M):
        seg = r.node_order[ns[v]:ns[v + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)  # ascending within a node
    adj = [set(r.adj[r.adj_off[v]:r.adj_off[v + 1]].tolist()) for v in range(M)]
    for v in range(M):
        assert r.adj[r.adj_off[v]] == v  # self first
        assert len(adj[v]) == r.adj_off[v + 1] - r.adj_off[v]  # no duplicates
        for w in adj[v]:
            assert v in adj[w]  # symmetric
    sizes = np.diff(ns)
    assert r.num_calcs == sum(int(sizes[v]) * sum(int(sizes[w]) for w in adj[v])
                              for v in range(M))
    # independent characterisation from the key tuples
    stop = [_stop_depth(r.node_keys[v]) for v in range(M)]
    for v in range(M):
        want = {w for w in range(M)
                if _adjacent_by_keys(r.node_keys[v], r.node_keys[w], stop[v], stop[w])}
        assert want == adj[v], v
    # node order is lexicographic in the key tuples (unsplit layers = UINT_MAX)
    tuples = [tuple(k) for k in r.node_keys.tolist()]
    assert tuples == sorted(tuples)
    # completeness: the join equals brute force
    bf = oracle_mod.brute_force(x, eps)
    assert np.array_equal(r.counts, bf.counts)
    assert np.array_equal(r.nbrs, bf.nbrs)
    # the last layer's selection statistic, recomputed in Python floats
    avg = n * 1.0 / M
    s = 0.0
    for sz in sizes:
        s += (float(sz) - avg) * (float(sz) - avg)
    assert s == r.sse[-1]


def test_sampled_join_is_a_subset(oracle_mod):
    x = datagen.mixture(5000, 8, 4)
    full = oracle_mod.self_join(x, 0.09)
    part = oracle_mod.self_join(x, 0.09, sample_every=7)
    rows = full.node_order[np.arange(0, len(x), 7)]
    assert part.evals_done < full.evals_done
    for row in rows:
        a = full.nbrs[full.offsets[row]:full.offsets[row + 1]]
        b = part.nbrs[part.offsets[row]:part.offsets[row + 1]]
        assert np.array_equal(a, b)
