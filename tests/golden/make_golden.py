#!/usr/bin/env python3
"""Writes tests/golden/golden.json: golden vectors for the self-join path.

For each case the input is generated from a seed (generators of
self-join-mistic_amd/datagen.py, dims put in the reference's sigma order,
main.cu:117-122), and the expected result is the FP64 brute force over all
pairs with the reference's per-pair operation sequence (kernel.cu:833-865):
no node net involved.  The node-net join of the oracle must give the same
result (this is the completeness certificate: every brute-force pair is a
candidate pair of the node net).  Recorded per case: SHA-256 of the input
bytes, of counts (u32 LE), offsets (u64 LE) and neighbour ids (u32 LE), the
pair total, the node-net summary, and the full counts array for the small
cases.  Also recorded: the first layer's reference points for configs 1 and
2 (utils.cu:66-137), from this platform's glibc rand().

usage: python tests/golden/make_golden.py   (rewrites golden.json)
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "self-join-mistic_amd"))

import datagen  # noqa: E402
import oracle  # noqa: E402

# name, generator, n, d, seed, eps (mean neighbours in the comment)
CASES = [
    ("uniform-2", "uniform", 8192, 2, 501, 0.02),           # ~10
    ("uniform-3", "uniform", 8192, 3, 502, 0.06),           # ~7
    ("mixture-8", "mixture", 8192, 8, 503, 0.09),
    ("mixture-18", "mixture", 8192, 18, 504, 0.25),
    ("mixture-57", "mixture", 6000, 57, 505, 0.553),
    ("correlated-90", "correlated", 5000, 90, 506, 1.14),
    ("clusters-384-f32", "clusters", 3000, 384, 507, 1.26),
    ("lattice-2", "lattice", 4096, 2, 0, 0.0625),           # exact distances
]


def make_input(kind, n, d, seed):
    if kind == "uniform":
        x = datagen.uniform(n, d, seed)
    elif kind == "mixture":
        x = datagen.mixture(n, d, seed)
    elif kind == "correlated":
        x = datagen.correlated(n, d, seed)
    elif kind == "clusters":
        x = datagen.clusters(n, d, seed, ncl=40)
    else:  # 64 x 64 grid with spacing 2^-6: every distance is exact
        g = np.arange(64, dtype=np.float64) / 64.0
        x = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).reshape(-1, 2)
    order, _ = oracle.dim_order(np.asarray(x, dtype=np.float64))
    return datagen.reorder_dims(x, order)


def sha(a, dt):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dt).tobytes()).hexdigest()


def main():
    out = {"cases": [], "rp": {}}
    for name, kind, n, d, seed, eps in CASES:
        x = make_input(kind, n, d, seed)
        x64 = np.asarray(x, dtype=np.float64)
        bf = oracle.brute_force(x64, eps)
        nn = oracle.self_join(x64, eps)
        assert np.array_equal(bf.counts, nn.counts) and np.array_equal(bf.nbrs, nn.nbrs), name
        case = {
            "name": name, "kind": kind, "n": n, "d": d, "seed": seed,
            "eps_hex": struct.pack(">d", eps).hex(),
            "storage": "f32" if x.dtype == np.float32 else "f64",
            "input_sha256": sha(x, x.dtype),
            "total_pairs": int(bf.total_pairs),
            "counts_sha256": sha(bf.counts, "<u4"),
            "offsets_sha256": sha(bf.offsets, "<u8"),
            "nbrs_sha256": sha(bf.nbrs, "<u4"),
            "num_nodes": int(nn.num_nodes), "num_calcs": int(nn.num_calcs),
            "chosen": [int(v) for v in nn.chosen],
        }
        if n <= 4096:
            case["counts"] = [int(v) for v in bf.counts]
        out["cases"].append(case)
        print("%-18s R=%d nodes=%d" % (name, bf.total_pairs, nn.num_nodes), flush=True)
    for cfg, (n, d, seed, kind) in {"c1": (65536, 2, 1001, "uniform"),
                                    "c2": (2000000, 18, 1002, "mixture")}.items():
        x = make_input(kind, n, d, seed)
        rp = oracle.rp_layers(np.asarray(x, dtype=np.float64), layers=1)[0]
        out["rp"][cfg] = {"n": n, "d": d, "seed": seed, "kind": kind,
                          "layer0_hex": [[struct.pack(">d", v).hex() for v in row] for row in rp]}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
