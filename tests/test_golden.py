"""Golden vectors (tests/golden/golden.json, written by make_golden.py from
the FP64 brute force with the reference's per-pair sequence): the oracle on
the CPU and libmsj.so on the GPU, every form, must reproduce the SHA-256 of
counts, offsets and neighbour lists, and the oracle the golden reference
points of configs 1 and 2."""
import hashlib
import json
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

with open(os.path.join(HERE, "golden", "golden.json")) as f:
    GOLDEN = json.load(f)
CASES = GOLDEN["cases"]


def _sha(a, dt):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dt).tobytes()).hexdigest()


def _eps(case):
    return struct.unpack(">d", bytes.fromhex(case["eps_hex"]))[0]


def _input(case):
    x = make_golden.make_input(case["kind"], case["n"], case["d"], case["seed"])
    assert _sha(x, x.dtype) == case["input_sha256"], "input generator drifted"
    return x


def _check(res, case):
    assert res.total_pairs == case["total_pairs"]
    assert _sha(res.counts, "<u4") == case["counts_sha256"]
    assert _sha(res.offsets, "<u8") == case["offsets_sha256"]
    assert _sha(res.nbrs, "<u4") == case["nbrs_sha256"]
    if "counts" in case:
        assert res.counts.tolist() == case["counts"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_golden(case, oracle_mod):
    x = _input(case)
    r = oracle_mod.self_join(np.asarray(x, dtype=np.float64), _eps(case))
    _check(r, case)
    assert r.num_nodes == case["num_nodes"] and r.num_calcs == case["num_calcs"]
    assert [int(v) for v in r.chosen] == case["chosen"]


def test_lattice_golden_closed_form():
    """The lattice case's total from the closed form: grid points at
    integer multiples of h = 1/64, eps = 4h, all distances exact."""
    case = [c for c in CASES if c["name"] == "lattice-2"][0]
    m, r = 64, 4
    offs = [(a, b) for a in range(-r, r + 1) for b in range(-r, r + 1) if a * a + b * b <= r * r]
    total = sum((m - abs(a)) * (m - abs(b)) for a, b in offs)
    assert case["total_pairs"] == total


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_reference_points_golden(cfg, oracle_mod):
    g = GOLDEN["rp"][cfg]
    x = make_golden.make_input(g["kind"], g["n"], g["d"], g["seed"])
    rp = oracle_mod.rp_layers(np.asarray(x, dtype=np.float64), layers=1)[0]
    want = np.array([[struct.unpack(">d", bytes.fromhex(v))[0] for v in row]
                     for row in g["layer0_hex"]])
    assert np.array_equal(rp, want)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "bf16", "split"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_golden(case, form, msj_mod):
    x = _input(case)
    print("\n[case] golden %s form=%s" % (case["name"], form), flush=True)
    kw = {"form": {"plain": 0, "bf16": 1, "split": 2}[form]}
    if case["storage"] == "f32":
        kw["storage"] = msj_mod.STORAGE_F32
    r = msj_mod.self_join(x, _eps(case), **kw)
    _check(r, case)
    assert r.stats["num_nodes"] == case["num_nodes"]
    assert r.stats["num_calcs"] == case["num_calcs"]
