#!/usr/bin/env python3
"""Benchmark: MiSTIC epsilon self-join on one MI355X, FP64 vs MFMA filter.

Default workload: 515,345 points, d = 90, FP64 correlated Gaussian
(x = Q diag(sqrt(lambda)) z, lambda_i = i^-1.2), min-max normalised (seed
1003); eps from a 10,000-query seeded sample (seed 2003) so the mean
neighbour count, self excluded, is 32.  Both forms run on the same data:
the plain FP64 join and the split-BF16 MFMA filter with its FP64 band
recheck.

A step is one whole join on data already resident in HBM: node-net build,
distance kernels and the ordered result arrays (counts, CSR offsets, sorted
neighbour lists), which stay in HBM.  Each form is timed over K steps after
W warmup joins.  `value` is candidate distance evaluations per second of
the MFMA-filter form's step (the plain form's figures sit under "forms").

Also printed (same JSON line):
  roofline      the dominant distance kernel of the headline form
                (k_mfma_join): algorithmic flops (2 d per unordered candidate
                pair and product plane) over its average duration from HIP
                events on the library's stream, against the dense BF16 MFMA
                peak; the plain kernel's FP64 figure is under forms.plain;
                algorithmic_bytes = rows of every candidate tile read once;
                traffic (counter-measured HBM bytes) is null here, the
                rocprofv3 summaries under profiles/ hold it;
  mismatch_rows_between_forms  rows whose neighbour lists differ between the
                forms, over all rows;
  mismatch_rows_vs_brute_force  (default workload) rows whose neighbour list
                from the node net differs from the brute-force candidate
                generator's (every pair a candidate), over all rows;
  cpu_baseline  the CPU restatement (oracle) on every CPU this process may
                use, on a sampled join (every k-th query point in node order,
                full candidate lists), as candidate evaluations per second;
  mismatch      rows of that sample whose GPU neighbour list differs.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]
                       [--forms plain,split]
"""
import argparse
import json
import os
import struct
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "self-join-mistic_amd"))

import numpy as np  # noqa: E402

import datagen  # noqa: E402
import msj  # noqa: E402

METRIC = "self-join time & distance evals/s at fixed ε: FP64 vs MFMA-filter, 1 MI355X"
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA
FORMS = {"plain": 0, "bf16": 1, "split": 2}

# c5: eps of the 64 instances of datagen.c5_specs(), from `--config c5 --calibrate`
C5_EPS_HEX = [
    "3fec838005945f0a", "3fcea8991317beb3", "3fb08267cd523431", "3fe918b28fcf9517",
    "3fdbbe4b2f08cde6", "3fa959684580808c", "3fe2adde4ab2fbc4", "3fe5eb7510578640",
    "3fda5e8a0d1c4b2f", "3fd9798454f8d1db", "3fc76122ed76f7a9", "3fe14f23fafd37a8",
    "3fd78fed1bb43427", "3fdd0cd02d0eee3b", "3fd80bdebfa016c5", "3fd86caf0789b6a7",
    "3fe3f994ccd6451a", "3fe4814ef3a652fa", "3fd5a429bd4e2c62", "3fe277ad5c37c067",
    "3fc3aa0df825adea", "3fe055f1819d25cb", "3fc4030ec26e59a7", "3fe1e9c9129c71d1",
    "3fd36a689aab73c6", "3fb86f229e3c2890", "3fca304e49b95a3b", "3fece1126e3e6031",
    "3fe5a6a5c9bb8edb", "3fc5a5ba8d33ec4a", "3fd95c243a061b78", "3fe7810f73a817b3",
    "3fb4d86d0fe8e46e", "3fe0dff264f47d3e", "3fe2c0ffe65f993c", "3fb87d168c9fe1c4",
    "3fb20d42b0738b13", "3fba4af6377eddec", "3fe54aa46ac5e1b3", "3fe60869db787f86",
    "3fe0e027ddbe6dbf", "3fdc7f9f1d9556df", "3fe2a1d939173411", "3fbf6c53ba02da1d",
    "3fda800529596860", "3fe57889421cddff", "3fb0fd46708eafc8", "3fd1b0946b0da8ce",
    "3fd21a5fe00f761f", "3fdc3df694df6d9a", "3fbbd2032fdda971", "3feabbff3edc3ef5",
    "3fe8dc7934c34397", "3fe4fdca8f82c06c", "3fec87ac0f716ec8", "3fe593f2d0478bd3",
    "3fe97faaa1c9ca9d", "3fe3341d2a99bd31", "3fe38a6fd6b86bc9", "3fc113ffdf8c9987",
    "3fdcd49b67dc3c3d", "3fb5814c5caf8ead", "3fe288f4ffa36e1e", "3fd23697d99e4faf",
]

CONFIGS = {
    "c1": dict(n=65536, d=2, kind="uniform", seed=1001, eps=0.01,
               workload="plain FP64 join, 65,536 x 2 uniform [0,1), eps 0.01"),
    # eps_hex: `python bench.py --calibrate` on an MI355X (10,000 seeded queries)
    "c2": dict(n=2_000_000, d=18, kind="mixture", seed=1002, target=64, cal_seed=2002,
               eps_hex="3fc1a4c7913e006e",
               workload="plain FP64 join, 2,000,000 x 18 Gaussian mixture + noise, "
                        "mean 64 neighbours"),
    "c3": dict(n=515_345, d=90, kind="correlated", seed=1003, target=32, cal_seed=2003,
               eps_hex="3fe8aa41cdaa80fa",
               workload="FP64 join, 515,345 x 90 correlated Gaussian, mean 32 neighbours"),
    "c4": dict(n=5_000_000, d=384, kind="clusters", seed=1004, target=16, cal_seed=2004,
               cal_sample=2000, cal_f32=True, storage="f32", eps_hex="3ff28ce26ea18cb2",
               workload="FP64 decisions on FP32-stored 5,000,000 x 384 Zipf clusters, "
                        "mean 16 neighbours"),
}

# Runs in a child process: torch brings its own HIP runtime, which must not
# share a process with libmsj.so's.
_CALIBRATE = r"""
import sys, json
import numpy as np
sys.path.insert(0, sys.argv[1])
import datagen
x = np.load(sys.argv[2])
f32 = bool(int(sys.argv[6]))
eps = datagen.calibrate_eps(x, int(sys.argv[3]), sample=int(sys.argv[5]), seed=int(sys.argv[4]),
                            device="cuda:0", f32=f32,
                            log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
print(json.dumps({"eps": eps}))
"""


_CALIBRATE_C5 = r"""
import sys, json
sys.path.insert(0, sys.argv[1])
import datagen
out = []
for k, (n, d, seed, cal_seed, target) in enumerate(datagen.c5_specs()):
    x = datagen.mixture(n, d, seed)
    out.append(datagen.calibrate_eps(x, target, sample=min(2000, n), seed=cal_seed,
                                     device="cuda:0"))
    print("[bench] c5 instance %d/64 calibrated" % (k + 1), file=sys.stderr, flush=True)
print(json.dumps({"eps": out}))
"""


def calibrate_c5():
    out = subprocess.run([sys.executable, "-c", _CALIBRATE_C5,
                          os.path.join(ROOT, "self-join-mistic_amd")],
                         stdout=subprocess.PIPE, check=True, text=True).stdout
    return json.loads(out.strip().splitlines()[-1])["eps"]


def calibrate(x, cfg):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "x.npy")
        np.save(path, x)
        out = subprocess.run([sys.executable, "-c", _CALIBRATE,
                              os.path.join(ROOT, "self-join-mistic_amd"), path,
                              str(cfg["target"]), str(cfg["cal_seed"]),
                              str(cfg.get("cal_sample", 10000)), str(int(cfg.get("cal_f32", False)))],
                             stdout=subprocess.PIPE, check=True, text=True).stdout
    return json.loads(out.strip().splitlines()[-1])["eps"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(cfg):
    n, d = cfg["n"], cfg["d"]
    if cfg["kind"] == "uniform":
        x = datagen.uniform(n, d, cfg["seed"])
    elif cfg["kind"] == "mixture":
        x = datagen.mixture(n, d, cfg["seed"])
    elif cfg["kind"] == "correlated":
        x = datagen.correlated(n, d, cfg["seed"])
    else:
        x = datagen.clusters(n, d, cfg["seed"])
    order, _ = msj.dim_order(x.astype(np.float64, copy=False))
    return datagen.reorder_dims(x, order)


def cpu_baseline(x, eps, num_calcs, gpu_rows, seconds, threads):
    """Oracle (CPU restatement) on a sampled join; returns (dict, mismatch)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: WPS433 - the checker and CPU baseline

    # pilot: about 2e8 evaluations give this machine's rate at this shape
    pilot = oracle.self_join(x, eps, sample_every=max(1, int(num_calcs // 2e8)),
                             nthreads=threads)
    rate_guess = pilot.evals_done / max(pilot.t_join, 1e-3)
    every = max(1, int(num_calcs // max(1.0, rate_guess * seconds)))
    log("[bench] CPU pilot: %.3g evals/s; sampling every %d-th query" % (rate_guess, every))
    t0 = time.perf_counter()
    r = oracle.self_join(x, eps, sample_every=every, nthreads=threads)
    wall = time.perf_counter() - t0
    rate = r.evals_done / r.t_join if r.t_join > 0 else 0.0
    # parity on the sampled rows (full neighbour lists on both sides)
    pos = np.arange(0, x.shape[0], every, dtype=np.int64)
    rows = r.node_order[pos]
    counts_g, offs_g, nbrs_g, _ = gpu_rows
    mism = 0
    for row in rows:
        a0, a1 = int(r.offsets[row]), int(r.offsets[row + 1])
        b0, b1 = int(offs_g[row]), int(offs_g[row + 1])
        if counts_g[row] != r.counts[row] or not np.array_equal(r.nbrs[a0:a1], nbrs_g[b0:b1]):
            mism += 1
    same_index = bool(np.array_equal(r.node_order, gpu_rows[3]))
    out = {
        "value": rate, "unit": "distance evals/s", "cores": threads, "kind": "port",
        "sample": "oracle (C/OpenMP restatement): full node-net build (%.1f s) + join of every "
                  "%d-th query point in node order (%d rows, %.3g candidate evaluations, "
                  "%.1f s); value = evaluations / join seconds" % (
                      r.t_construct, every, len(rows), r.evals_done, r.t_join),
        "construct_s": r.t_construct, "join_s": r.t_join, "wall_s": wall,
        "sample_rows": int(len(rows)), "node_order_equal": same_index,
    }
    return out, mism


def host_cpus():
    """CPUs this process may use: the smaller of its affinity set and its
    cgroup CPU quota (a GPU box grants each GPU a share of its cores)."""
    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = os.cpu_count()
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    info["omp_num_threads_env"] = os.environ.get("OMP_NUM_THREADS")
    try:
        out = subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout
        info["nproc"] = int(out.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        info["nproc"] = None
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    info["model"] = model
    usable = info["affinity"] or 1
    if quota:
        usable = max(1, min(usable, int(quota)))
    if info["nproc"]:
        usable = min(usable, info["nproc"])
    return usable, info


def kernel_name(st, d):
    if st["form_used"] == 0:
        return "k_join_sym" if (d + 3) // 4 * 4 <= 32 else "k_join_hd"
    return "k_mfma_join"


def run_form(x, eps, cfg, form_name, args, storage):
    """Counting pass (untimed, host output), then `warmup` + `steps` timed
    joins on resident data.  Returns (summary dict, host result of the
    counting pass)."""
    form = FORMS[form_name]
    verbose = 1 if args.verbose else 0
    n, d = x.shape
    with msj.Context(output=msj.OUT_HOST, count_work=1, storage=storage, form=form,
                     verbose=verbose) as cc:
        cc.load(x)
        ref_run = cc.join(eps)
    rs = ref_run.stats
    log("[bench] %s counting pass: R=%d nodes=%d candidates=%d dims=%d form_used=%d" % (
        form_name, ref_run.total_pairs, rs["num_nodes"], rs["num_calcs"], rs["dims_evaluated"],
        rs["form_used"]))
    if form != 0:
        log("[bench] %s filter check: band=%d checked=%d violations=%d max|err|=%.3g "
            "mean|err|=%.3g" % (form_name, rs["band_pairs"], rs["filter_checked"],
                                rs["filter_violations"], rs["max_score_err"],
                                rs["mean_score_err"]))
    if args.single_pass:
        r = ref_run
        kms, mms = [r.stats["ms_join_kernel"]], [r.stats["ms_main_kernel"]]
        stage = [(r.stats["ms_index"], r.stats["ms_join"], r.stats["ms_output"])]
        ms_step = r.stats["ms_total"] - r.stats["ms_d2h"]
        steps, warmup = 1, 0
    else:
        steps, warmup = args.steps, args.warmup
        # msj_ctx_join returns after its stream has drained: the clock below
        # brackets exactly the K joins
        with msj.Context(output=msj.OUT_DEVICE, storage=storage, form=form,
                         verbose=verbose) as ctx:
            ctx.load(x)
            for w in range(warmup):
                r = ctx.join(eps, copy_arrays=False)
                log("[bench] %s warmup %d: %.1f ms" % (form_name, w, r.stats["ms_total"]))
            kms, mms, stage = [], [], []
            t0 = time.perf_counter()
            for k in range(steps):
                r = ctx.join(eps, copy_arrays=False)
                kms.append(r.stats["ms_join_kernel"])
                mms.append(r.stats["ms_main_kernel"])
                stage.append((r.stats["ms_index"], r.stats["ms_join"], r.stats["ms_output"]))
            wall = time.perf_counter() - t0
        ms_step = wall * 1e3 / steps
    st = r.stats
    C, R = st["num_calcs"], r.total_pairs
    k_ms, m_ms = float(np.mean(kms)), float(np.mean(mms))
    ms_index, ms_join, ms_output = (float(v) for v in np.mean(stage, axis=0))
    log("[bench] %s: %d steps, %.2f ms/step (index %.2f, join %.2f [kernels %.2f, %s %.2f], "
        "output %.2f ms)" % (form_name, steps, ms_step, ms_index, ms_join, k_ms,
                             kernel_name(st, d), m_ms, ms_output))
    kname = kernel_name(st, d)
    dpad = (d + 3) // 4 * 4
    if st["form_used"] == 0:
        flops = 3.0 * rs["dims_evaluated"]  # 3 flops per FP64 dimension consumed
        roof = {"bound": "fp64_valu", "peak": FP64_PEAK_TFLOPS,
                "kernel": kname + " (FP64 vector ALU; MI355X FP64 vector peak)"}
        alg_bytes = 8.0 * d * st["tile_rows"] + 8.0 * R + 4.0 * n
    else:
        # every unordered candidate pair (self pairs included) scored over d
        # dimensions, 2 flops per product; split-BF16 adds the two cross
        # products where they ran (every pair up to d = 128; above, the wave
        # blocks of tiles that the BF16 scores do not settle)
        mult = 1.0 + 2.0 * st["split_share_ppm"] / 1e6 if st["form_used"] == 2 else 1.0
        flops = (C + n) / 2.0 * d * 2.0 * mult
        roof = {"bound": "mfma", "peak": BF16_PEAK_TFLOPS,
                "kernel": kname + " (BF16 MFMA, FP32 accumulate; MI355X dense BF16 peak)"}
        planes = 2 if st["form_used"] == 2 else 1
        alg_bytes = 2.0 * planes * ((dpad + 15) // 16 * 16) * st["tile_rows"] + 8.0 * R
    achieved = flops / (m_ms / 1e3) / 1e12
    roof.update({"achieved": achieved, "unit": "TFLOP/s", "frac": achieved / roof["peak"],
                 "traffic": None, "kernel_ms": m_ms, "algorithmic_flops": flops,
                 "algorithmic_bytes": alg_bytes,
                 "algorithmic_bytes_per_s": alg_bytes / (m_ms / 1e3)})
    out = {
        "form": form_name, "form_used": st["form_used"], "ms_per_step": ms_step,
        "steps": steps, "warmup": warmup, "join_time_s": ms_step / 1e3,
        "evals_per_s": C / (ms_step / 1e3), "pairs_per_s": R / (ms_step / 1e3),
        "total_pairs": R, "candidate_evals": C, "num_nodes": st["num_nodes"],
        "ms_index": ms_index, "ms_join": ms_join, "ms_output": ms_output,
        "ms_join_kernels": k_ms, "roofline": roof,
        "dtype": {0: "f64", 1: "bf16 MFMA (f32 acc) + f64 band",
                  2: "split-bf16 MFMA (f32 acc) + f64 band"}[st["form_used"]],
    }
    if st["form_used"] != 0:
        out["filter"] = {
            "band_pairs": rs["band_pairs"],
            "band_fraction": rs["band_pairs"] / max(1.0, (C + n) / 2.0),
            "max_score_err": rs["max_score_err"], "mean_score_err": rs["mean_score_err"],
            "sampled_pairs_checked": rs["filter_checked"],
            "violations": rs["filter_violations"],
            "mfma_flops_issued": st["mfma_flops"],
            "split_share": st["split_share_ppm"] / 1e6,
            "mfma_issued_frac_of_peak": st["mfma_flops"] / (m_ms / 1e3) / 1e12 / BF16_PEAK_TFLOPS,
            "ms_band_kernel": k_ms - m_ms,
        }
    return out, ref_run


def same_result(a, b):
    """Rows whose neighbour lists differ between two host results."""
    if a.total_pairs == b.total_pairs and np.array_equal(a.counts, b.counts) \
            and np.array_equal(a.nbrs, b.nbrs):
        return 0
    bad = 0
    for row in range(a.n):
        a0, a1 = int(a.offsets[row]), int(a.offsets[row + 1])
        b0, b1 = int(b.offsets[row]), int(b.offsets[row + 1])
        if a.counts[row] != b.counts[row] or not np.array_equal(a.nbrs[a0:a1], b.nbrs[b0:b1]):
            bad += 1
    return bad


DEFAULT_FORMS = {"c1": "plain", "c2": "plain", "c3": "plain,split", "c4": "split"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS) + ["c5"])
    ap.add_argument("--forms", default=None,
                    help="comma list of forms to time (plain, bf16, split); default: "
                         "plain,split for c3, plain for c1/c2, split for c4")
    ap.add_argument("--form", default=None, choices=sorted(FORMS),
                    help="time one form only (same as --forms FORM)")
    ap.add_argument("--eps", type=float, default=None, help="override the calibrated eps")
    ap.add_argument("--calibrate", action="store_true",
                    help="recompute eps from the seeded 10,000-query sample, print it, exit")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--single-pass", action="store_true",
                    help="time the counting join itself (one join instead of 1 + warmup + "
                         "steps; its kernel keeps work counters, a few percent slower)")
    ap.add_argument("--rows", type=int, default=0,
                    help="join only the first ROWS rows of the configuration's data")
    ap.add_argument("--verbose", action="store_true",
                    help="library progress lines (layers, join item ranges) on stderr")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: every CPU this process may use)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (torch), for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 18  # float32 elements
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 5}), flush=True)
        return

    if args.config == "c5":
        return run_c5(args)
    cfg = CONFIGS[args.config]
    forms = (args.form or args.forms or DEFAULT_FORMS[args.config]).split(",")
    for f in forms:
        if f not in FORMS:
            ap.error("unknown form %r" % f)
    log("[bench] generating %s: n=%d d=%d" % (args.config, cfg["n"], cfg["d"]))
    x = make_data(cfg)
    if args.rows:
        x = np.ascontiguousarray(x[:args.rows])
        cfg = dict(cfg, n=x.shape[0], workload=cfg["workload"] + " (first %d rows)" % x.shape[0])
        log("[bench] joining the first %d rows" % x.shape[0])
    if args.eps is not None:
        eps = args.eps
    elif "eps" in cfg and not args.calibrate:
        eps = cfg["eps"]
    elif "eps_hex" in cfg and not args.calibrate:
        eps = struct.unpack(">d", bytes.fromhex(cfg["eps_hex"]))[0]
    else:
        log("[bench] calibrating eps on the GPU (%d seeded queries)" % cfg.get("cal_sample", 10000))
        eps = calibrate(x, cfg)
    eps_hex = struct.pack(">d", eps).hex()
    log("[bench] eps = %r (0x%s)" % (eps, eps_hex))
    if args.calibrate:
        print(json.dumps({"config": args.config, "eps": eps, "eps_hex": eps_hex}), flush=True)
        return

    storage = msj.STORAGE_F32 if cfg.get("storage") == "f32" else msj.STORAGE_F64
    results, hosts = {}, {}
    for f in forms:
        results[f], hosts[f] = run_form(x, eps, cfg, f, args, storage)
    # headline: the MFMA-filter form when one was timed, else the plain form
    head = next((f for f in reversed(forms) if results[f]["form_used"] != 0), forms[0])
    H = results[head]
    line = {
        "metric": METRIC, "value": H["evals_per_s"], "unit": "distance evals/s", "n_gpus": 1,
        "steps": H["steps"], "warmup": H["warmup"], "ms_per_step": H["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": H["dtype"], "mean_neighbours": H["total_pairs"] / cfg["n"] - 1.0,
        "data": "synthetic (seeded, generated in-process)",
        "config": {"workload": cfg["workload"], "n": cfg["n"], "d": cfg["d"], "eps": eps,
                   "eps_hex": eps_hex, "form": head, "forms_timed": forms,
                   "parallelism": "none (1 GPU)",
                   "timing": "one counting join (work counters on)" if args.single_pass
                   else "%d joins on resident data per form" % H["steps"]},
        "join_time_s": H["join_time_s"], "pairs_per_s": H["pairs_per_s"],
        "total_pairs": H["total_pairs"], "candidate_evals": H["candidate_evals"],
        "num_nodes": H["num_nodes"], "ms_index": H["ms_index"], "ms_join": H["ms_join"],
        "ms_output": H["ms_output"], "roofline": H["roofline"], "forms": results,
    }
    if "filter" in H:
        line["filter"] = H["filter"]
    if "plain" in results and head != "plain":
        line["speedup_filter_vs_plain"] = results["plain"]["ms_per_step"] / H["ms_per_step"]
        line["speedup_filter_vs_plain_join_kernels"] = (results["plain"]["ms_join_kernels"]
                                                        / H["ms_join_kernels"])
    # full-size parity between the forms (every row, both host results)
    if len(forms) > 1:
        base = hosts[forms[0]]
        line["mismatch_rows_between_forms"] = {
            f: same_result(base, hosts[f]) for f in forms[1:]}
        line["node_order_equal_between_forms"] = all(
            np.array_equal(base.node_order, hosts[f].node_order) for f in forms[1:])
        log("[bench] forms compared on every row: %s" % line["mismatch_rows_between_forms"])
    # full-size parity of the candidate generator: the node net's join against
    # the brute-force generator (one node, every pair a candidate, the same
    # FP64 decisions) on every row, in the fastest timed form (untimed)
    if args.config == "c3" and not args.rows:
        fb = head
        log("[bench] brute-force generator (%s form), all rows" % fb)
        bf = msj.self_join(x, eps, form=FORMS[fb], index=msj.INDEX_BRUTE,
                           storage=storage)
        line["mismatch_rows_vs_brute_force"] = same_result(hosts[fb], bf)
        line["brute_force_candidate_evals"] = bf.stats["num_calcs"]
        log("[bench] node net vs brute force on every row: %d rows differ"
            % line["mismatch_rows_vs_brute_force"])
    if not args.no_cpu_baseline:
        threads, hostinfo = host_cpus()
        if args.cpu_threads:
            threads = args.cpu_threads
        log("[bench] CPU baseline (oracle, %d threads of %s, ~%.0f s)" % (
            threads, hostinfo, args.cpu_seconds))
        ref = hosts[forms[0]]
        gpu_rows = (ref.counts, ref.offsets, ref.nbrs, ref.node_order)
        cb, mism = cpu_baseline(x, eps, H["candidate_evals"], gpu_rows, args.cpu_seconds,
                                threads)
        cb["host"] = hostinfo
        line["cpu_baseline"] = cb
        line["mismatch"] = mism
        log("[bench] CPU %.3g evals/s; sampled-row mismatch %d" % (cb["value"], mism))
    print(json.dumps(line), flush=True)


def run_c5(args):
    if args.calibrate or not C5_EPS_HEX:
        log("[bench] calibrating the 64 instances of c5 on the GPU")
        eps = calibrate_c5()
        hexes = [struct.pack(">d", e).hex() for e in eps]
        print(json.dumps({"config": "c5", "eps_hex": hexes}), flush=True)
        return
    specs = datagen.c5_specs()
    eps = [struct.unpack(">d", bytes.fromhex(h))[0] for h in C5_EPS_HEX]
    log("[bench] generating c5: %d instances, %d points" % (len(specs), sum(p[0] for p in specs)))
    datas = []
    for (n, d, seed, _, _) in specs:
        x = datagen.mixture(n, d, seed)
        order, _ = msj.dim_order(x)
        datas.append(datagen.reorder_dims(x, order))
    # counting pass (untimed): FP64 work of the plain kernel, per job
    dims_eval, calcs = 0, 0
    for x, e in zip(datas, eps):
        r = msj.self_join(x, e, count_work=1)
        dims_eval += r.stats["dims_evaluated"]
        calcs += r.stats["num_calcs"]
    log("[bench] counting pass: candidates=%d dims=%d" % (calcs, dims_eval))

    form = args.form or "plain"
    fcode = FORMS[form]
    with msj.Batch(output=msj.OUT_DEVICE, form=fcode) as b:
        b.load(datas, eps)
        for w in range(max(1, args.warmup)):
            _, st = b.run(copy_arrays=False)
            log("[bench] warmup %d: %.1f ms (tune %.1f ms, solo jobs %d, pooled classes %s, "
                "epi log2 %s, slab dims %s, grid %d)" % (
                    w, st["ms_total"], st["ms_tune"], st["solo_jobs"], st["class_ds"],
                    st["class_epi_log2"], st["class_slab"], st["grid_blocks"]))
        stage = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            _, st = b.run(copy_arrays=False)
            stage.append((st["ms_index"], st["ms_join"], st["ms_output"]))
        wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    ms_index, ms_join, ms_output = (float(v) for v in np.mean(stage, axis=0))
    log("[bench] batch: %.2f ms/step (builds done after %.2f ms; join stream %.2f ms of device "
        "time; output stage busy %.2f ms; the three overlap)" % (ms_step, ms_index, ms_join,
                                                                 ms_output))

    # the same 64 joins as sequential plain launches, data resident
    ctxs = []
    try:
        for x in datas:
            c = msj.Context(output=msj.OUT_DEVICE, form=fcode)
            c.load(x)
            ctxs.append(c)
        for c, e in zip(ctxs, eps):  # warm
            c.join(e, copy_arrays=False)
        t0 = time.perf_counter()
        for k in range(args.steps):
            for c, e in zip(ctxs, eps):
                c.join(e, copy_arrays=False)
        seq_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    finally:
        for c in ctxs:
            c.close()
    log("[bench] 64 sequential %s joins: %.2f ms/step" % (form, seq_ms))

    # parity: batch (host output) against the sequential single joins
    log("[bench] parity: batch with host output vs 64 single joins")
    res = msj.join_batch(datas, eps, form=fcode)
    mism = 0
    for k, (x, e, r) in enumerate(zip(datas, eps, res)):
        s1 = msj.self_join(x, e)
        if not (r.total_pairs == s1.total_pairs and np.array_equal(r.counts, s1.counts)
                and np.array_equal(r.nbrs, s1.nbrs)):
            mism += 1
        if k % 16 == 15:
            log("[bench] parity: %d/64 jobs compared, %d mismatching" % (k + 1, mism))
    R = sum(r.total_pairs for r in res)
    flops = 3.0 * dims_eval
    achieved = flops / (ms_join / 1e3) / 1e12
    line = {
        "metric": METRIC, "value": calcs / (ms_step / 1e3), "unit": "distance evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if form == "plain" else
                 "f64 pooled jobs; %s MFMA filter + f64 band on the solo jobs" % form,
        "data": "synthetic (seeded, generated in-process)",
        "config": {"workload": "batching form: 64 independent FP64 joins, sizes "
                               "log-uniform in [1e3, 1e6], d in {8, 18, 57, 90}, mixture data, "
                               "mean 16-64 neighbours", "jobs": len(datas),
                   "total_points": int(sum(p[0] for p in specs)), "form": "batch",
                   "join_form": form,
                   "parallelism": "none (1 GPU)"},
        "join_time_s": ms_step / 1e3, "pairs_per_s": R / (ms_step / 1e3), "total_pairs": R,
        "candidate_evals": calcs, "ms_index": ms_index, "ms_join": ms_join,
        "ms_output": ms_output, "sequential_ms_per_step": seq_ms,
        "speedup_vs_sequential": seq_ms / ms_step, "mismatch_jobs": mism,
        "tuned": {"class_ds": st["class_ds"], "epi_log2": st["class_epi_log2"],
                  "slab_dims": st["class_slab"],
                  "grid_blocks": st["grid_blocks"], "work_items": st["work_items"]},
        "solo_jobs": st["solo_jobs"],
        "stages_overlap": "builds (2 builder streams) overlap the joins (join stream) and the "
                          "outputs (output stream)",
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "join stream: solo k_join_hd / k_join_sym launches + persistent "
                               "k_join_batch (FP64 vector ALU; MI355X FP64 peak)",
                     "kernel_ms": ms_join, "algorithmic_flops": flops},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # noqa: WPS433 - the CPU baseline
        threads, hostinfo = host_cpus()
        if args.cpu_threads:
            threads = args.cpu_threads
        # about 2e8 candidate evaluations/s on the box's cores at d = 57..90
        budget = max(1.0, 2e8 * args.cpu_seconds)
        every = max(1, int(calcs // budget))
        done, tj = 0.0, 0.0
        log("[bench] CPU baseline (oracle, %d threads): every %d-th query" % (
            threads, every))
        for k, (x, e) in enumerate(zip(datas, eps)):
            r = oracle.self_join(x, e, sample_every=every, nthreads=threads)
            done += r.evals_done
            tj += r.t_join
            if k % 8 == 7:
                log("[bench] CPU baseline: %d/64 instances, %.1f s of join" % (k + 1, tj))
        line["cpu_baseline"] = {
            "value": done / tj if tj > 0 else 0.0, "unit": "distance evals/s",
            "cores": threads, "host": hostinfo, "kind": "port",
            "sample": "oracle (C/OpenMP restatement) on all 64 instances: node nets + join of "
                      "every %d-th query point in node order (%.3g candidate evaluations, "
                      "%.1f s of join)" % (every, done, tj)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
