"""Seeded synthetic inputs for the benchmark configurations and the tests.

All generators use numpy's PCG64 (numpy.random.default_rng), so a seed gives
the same FP64 values on every machine.  Data seeds are 1000 + config number,
calibration-sample seeds 2000 + config number.

  c1  uniform [0,1)^2, 65,536 points, eps = 0.01
  c2  8-Gaussian mixture + 10% uniform noise, d = 18, min-max normalised
  c3  correlated Gaussian, covariance Q diag(i^-1.2) Q^T, d = 90, normalised
  c4  1,000 isotropic clusters (sigma 0.05, Zipf(1.1) sizes) clipped to [0,1],
      d = 384, stored as FP32
"""
import numpy as np


def uniform(n, d, seed):
    return np.random.default_rng(seed).random((n, d))


def _minmax(x):
    lo = x.min(axis=0)
    hi = x.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    return (x - lo) / span


def mixture(n, d, seed, k=8, noise=0.1, sig_lo=0.02, sig_hi=0.15):
    """c2 generator: k equal-weight Gaussians with per-(component, dim) sigma,
    plus `noise` fraction of uniform points, then per-dimension min-max."""
    rng = np.random.default_rng(seed)
    means = rng.random((k, d))
    sigmas = rng.uniform(sig_lo, sig_hi, size=(k, d))
    is_noise = rng.random(n) < noise
    comp = rng.integers(0, k, size=n)
    x = means[comp] + sigmas[comp] * rng.standard_normal((n, d))
    nn = int(is_noise.sum())
    x[is_noise] = rng.random((nn, d))
    return _minmax(x)


def correlated(n, d, seed):
    """c3 generator: x = Q diag(sqrt(lambda)) z, lambda_i = i^-1.2, Q from the QR
    of a Gaussian matrix with diag(R) > 0, then per-dimension min-max."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((d, d))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))[None, :]
    lam = np.arange(1, d + 1, dtype=np.float64) ** -1.2
    z = rng.standard_normal((n, d))
    x = (z * np.sqrt(lam)[None, :]) @ q.T
    return _minmax(x)


def clusters(n, d, seed, ncl=1000, zipf=1.1, sigma=0.05):
    """c4 generator: Zipf-sized isotropic clusters clipped to [0,1], FP32."""
    rng = np.random.default_rng(seed)
    centres = rng.random((ncl, d))
    w = np.arange(1, ncl + 1, dtype=np.float64) ** -zipf
    w /= w.sum()
    comp = rng.choice(ncl, size=n, p=w)
    out = np.empty((n, d), dtype=np.float32)
    step = 65536
    for s in range(0, n, step):
        e = min(n, s + step)
        blk = centres[comp[s:e]] + sigma * rng.standard_normal((e - s, d))
        out[s:e] = np.clip(blk, 0.0, 1.0).astype(np.float32)
    return out


def c5_specs(seed=1005, jobs=64):
    """Config 5: `jobs` independent instances; sizes log-uniform in
    [10^3, 10^6], d from {8, 18, 57, 90}, mixture data (Gaussian clusters +
    uniform noise), target mean neighbours uniform in [16, 64].  Returns
    (n, d, data seed, calibration seed, target) per instance."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(jobs):
        n = int(round(10 ** rng.uniform(3.0, 6.0)))
        d = int(rng.choice([8, 18, 57, 90]))
        target = int(rng.integers(16, 65))
        out.append((n, d, seed * 1000 + k, seed * 1000 + 500 + k, target))
    return out


def reorder_dims(x, order):
    """main.cu:117-122: column j of the result is column order[j] of x."""
    return np.ascontiguousarray(x[:, np.asarray(order, dtype=np.int64)])


def calibrate_eps(x, target, sample=10000, seed=2000, device=None, log=None, f32=False):
    """eps such that the mean neighbour count (self excluded) over a seeded
    sample of `sample` query rows is `target`: eps^2 is the (target*sample)-th
    smallest of the pooled query-to-all squared distances (self excluded).

    On a torch `device` the squared distances come from an FP64 GEMM,
    |q|^2 + |x|^2 - 2 q.x (clamped at 0), and the selection is an exact radix
    select over their IEEE bit patterns: three streaming passes, each a
    histogram of the next 22/21/21 bits among the values sharing the prefix
    found so far; nothing of size sample*n is stored.  With `f32`, an FP32
    GEMM and a 32-bit select instead (see _calibrate_f32).  Without a device,
    numpy computes the pool directly (small inputs only)."""
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    s = min(sample, n)
    qidx = np.sort(rng.choice(n, size=s, replace=False))
    want = target * s
    if device is None:
        X = np.ascontiguousarray(x, dtype=np.float64)
        pool = []
        for i in qidx:
            d2 = ((X - X[i]) ** 2).sum(axis=1)
            d2[i] = np.inf
            pool.append(d2)
        pool = np.concatenate(pool)
        return float(np.sqrt(np.partition(pool, want - 1)[want - 1]))

    import torch
    if f32:
        return _calibrate_f32(x, qidx, want, device, log)
    X = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)
    x2 = (X * X).sum(dim=1)
    Q = torch.as_tensor(qidx, device=device)
    chunk = max(1, int(4e9 // (8 * n)))
    inf_bits = 0x7FF0000000000000
    prefix, used, remaining = 0, 0, want
    for width in (22, 21, 21):
        shift = 64 - used - width
        hist = torch.zeros(1 << width, dtype=torch.int64, device=device)
        for a in range(0, s, chunk):
            rows = Q[a:a + chunk]
            q = X[rows]
            d2 = torch.addmm(x2[None, :], q, X.t(), beta=1.0, alpha=-2.0)
            d2 += x2[rows][:, None]
            d2.clamp_(min=0.0)
            d2[torch.arange(len(rows), device=device), rows] = float("inf")  # self
            b = d2.view(torch.int64)
            if used == 0:
                sel = b[b < inf_bits]
            else:
                sel = b[(b >> (64 - used)) == prefix]
            hist += torch.bincount((sel >> shift) & ((1 << width) - 1), minlength=1 << width)
            del d2, b, sel
        cum = torch.cumsum(hist, 0)
        digit = int(torch.searchsorted(cum, torch.tensor([remaining], device=device)).item())
        if digit > 0:
            remaining -= int(cum[digit - 1].item())
        prefix = (prefix << width) | digit
        used += width
        if log:
            log("calibration pass %d/3 done" % (used // 21))
    e2 = float(np.array([prefix], dtype=np.int64).view(np.float64)[0])
    return float(np.sqrt(e2))


def _calibrate_f32(x, qidx, want, device, log):
    """FP32 variant for large n*d: column-mean-centred rows, squared distances
    |q|^2 + |x|^2 - 2 q.x (clamped at 0) from an FP32 GEMM, and an exact radix
    select of the want-th smallest over their 32-bit patterns (two streaming
    passes of 16 bits)."""
    import torch
    n = x.shape[0]
    mu = np.zeros(x.shape[1], dtype=np.float64)
    for a in range(0, n, 65536):
        mu += np.asarray(x[a:a + 65536], dtype=np.float64).sum(axis=0)
    mu /= max(n, 1)
    X = torch.empty((n, x.shape[1]), dtype=torch.float32, device=device)
    mu_t = torch.as_tensor(mu, dtype=torch.float64, device=device)
    for a in range(0, n, 1 << 20):
        blk = torch.as_tensor(np.ascontiguousarray(x[a:a + (1 << 20)]), device=device)
        X[a:a + blk.shape[0]] = (blk.to(torch.float64) - mu_t).to(torch.float32)
    x2 = (X * X).sum(dim=1)
    Q = torch.as_tensor(qidx, device=device)
    s = len(qidx)
    chunk = max(1, min(s, int(2e9 // (4 * n))))
    inf_bits = 0x7F800000
    prefix, remaining = 0, want
    for p in range(2):
        hist = torch.zeros(1 << 16, dtype=torch.int64, device=device)
        for a in range(0, s, chunk):
            rows = Q[a:a + chunk]
            d2 = torch.addmm(x2[None, :], X[rows], X.t(), beta=1.0, alpha=-2.0)
            d2 += x2[rows][:, None]
            d2.clamp_(min=0.0)
            d2[torch.arange(len(rows), device=device), rows] = float("inf")
            b = d2.view(torch.int32)
            sel = b[b < inf_bits] if p == 0 else b[(b >> 16) == prefix]
            hist += torch.bincount((sel >> (16 - 16 * p)) & 0xFFFF, minlength=1 << 16)
            del d2, b, sel
            if log:
                log("calibration pass %d/2: %d/%d queries" % (p + 1, min(s, a + chunk), s))
        cum = torch.cumsum(hist, 0)
        digit = int(torch.searchsorted(cum, torch.tensor([remaining], device=device)).item())
        if digit > 0:
            remaining -= int(cum[digit - 1].item())
        prefix = (prefix << 16) | digit
    e2 = float(np.array([prefix], dtype=np.uint32).view(np.float32)[0])
    return float(np.sqrt(e2))
