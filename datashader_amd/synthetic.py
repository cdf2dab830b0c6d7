"""Seeded synthetic inputs of the benchmark workloads (no datasets are fetched).

Host generators (numpy ``default_rng``) produce inputs small enough to check against the
CPU oracle; device generators (torch CUDA generators) produce the full-size inputs
directly in HBM.  Workloads:

  points_mixture   5-component 2-D Gaussian mixture, means U(-6,6)^2, sigmas
                   U(0.2,1.5), equal weights, a fraction of NaN rows
                   (800x600 count, 10M float64 rows; and the 1B-row serving set)
  points_skewed    90% N((0.5,0.5), 0.01^2), 10% U(0,1)^2, value lognormal(0, 2)
                   (1920x1080 float reductions, 100M float32 rows)
  points_ridges    x U(0,1); y 40% U(0,1), 60% on three ridges
                   0.25(k+1) + 0.1 sin(2 pi (k+1) x) + N(0, 0.02) (redrawn uniform when
                   outside [0,1]); category Zipf(1.2) over 16 categories
                   (1024x1024x16 by-category count, 500M rows)
  random_walks     lines of 256 samples, y0 U(-1,1), steps N(0, 0.02^2), shared
                   x = linspace(0, 1, 256) (2048x1024 line count, 1M lines)
"""
from __future__ import annotations

import math

import numpy as np


# ---- host (numpy) ------------------------------------------------------------------

def points_mixture_np(n, seed=1, nan_frac=0.001, dtype=np.float64, ncomp=5):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-6, 6, (ncomp, 2))
    sig = rng.uniform(0.2, 1.5, ncomp)
    comp = rng.integers(0, ncomp, n)
    x = means[comp, 0] + sig[comp] * rng.standard_normal(n)
    y = means[comp, 1] + sig[comp] * rng.standard_normal(n)
    k = int(n * nan_frac)
    if k:
        idx = rng.choice(n, k, replace=False)
        x[idx] = np.nan
        y[idx] = np.nan
    return x.astype(dtype), y.astype(dtype)


def points_skewed_np(n, seed=2, hot=0.9, sigma=0.01, dtype=np.float32):
    rng = np.random.default_rng(seed)
    nh = int(n * hot)
    x = np.concatenate([0.5 + sigma * rng.standard_normal(nh), rng.uniform(0, 1, n - nh)])
    y = np.concatenate([0.5 + sigma * rng.standard_normal(nh), rng.uniform(0, 1, n - nh)])
    perm = rng.permutation(n)
    v = rng.lognormal(0.0, 2.0, n)
    return x[perm].astype(dtype), y[perm].astype(dtype), v.astype(dtype)


def zipf_probs(ncat=16, s=1.2):
    p = (np.arange(ncat) + 1.0) ** -s
    return p / p.sum()


def points_ridges_np(n, seed=3, ncat=16):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, n)
    y = rng.uniform(0, 1, n)
    ridge = rng.uniform(0, 1, n) < 0.6
    k = rng.integers(0, 3, n)
    yr = 0.25 * (k + 1) + 0.1 * np.sin(2 * math.pi * (k + 1) * x) + rng.normal(0, 0.02, n)
    ok = ridge & (yr >= 0) & (yr <= 1)
    y = np.where(ok, yr, y)
    cat = rng.choice(ncat, n, p=zipf_probs(ncat)).astype(np.int8)
    return x.astype(np.float32), y.astype(np.float32), cat


def random_walks_np(nlines, nverts=256, seed=4, dtype=np.float32):
    rng = np.random.default_rng(seed)
    y0 = rng.uniform(-1, 1, (nlines, 1))
    steps = rng.normal(0, 0.02, (nlines, nverts - 1))
    y = np.concatenate([y0, y0 + np.cumsum(steps, axis=1)], axis=1)
    x = np.linspace(0, 1, nverts)
    return x, y.astype(dtype)


# ---- device (torch) ------------------------------------------------------------------

def _gen(seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def points_mixture_dev(n, seed=1, nan_frac=0.001, dtype=None, device="cuda", ncomp=5,
                       with_value=False):
    import torch
    dtype = dtype or torch.float64
    g = _gen(seed, device)
    means = (torch.rand(ncomp, 2, generator=g, device=device, dtype=torch.float64) * 12 - 6)
    sig = torch.rand(ncomp, generator=g, device=device, dtype=torch.float64) * 1.3 + 0.2
    out_x = torch.empty(n, dtype=dtype, device=device)
    out_y = torch.empty(n, dtype=dtype, device=device)
    out_v = torch.empty(n, dtype=torch.float32, device=device) if with_value else None
    chunk = 1 << 26
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        comp = torch.randint(0, ncomp, (m,), generator=g, device=device)
        zx = torch.randn(m, generator=g, device=device, dtype=torch.float64)
        zy = torch.randn(m, generator=g, device=device, dtype=torch.float64)
        out_x[s:s + m] = (means[comp, 0] + sig[comp] * zx).to(dtype)
        out_y[s:s + m] = (means[comp, 1] + sig[comp] * zy).to(dtype)
        if with_value:
            out_v[s:s + m] = torch.rand(m, generator=g, device=device)
        del comp, zx, zy
    k = int(n * nan_frac)
    if k:
        idx = torch.randint(0, n, (k,), generator=g, device=device)
        out_x[idx] = float("nan")
        out_y[idx] = float("nan")
    return out_x, out_y, out_v


def points_skewed_dev(n, seed=2, hot=0.9, sigma=0.01, device="cuda"):
    import torch
    g = _gen(seed, device)
    nh = int(n * hot)
    x = torch.empty(n, dtype=torch.float32, device=device)
    y = torch.empty(n, dtype=torch.float32, device=device)
    x[:nh] = 0.5 + sigma * torch.randn(nh, generator=g, device=device)
    y[:nh] = 0.5 + sigma * torch.randn(nh, generator=g, device=device)
    x[nh:] = torch.rand(n - nh, generator=g, device=device)
    y[nh:] = torch.rand(n - nh, generator=g, device=device)
    perm = torch.randperm(n, generator=g, device=device)
    x = x[perm]
    y = y[perm]
    del perm
    v = torch.exp(2.0 * torch.randn(n, generator=g, device=device))
    return x, y, v


def points_ridges_dev(n, seed=3, ncat=16, device="cuda"):
    import torch
    g = _gen(seed, device)
    x = torch.empty(n, dtype=torch.float32, device=device)
    y = torch.empty(n, dtype=torch.float32, device=device)
    cat = torch.empty(n, dtype=torch.int8, device=device)
    cdf = torch.tensor(np.cumsum(zipf_probs(ncat)), dtype=torch.float32, device=device)
    chunk = 1 << 26
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        xs = torch.rand(m, generator=g, device=device)
        yu = torch.rand(m, generator=g, device=device)
        ridge = torch.rand(m, generator=g, device=device) < 0.6
        k = torch.randint(0, 3, (m,), generator=g, device=device).float()
        yr = 0.25 * (k + 1) + 0.1 * torch.sin(2 * math.pi * (k + 1) * xs) \
            + 0.02 * torch.randn(m, generator=g, device=device)
        ok = ridge & (yr >= 0) & (yr <= 1)
        x[s:s + m] = xs
        y[s:s + m] = torch.where(ok, yr, yu)
        u = torch.rand(m, generator=g, device=device)
        cat[s:s + m] = torch.searchsorted(cdf, u).clamp_(max=ncat - 1).to(torch.int8)
    return x, y, cat


def random_walks_dev(nlines, nverts=256, seed=4, device="cuda"):
    """Returns (x float64[nverts] on the host, ys float32 block [nverts, nlines]):
    row j of the block is sample j of every line (one contiguous column per sample)."""
    import torch
    g = _gen(seed, device)
    block = torch.empty(nverts, nlines, dtype=torch.float32, device=device)
    block[0] = torch.rand(nlines, generator=g, device=device) * 2 - 1
    for j in range(1, nverts):
        block[j] = block[j - 1] + 0.02 * torch.randn(nlines, generator=g, device=device)
    return np.linspace(0, 1, nverts), block
