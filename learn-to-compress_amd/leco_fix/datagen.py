"""Seeded synthetic columns for the codec's workloads.

Each generator returns a numpy array (host) or, with device="cuda", a torch
tensor in HBM holding the same kind of data.  Host and device draws use
different random streams, so the two are alike in distribution, not in bits.

- linear_noise_u32: y_i = clamp(round(1000*i + N(0, 50^2)), 0, 2^32-1)
  (10M values at block size 10,000 is the reference-CPU workload).
- sorted_pwl_u64: sorted keys from piecewise-linear segments; segment lengths
  geometric with mean 2^20, slopes log-uniform in [1, 2^20], per-segment noise
  sigma log-uniform in [1, 2^12]; each segment starts where the previous one
  ended; noise is added, the keys are clamped at 0 and sorted.
- sorted_uniform, sorted_zipf_gaps, linear_sigma, stepwise: the four column
  distributions of the batched workload.  The stepwise generator follows the
  reference's scripts/generate_stepwise.py idea (runs of a constant value,
  then a jump), with run lengths Poisson(500) and jumps uniform in [1, 2^16].
- drift_u64: y = a*i + b*i^2 + N(0, 2^10) (variance 2^10, sigma 32), clamped at 0.
"""
from __future__ import annotations

import math

import numpy as np

U32_MAX = (1 << 32) - 1


def _finish(vals: np.ndarray, dtype) -> np.ndarray:
    hi = float(U32_MAX) if np.dtype(dtype) == np.uint32 else 1.8446744073709550e19
    return np.clip(np.rint(vals), 0, hi).astype(dtype)


def linear_noise_u32(n: int, seed: int = 1, start: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    i = np.arange(start, start + n, dtype=np.float64)
    return _finish(1000.0 * i + rng.normal(0.0, 50.0, n), np.uint32)


def pwl_segments(n: int, seed: int):
    """Segment (start, length, slope, sigma) list covering n values."""
    rng = np.random.default_rng(seed)
    segs, pos = [], 0
    while pos < n:
        length = int(min(rng.geometric(1.0 / (1 << 20)), n - pos))
        slope = math.exp(rng.uniform(0.0, math.log(1 << 20)))
        sigma = math.exp(rng.uniform(0.0, math.log(1 << 12)))
        segs.append((pos, length, slope, sigma))
        pos += length
    return segs


def sorted_pwl_u64(n: int, seed: int = 2, device: str = "cpu"):
    segs = pwl_segments(n, seed)
    if device == "cpu":
        rng = np.random.default_rng(seed + 1000)
        out = np.empty(n, np.float64)
        base = 0.0
        for pos, length, slope, sigma in segs:
            t = np.arange(length, dtype=np.float64)
            out[pos:pos + length] = base + slope * t + rng.normal(0.0, sigma, length)
            base += slope * length
        vals = _finish(out, np.uint64)
        vals.sort()
        return vals
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed + 1000)
    out = torch.empty(n, dtype=torch.float64, device=device)
    base = 0.0
    for pos, length, slope, sigma in segs:
        t = torch.arange(length, dtype=torch.float64, device=device)
        noise = torch.randn(length, dtype=torch.float64, device=device, generator=g)
        out[pos:pos + length] = base + slope * t + sigma * noise
        base += slope * length
    vals = torch.clamp(torch.round(out), min=0.0).to(torch.int64)
    del out
    vals, _ = torch.sort(vals)
    return vals


def sorted_uniform(n: int, width: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    hi = U32_MAX if width == 32 else (1 << 62)
    v = rng.integers(0, hi, n, dtype=np.uint64)
    v.sort()
    return v.astype(np.uint32 if width == 32 else np.uint64)


def sorted_zipf_gaps(n: int, width: int, seed: int, a: float = 1.2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    gaps = np.minimum(rng.zipf(a, n), 1 << 20).astype(np.uint64)
    v = np.cumsum(gaps, dtype=np.uint64)
    if width == 32:
        v = np.minimum(v, U32_MAX)
    return v.astype(np.uint32 if width == 32 else np.uint64)


def linear_sigma(n: int, width: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    hi = float(U32_MAX) if width == 32 else float(1 << 52)
    slope = math.exp(rng.uniform(0.0, math.log(max(hi / max(n, 1), 2.0))))
    sigma = math.exp(rng.uniform(0.0, math.log(1 << 10)))
    i = np.arange(n, dtype=np.float64)
    return _finish(np.minimum(slope * i + rng.normal(0.0, sigma, n), hi), np.uint32 if width == 32 else np.uint64)


def stepwise(n: int, width: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    runs = rng.poisson(500, n // 200 + 2) + 1
    jumps = rng.integers(1, 1 << 16, len(runs), dtype=np.uint64)
    levels = np.cumsum(jumps, dtype=np.uint64)
    v = np.repeat(levels, runs)[:n]
    if len(v) < n:
        v = np.concatenate([v, np.full(n - len(v), v[-1] if len(v) else 0, np.uint64)])
    if width == 32:
        v = np.minimum(v, U32_MAX)
    return v.astype(np.uint32 if width == 32 else np.uint64)


def drift_u64(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.uniform(1.0, float(1 << 16))
    b = math.exp(rng.uniform(math.log(1e-6), math.log(1e-3)))
    i = np.arange(n, dtype=np.float64)
    return _finish(a * i + b * i * i + rng.normal(0.0, 32.0, n), np.uint64)


DISTRIBUTIONS = {
    "sorted_uniform": sorted_uniform,
    "sorted_zipf": sorted_zipf_gaps,
    "linear_sigma": linear_sigma,
    "stepwise": stepwise,
}
