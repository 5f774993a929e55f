"""Writes the golden fixtures under tests/golden/ from tests/spec_leco.py.

    python tests/golden/make_golden.py

The reference cannot be compiled here (its common.h needs Boost, GMP and
Eigen), so the expected streams come from the exact-arithmetic spec, which is
an implementation independent of the C oracle.  Each <case>.npz holds the
input `values`, the block size `bs`, the expected `stream` bytes, the block
`offsets`, and `lookup_idx` / `lookup_val` pairs; cases.json lists what each
case covers.  Inputs are seeded and small enough for pure Python.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "learn-to-compress_amd"))

import spec_leco as spec  # noqa: E402
from leco_fix import datagen  # noqa: E402

U64 = np.uint64
U32 = np.uint32


def crafted_lw1(width: int, n: int) -> np.ndarray:
    """Alternating 0 / 2^(w-2): residuals about +-2^(w-3) after centring, so
    the bit width comes out as w-1 (the block the reference decoder mis-reads)."""
    hi = 1 << (width - 2)
    return np.array([0 if k % 2 == 0 else hi for k in range(n)], U32 if width == 32 else U64)


def cases():
    rng = np.random.default_rng(20240611)
    yield "cfg1_head_u32", "linear 1000*i + N(0,50^2), exact fit regime, ragged last block", \
        datagen.linear_noise_u32(25000, seed=1), 10000
    yield "cfg1_kink_u32", "same generator at i ~ 4.29M: rounding regime, clamp kink, saturated blocks", \
        datagen.linear_noise_u32(30000, seed=1, start=4_280_000), 10000
    yield "pwl_u64", "sorted piecewise-linear keys ~2^45 (rounding regime), bs 4096, ragged last block", \
        datagen.sorted_pwl_u64(3 * 4096 + 1000, seed=7), 4096
    yield "bs1_u32", "one-value blocks: NaN theta, l up to raw", \
        np.array([0, 1, 5, 1 << 29, (1 << 30) + 3, (1 << 31) - 1, (1 << 32) - 1], U32), 1
    yield "bs1_u64", "one-value blocks incl. >= 2^63 (INT64_MIN predictions)", \
        np.array([0, 1, (1 << 53) + 1, 1 << 63, (1 << 63) + 5, (1 << 64) - 1], U64), 1
    yield "bs2_bs3_u64", "tiny blocks, bs 3, ragged", np.sort(rng.integers(0, 1 << 60, 11, dtype=U64)), 3
    yield "const_u64", "constant blocks (l = 0, 17 bytes)", \
        np.concatenate([np.full(256, 7, U64), np.full(256, 1 << 50, U64), np.full(100, 0, U64)]), 256
    yield "linear_exact_u64", "exact line 123456789 + 37 i (theta exact, l = 0)", \
        (123456789 + 37 * np.arange(3000, dtype=U64)).astype(U64), 1000
    yield "uniform_u64", "full-range random: raw blocks (l = 64)", rng.integers(0, 1 << 64, 600, dtype=U64, endpoint=False), 255
    yield "uniform_u32", "full-range random: raw blocks (l = 32)", rng.integers(0, 1 << 32, 700, dtype=U64).astype(U32), 256
    yield "negpred_u64", "uint64 block whose fitted line goes negative: raw by unsigned compare", \
        np.array([46] * 12 + [80, 16817, 27780, 40838] + [28988, 18080, 13552] + [43] * 13, U64), 16
    yield "high_u64", "values >= 2^63: out-of-range predictions, wrapping residuals", \
        np.sort(rng.integers(1 << 63, (1 << 64) - 1, 300, dtype=U64)), 100
    yield "lw1_u32", "crafted l = 31 block (reference decodeArray8 reads it as raw)", crafted_lw1(32, 64), 64
    yield "lw1_u64", "crafted l = 63 block (reference decodeArray8 reads it as raw)", crafted_lw1(64, 64), 64
    yield "zipf_u64", "sorted Zipf(1.2)-gapped keys", datagen.sorted_zipf_gaps(2000, 64, seed=11), 512
    yield "stepwise_u32", "stepwise runs Poisson(500), jumps U[1,2^16]", datagen.stepwise(3000, 32, seed=12), 1024
    yield "linsigma_u32", "linear + N(0, sigma)", datagen.linear_sigma(2500, 32, seed=13), 512
    yield "drift_u64", "a*i + b*i^2 + N(0, 2^10), bs 1024", datagen.drift_u64(2500, seed=14), 1024
    yield "ragged_u64", "n = 1500 at bs 1000 (last block 500)", datagen.sorted_uniform(1500, 64, seed=15), 1000


def main():
    manifest = []
    for name, desc, values, bs in cases():
        width = 8 * values.dtype.itemsize
        vals = [int(v) for v in values]
        stream, offs = spec.encode_column(vals, width, bs)
        # decode every block with the spec and check the round trip
        dec = []
        for b in range(len(offs) - 1):
            m = min(bs, len(vals) - b * bs)
            dec += spec.decode_block(stream[offs[b]:offs[b + 1]], m, width)
        assert dec == vals, name
        rng = np.random.default_rng(len(manifest) + 1)
        idx = rng.integers(0, len(vals), min(64, len(vals)), dtype=np.uint64)
        bits = [stream[offs[b]] for b in range(len(offs) - 1)]
        np.savez(os.path.join(HERE, f"{name}.npz"), values=values, bs=np.array(bs, np.uint64),
                 stream=np.frombuffer(stream, np.uint8), offsets=np.array(offs, np.uint64),
                 lookup_idx=idx, lookup_val=values[idx])
        manifest.append({"name": name, "width": width, "bs": bs, "n": len(vals), "bytes": len(stream),
                         "bits": sorted(set(bits)), "covers": desc})
        print(f"{name:18s} n={len(vals):6d} bs={bs:5d} bytes={len(stream):7d} l={sorted(set(bits))}")
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
