"""Writes the FOR / Delta golden fixtures under tests/golden/ from
tests/spec_for_delta.py (an implementation independent of the C oracle).

    python tests/golden/make_golden_base.py

Each base_<codec>_<case>.npz holds the input `values`, block size `bs` and the
expected `stream` and `offsets`; base_cases.json lists what each covers."""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import spec_for_delta as spec  # noqa: E402


def cases():
    rng = np.random.default_rng(20240612)
    out = []
    for width in (32, 64):
        dt = np.uint32 if width == 32 else np.uint64
        top = (1 << width) - 1
        out.append((f"sorted_{width}", np.sort(rng.integers(0, top >> 8, 700, dtype=np.uint64)).astype(dt), 128,
                    "sorted keys, several bit widths"))
        out.append((f"const_{width}", np.full(300, 77, dt), 100, "constant blocks: l = 0"))
        out.append((f"raw_{width}", rng.integers(0, top, 200, dtype=np.uint64, endpoint=True).astype(dt), 64,
                    "full-range values: raw blocks (FOR l = w; Delta l = w from deltas >= 2^(w-1))"))
        alt = np.array([0 if k % 2 == 0 else (1 << (width - 2)) + k for k in range(96)], np.uint64).astype(dt)
        out.append((f"wm1_{width}", alt, 32, "FOR l = w-1 (max - min in [2^(w-2), 2^(w-1))); Delta l = w"))
        out.append((f"ragged_{width}", rng.integers(0, 5000, 1001, dtype=np.uint64).astype(dt), 250,
                    "unsorted small values, a 1-value last block"))
        out.append((f"bits33_{width}", (np.arange(40, dtype=np.uint64) * np.uint64(1 << 27)).astype(dt), 40,
                    "FOR widths around 32 (the writer's (32 - l) / 8 pad)"))
    return out


def main():
    meta = []
    for name, vals, bs, what in cases():
        width = 32 if vals.dtype == np.uint32 else 64
        y = [int(v) for v in vals]
        for codec, enc in (("for", spec.for_encode), ("delta", spec.delta_encode)):
            blocks = [enc(y[s:s + bs], width) for s in range(0, len(y), bs)]
            offs = np.cumsum([0] + [len(b) for b in blocks]).astype(np.uint64)
            np.savez(os.path.join(HERE, f"base_{codec}_{name}.npz"), values=vals, bs=np.int64(bs),
                     stream=np.frombuffer(b"".join(blocks), np.uint8), offsets=offs)
            meta.append({"name": f"base_{codec}_{name}", "codec": codec, "width": width, "bs": bs, "covers": what,
                         "bits": sorted({b[0] for b in blocks})})
    with open(os.path.join(HERE, "base_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
