"""Writes the piecewise_fix_op_max golden fixtures under tests/golden/ from
tests/spec_opmax.py (an implementation independent of the C oracle).

    python tests/golden/make_golden_opmax.py

Each opmax_<case>.npz holds the input `values`, block size `bs`, the expected
`stream`, `offsets`, fix-up lists `plus` / `minus` (the reference's
mul_add_diff_set / mul_add_diff_set_minus) and `decoded` (the reference
decoder's output before the lists are applied); opmax_cases.json lists what
each case covers."""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import spec_opmax as spec  # noqa: E402


def cases():
    rng = np.random.default_rng(20240615)
    i = np.arange(60000, dtype=np.float64)
    return [
        ("linear", np.round(1e6 + 21.3 * i[:4000] + rng.integers(-3, 4, 4000)).astype(np.uint32), 200,
         "linear keys with noise, bs 200 (the published 1M-block configuration's block size)"),
        ("drift_plus", np.round(4.1e9 + 0.0271828 * i).astype(np.uint32), 30000,
         "long blocks whose additive chain lags: mul_add_diff_set entries"),
        ("drift_minus", np.round(3.0e9 + 0.4123457 * i[:30000]).astype(np.uint32), 30000,
         "long blocks whose additive chain leads: mul_add_diff_set_minus entries"),
        ("raw", rng.integers(0, 2**32, 900, dtype=np.uint64).astype(np.uint32), 300,
         "full-range values: raw blocks (tmp_bit 32)"),
        ("raw31", rng.integers(0, 2**31, 600, dtype=np.uint64).astype(np.uint32), 200,
         "tmp_bit 31 blocks: stored raw"),
        ("ragged", np.sort(rng.integers(0, 1 << 28, 1234, dtype=np.uint64)).astype(np.uint32), 100,
         "sorted keys, a 34-value last block"),
        ("tiny", rng.integers(0, 1000, 11, dtype=np.uint64).astype(np.uint32), 2,
         "two-value blocks and a one-value last block (NaN model)"),
        ("const", np.full(700, 123456, np.uint32), 256, "constant blocks: tmp_bit 0"),
        ("near_zero", (np.abs(np.round(50 - 0.37 * i[:800])) + rng.integers(0, 3, 800)).astype(np.uint32), 160,
         "predictions crossing zero"),
    ]


def main():
    meta = []
    for name, vals, bs, what in cases():
        y = [int(v) for v in vals]
        stream, offs, plus, minus = spec.encode_column(y, bs)
        dec = []
        for b in range(len(offs) - 1):
            m = min(bs, len(y) - b * bs)
            dec += spec.decode_block(stream[offs[b]:offs[b + 1]], m)
        np.savez(os.path.join(HERE, f"opmax_{name}.npz"), values=vals, bs=np.int64(bs),
                 stream=np.frombuffer(stream, np.uint8), offsets=np.array(offs, np.uint64),
                 plus=np.array(plus, np.int32), minus=np.array(minus, np.int32), decoded=np.array(dec, np.uint32))
        meta.append({"name": f"opmax_{name}", "bs": bs, "n": len(y), "covers": what,
                     "bits": sorted({stream[o] for o in offs[:-1]}), "plus": len(plus), "minus": len(minus)})
    with open(os.path.join(HERE, "opmax_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
