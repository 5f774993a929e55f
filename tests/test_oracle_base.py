"""CPU checks that pin the oracle's FOR_int / Delta_int baselines
(headers/FOR_integer_template.h, delta_integer_template.h): the golden
fixtures written by tests/spec_for_delta.py, the spec on random blocks,
closed-form sizes, round trips, the FOR filters against brute force, and the
documented deviations."""
import json
import os

import numpy as np
import pytest

import spec_for_delta as spec
from conftest import GOLDEN

with open(os.path.join(GOLDEN, "base_cases.json")) as _f:
    BASE = json.load(_f)


def _load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", [c["name"] for c in BASE])
def test_base_golden(oracle_mod, case):
    g = _load(case)
    codec = case.split("_")[1]
    vals, bs = g["values"], int(g["bs"])
    st, off = oracle_mod.base_encode_column(codec, vals, bs)
    assert np.array_equal(off, g["offsets"])
    assert st.tobytes() == g["stream"].tobytes()
    dec = oracle_mod.base_decode_column(codec, g["stream"], g["offsets"], len(vals), bs, vals.dtype)
    assert np.array_equal(dec, vals)


def _random_block(rng, width, n, kind):
    top = (1 << width) - 1
    if kind == 0:
        v = np.sort(rng.integers(0, top >> int(rng.integers(0, width - 1)), n, dtype=np.uint64, endpoint=True))
    elif kind == 1:
        v = rng.integers(0, top, n, dtype=np.uint64, endpoint=True)
    elif kind == 2:
        v = np.full(n, rng.integers(0, top, dtype=np.uint64, endpoint=True), np.uint64)
    else:
        v = rng.integers(0, 1 << int(rng.integers(1, width)), n, dtype=np.uint64)
    return v.astype(np.uint32 if width == 32 else np.uint64)


def test_base_oracle_matches_spec_on_random_blocks(oracle_mod):
    rng = np.random.default_rng(17)
    for _ in range(300):
        width = int(rng.choice([32, 64]))
        n = int(rng.choice([1, 2, 3, 17, 100, 257, 1000]))
        y = _random_block(rng, width, n, int(rng.integers(0, 4)))
        yl = [int(v) for v in y]
        for codec, enc, dec in (("for", spec.for_encode, spec.for_decode),
                                ("delta", spec.delta_encode, spec.delta_decode)):
            blk = oracle_mod.base_encode_block(codec, y)
            assert blk == enc(yl, width), (codec, width, n)
            assert len(blk) == oracle_mod.base_block_bytes(codec, blk[0], width, n)
            got = oracle_mod.base_decode_block(codec, blk, n, y.dtype)
            assert np.array_equal(got, y) and dec(blk, n, width) == yl
            for j in sorted({0, n // 2, n - 1}):
                assert oracle_mod.base_lookup_block(codec, blk, j, y.dtype) == yl[j]


def test_base_sizes_closed_form(oracle_mod):
    # FOR: l = 4, n = 2 -> header 1 + 2*8, E + 1 = 2 bytes, (32 - 4) / 8 = 3 pad bytes
    assert oracle_mod.base_block_bytes("for", 4, 64, 2) == 1 + 16 + 2 + 3
    assert oracle_mod.base_block_bytes("for", 33, 64, 10) == 1 + 16 + 42 + 1 + 0
    assert oracle_mod.base_block_bytes("for", 0, 32, 100) == 9
    assert oracle_mod.base_block_bytes("for", 32, 32, 100) == 401
    # Delta: l = 9, n = 10 -> E = 12, E + 1 + (8 - 12 % 8) = 17
    assert oracle_mod.base_block_bytes("delta", 9, 64, 10) == 1 + 8 + 13 + 4
    assert oracle_mod.base_block_bytes("delta", 0, 32, 7) == 5


def test_for_width_minus_one_deviation(oracle_mod):
    # FOR's decodeArray8 copies an l = w-1 block as raw values (:88); it is packed
    for width, dt in ((32, np.uint32), (64, np.uint64)):
        y = np.array([0, 1 << (width - 2), 5, (1 << (width - 2)) + 7], dt)
        blk = oracle_mod.base_encode_block("for", y)
        assert blk[0] == width - 1
        assert np.array_equal(oracle_mod.base_decode_block("for", blk, 4, dt), y)
        assert not np.array_equal(oracle_mod.base_decode_block("for", blk, 4, dt, ref_compat=True), y)


@pytest.mark.parametrize("width", [32, 64])
def test_for_filters_equal_brute_force(oracle_mod, width):
    rng = np.random.default_rng(width)
    dt = np.uint32 if width == 32 else np.uint64
    vals = np.concatenate([rng.integers(0, 1 << (width - 4), 5000, dtype=np.uint64),
                           np.full(700, 12345, np.uint64),
                           rng.integers(0, (1 << width) - 1, 800, dtype=np.uint64)]).astype(dt)
    st, off = oracle_mod.base_encode_column("for", vals, 300)
    v = vals.astype(np.uint64)
    for f in (0, 12344, 12345, int(np.median(v)), (1 << width) - 1):
        got = oracle_mod.for_filter_column(st, off, len(vals), 300, dt, filt=f)
        assert np.array_equal(got, np.nonzero(v > np.uint64(f))[0].astype(np.uint32))
    for base, f1, f2 in ((1000, 100, 700), (7, 0, 3), ((1 << 31) - 1, 5, 1 << 30)):
        got = oracle_mod.for_filter_column(st, off, len(vals), 300, dt, close=True, f1=f1, f2=f2, base=base)
        r = v % np.uint64(base)
        assert np.array_equal(got, np.nonzero((r > f1) & (r < f2))[0].astype(np.uint32))
