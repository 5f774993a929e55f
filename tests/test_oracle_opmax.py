"""The oracle's piecewise_fix_op_max restatement, pinned (CPU only):

- against an independent exact-arithmetic spec (tests/spec_opmax.py), block
  bytes and fix-up lists for every shape the codec has: packed widths 2..30,
  raw blocks (tmp_bit 31 and 32), l = 0, one- and two-value blocks, negative
  predictions, values near 2^32, tie-prone float slopes;
- against what the reference harness checks (experiments/fix_int.cpp:95-135):
  decode + fix-up lists == input, and random access == input;
- against closed forms: an exact arithmetic progression encodes to a 13-byte
  block whose header holds the line, and a chain with no rounding needs no fix-ups.
"""
import struct

import numpy as np
import pytest

from oracle import oracle
import spec_opmax as so

RNG_SEED = 20240601


def _blocks():
    rng = np.random.default_rng(RNG_SEED)
    out = []
    i = np.arange(2000, dtype=np.float64)
    out.append(("linear_noise", np.round(1e6 + 21.3 * i + rng.integers(-3, 4, 2000)).astype(np.uint32)))
    out.append(("high_drift", np.round(3.0e9 + 2.7 * i).astype(np.uint32)))              # chain rounding -> fix-ups
    out.append(("sorted_uniform", np.sort(rng.integers(0, 2**32, 700, dtype=np.uint64)).astype(np.uint32)))
    out.append(("uniform_raw", rng.integers(0, 2**32, 300, dtype=np.uint64).astype(np.uint32)))
    out.append(("uniform_31", rng.integers(0, 2**31, 300, dtype=np.uint64).astype(np.uint32)))
    out.append(("const", np.full(500, 77, np.uint32)))
    out.append(("one", np.array([123456789], np.uint32)))
    out.append(("two", np.array([5, 4000000000], np.uint32)))
    out.append(("near_zero", (np.abs(np.round(50 - 0.37 * i[:400])) + rng.integers(0, 3, 400)).astype(np.uint32)))
    out.append(("desc_high", np.round(4.2e9 - 1000.25 * i[:1000]).astype(np.uint32)))
    out.append(("zigzag", (1000 + (np.arange(600) % 2) * 900).astype(np.uint32)))
    # long blocks whose chain drifts (float slope bits below the ulp of p): fix-ups both ways
    j = np.arange(60000, dtype=np.float64)
    out.append(("drift_plus", np.round(4.1e9 + 0.0271828 * j).astype(np.uint32)))
    out.append(("drift_minus", np.round(4.0e9 + 0.1234567 * j).astype(np.uint32)))
    for w in (2, 3, 7, 8, 13, 16, 24, 29):
        base = np.round(5e8 + 123.456 * i[:513])
        noise = rng.integers(-(1 << (w - 2)), 1 << (w - 2), 513) if w > 2 else rng.integers(-1, 2, 513)
        out.append((f"width{w}", (base + noise).astype(np.uint32)))
    return out


BLOCKS = _blocks()


@pytest.mark.parametrize("name,y", BLOCKS, ids=[b[0] for b in BLOCKS])
def test_encode_matches_spec(name, y):
    blk, fix = oracle.opmax_encode_block(y)
    sblk, sfix = so.encode_block([int(v) for v in y])
    assert blk == sblk
    assert fix.tolist() == sfix
    assert len(blk) == oracle.opmax_block_bytes(blk[0], len(y)) == so.block_bytes(blk[0], len(y))


@pytest.mark.parametrize("name,y", BLOCKS, ids=[b[0] for b in BLOCKS])
def test_decode_and_lookup(name, y):
    blk, fix = oracle.opmax_encode_block(y)
    dec = oracle.opmax_decode_block(blk, len(y))
    assert dec.tolist() == so.decode_block(blk, len(y))
    l = blk[0]
    if l < 31:  # the harness's check: decodeArray8 + fix-ups == input
        assert ((dec.astype(np.int64) + fix) % 2**32).astype(np.uint32).tolist() == y.tolist()
    for j in range(0, len(y), max(1, len(y) // 97)):
        v = oracle.opmax_lookup_block(blk, j)
        assert v == so.lookup_block(blk, j) == int(y[j])


def test_widths_and_raw_cover_every_shape():
    seen = {oracle.opmax_encode_block(y)[0][0] for _, y in BLOCKS}
    assert {0, 2, 31, 32} <= seen
    assert len([l for l in seen if 2 <= l <= 30]) >= 8


def test_fixups_are_exercised():
    fixes = [oracle.opmax_encode_block(y)[1] for _, y in BLOCKS]
    assert sum(int((f > 0).sum()) for f in fixes) > 0
    assert sum(int((f < 0).sum()) for f in fixes) > 0


def test_exact_line_closed_form():
    y = (1000 + 4 * np.arange(64)).astype(np.uint32)
    blk, fix = oracle.opmax_encode_block(y)
    assert blk == bytes([0]) + struct.pack("<d", 1000.0) + struct.pack("<f", 4.0)
    assert not fix.any()
    assert oracle.opmax_decode_block(blk, 64).tolist() == y.tolist()


def test_single_value_block_nan_model():
    blk, fix = oracle.opmax_encode_block(np.array([7], np.uint32))
    # lr.h:103-108 on one value is 0/0: x86's default NaN in theta0, its float narrowing in theta1
    assert blk[1:9] == struct.pack("<Q", 0xFFF8000000000000)
    assert blk[9:13] == struct.pack("<I", 0xFFC00000)
    assert oracle.opmax_decode_block(blk, 1).tolist() == [7]
    assert oracle.opmax_lookup_block(blk, 0) == 7


@pytest.mark.parametrize("bs", [1, 7, 200, 2000])
def test_column_and_harness_lists(bs):
    rng = np.random.default_rng(bs)
    n = 6000
    y = np.round(2.9e9 + 3.3 * np.arange(n) + rng.integers(-40, 40, n)).astype(np.uint32)
    stream, offs, fix = oracle.opmax_encode_column(y, bs)
    assert offs[-1] == len(stream)
    dec = oracle.opmax_decode_column(stream, offs, n, bs)
    plus, minus = oracle.opmax_fix_lists(fix)
    assert oracle.opmax_apply_lists(dec, plus, minus).tolist() == y.tolist()
    if bs >= 200:
        s_stream, s_offs, s_plus, s_minus = so.encode_column([int(v) for v in y[:2 * bs]], bs)
        k = int(offs[2])
        assert stream[:k].tobytes() == s_stream and offs[:3].tolist() == s_offs
        assert plus[plus < 2 * bs].tolist() == s_plus and minus[minus < 2 * bs].tolist() == s_minus
    idx = rng.integers(0, n, 300).astype(np.uint64)
    assert oracle.opmax_lookup_column(stream, offs, n, bs, idx).tolist() == y[idx].tolist()


def _golden_opmax():
    import json
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(here, "opmax_cases.json")) as f:
        return [c["name"] for c in json.load(f)]


@pytest.mark.parametrize("name", _golden_opmax())
def test_oracle_matches_golden(name):
    from conftest import load_golden
    g = load_golden(name)
    y, bs = g["values"], int(g["bs"])
    stream, offs, fix = oracle.opmax_encode_column(y, bs)
    assert stream.tobytes() == g["stream"].tobytes()
    assert offs.tolist() == g["offsets"].tolist()
    plus, minus = oracle.opmax_fix_lists(fix)
    assert plus.tolist() == g["plus"].tolist() and minus.tolist() == g["minus"].tolist()
    dec = oracle.opmax_decode_column(stream, offs, len(y), bs)
    assert dec.tolist() == g["decoded"].tolist()
    fixed = oracle.opmax_apply_lists(dec, plus, minus)
    raw_pos = np.concatenate([np.arange(b * bs, min((b + 1) * bs, len(y)))
                              for b in range(len(offs) - 1) if stream[int(offs[b])] >= 31] or [np.array([], int)])
    ok = np.ones(len(y), bool)
    ok[raw_pos.astype(np.int64)] = False
    assert (fixed[ok] == y[ok]).all()
