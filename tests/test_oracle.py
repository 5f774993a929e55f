"""CPU checks that pin the oracle (oracle/leco_fix_oracle.c) before it is
trusted as the GPU library's checker.

- golden fixtures written by the exact-arithmetic spec (tests/spec_leco.py);
- the spec itself on seeded random blocks of every size class;
- closed forms of the least-squares fit, size formulas, bit-width invariants;
- the reference's own CPU workload (10M uint32, block size 10,000): round trip,
  lookups and compression ratio.
"""
import math
import struct

import numpy as np
import pytest

import spec_leco as spec
from conftest import golden_cases, load_golden

CASES = [c["name"] for c in golden_cases()]


@pytest.mark.parametrize("name", CASES)
def test_oracle_stream_matches_golden(oracle_mod, name):
    g = load_golden(name)
    stream, offs = oracle_mod.encode_column(g["values"], int(g["bs"]))
    assert np.array_equal(offs, g["offsets"])
    assert stream.tobytes() == g["stream"].tobytes()


@pytest.mark.parametrize("name", CASES)
def test_oracle_decode_and_lookup_golden(oracle_mod, name):
    g = load_golden(name)
    vals, bs = g["values"], int(g["bs"])
    out = oracle_mod.decode_column(g["stream"], g["offsets"], len(vals), bs, vals.dtype)
    assert np.array_equal(out, vals)
    got = oracle_mod.lookup_column(g["stream"], g["offsets"], len(vals), bs, g["lookup_idx"], vals.dtype)
    assert np.array_equal(got, g["lookup_val"])


@pytest.mark.parametrize("name", CASES)
def test_refill_decode_loop_matches_golden(oracle_mod, name):
    """The timing loop shaped like read_all_bit_fix (bit_read.h:222-275):
    64-bit words into a 128-bit register.  It must decode every golden
    stream to its values (the checker's byte-window decode does the same)."""
    g = load_golden(name)
    vals, bs = g["values"], int(g["bs"])
    out = oracle_mod.decode_column(g["stream"], g["offsets"], len(vals), bs, vals.dtype, refill=True)
    assert np.array_equal(out, vals)


def test_refill_decode_loop_matches_window_decode_on_every_width(oracle_mod):
    """Blocks forced to every width l = 0..W (a block's residual range sets
    l), odd lengths, so fields straddle the 64-bit words in every way."""
    rng = np.random.default_rng(12)
    for width in (32, 64):
        dt = np.uint32 if width == 32 else np.uint64
        for l in range(0, width + 1):
            for n in (1, 3, 64, 65, 1023):
                base = np.arange(n, dtype=np.float64) * 97.0 + 1e6
                amp = 0 if l <= 1 else float(2 ** (l - 2))
                noise = rng.integers(-int(amp), int(amp) + 1, n) if amp else np.zeros(n, np.int64)
                v = np.clip(base + noise, 0, float((1 << width) - 1)).astype(np.uint64).astype(dt)
                if l >= width - 1:
                    v = rng.integers(0, (1 << width) - 1, n, dtype=np.uint64, endpoint=True).astype(dt)
                st, off = oracle_mod.encode_column(v, n)
                a = oracle_mod.decode_column(st, off, n, n, dt)
                b = oracle_mod.decode_column(st, off, n, n, dt, refill=True)
                assert np.array_equal(a, v) and np.array_equal(b, v), (width, l, n)


def _random_block(rng, width, n, kind):
    top = (1 << width) - 1
    if kind == "uniform":
        v = rng.integers(0, top, n, dtype=np.uint64, endpoint=True)
    elif kind == "sorted":
        v = np.sort(rng.integers(0, top >> int(rng.integers(0, width - 1)), n, dtype=np.uint64, endpoint=True))
    elif kind == "linear":
        a = int(rng.integers(0, top >> 2))
        b = float(rng.uniform(0, 2.0 ** (width - 12)))
        noise = rng.normal(0, 2.0 ** int(rng.integers(0, 20)), n)
        v = np.clip(a + b * np.arange(n) + noise, 0, float(top) * 0.99).astype(np.uint64)
    elif kind == "high":
        v = rng.integers(top >> 1, top, n, dtype=np.uint64, endpoint=True)
    else:  # "const"
        v = np.full(n, rng.integers(0, top, dtype=np.uint64, endpoint=True), np.uint64)
    return v.astype(np.uint32 if width == 32 else np.uint64)


def test_oracle_matches_spec_on_random_blocks(oracle_mod):
    rng = np.random.default_rng(5)
    for trial in range(400):
        width = 32 if trial % 2 else 64
        n = int(rng.choice([1, 2, 3, 7, 8, 63, 64, 65, 255, 256, 257, 300]))
        kind = ["uniform", "sorted", "linear", "high", "const"][trial % 5]
        y = _random_block(rng, width, n, kind)
        want = spec.encode_block([int(v) for v in y], width)
        got = oracle_mod.encode_block(y)
        assert got == want, (trial, width, n, kind)
        dec = oracle_mod.decode_block(got, n, y.dtype)
        assert np.array_equal(dec, y)
        for j in rng.integers(0, n, 4):
            assert oracle_mod.lookup_block(got, int(j), y.dtype) == int(y[j])


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("m", [2, 3, 10, 256, 1000])
def test_fit_exact_line_closed_form(oracle_mod, width, m):
    # y = a + b*i with all partial sums below 2^53: the fit must return b and a
    # exactly, and the block must have l = 0 with theta0 = a stored.
    a, b = 987654, 321
    y = (a + b * np.arange(m)).astype(np.uint32 if width == 32 else np.uint64)
    t0, t1 = oracle_mod.fit(y)
    assert t1 == float(b) and t0 == float(a)
    blk = oracle_mod.encode_block(y)
    assert blk[0] == 0 and len(blk) == 17
    assert struct.unpack("<dd", blk[1:17]) == (float(a), float(b))


def test_fit_constant_block(oracle_mod):
    # exact regime (sum of i*y < 2^53): slope exactly 0, l = 0
    y = np.full(500, 123456789, np.uint64)
    t0, t1 = oracle_mod.fit(y)
    assert t1 == 0.0 and t0 == 123456789.0
    assert oracle_mod.encode_block(y)[0] == 0
    # rounding regime: sum of i*y = 1.5e16 > 2^53, so the sequential FP64
    # chain rounds and the fitted slope is a tiny nonzero number; the stream
    # must still match the spec and round-trip
    y = np.full(500, 123456789012, np.uint64)
    t0, t1 = oracle_mod.fit(y)
    assert t1 != 0.0 and abs(t1) < 1e-6
    blk = oracle_mod.encode_block(y)
    assert blk == spec.encode_block([int(v) for v in y], 64)
    assert np.array_equal(oracle_mod.decode_block(blk, 500, np.uint64), y)


def test_fit_agrees_with_least_squares(oracle_mod):
    # textbook OLS (numpy polyfit) within 1e-9 relative: catches a dropped or
    # sign-flipped term of the normal equations in any regime
    rng = np.random.default_rng(9)
    for m in [5, 100, 4096, 10000]:
        y = np.sort(rng.integers(0, 1 << 40, m, dtype=np.uint64))
        t0, t1 = oracle_mod.fit(y)
        b, a = np.polyfit(np.arange(m, dtype=np.float64), y.astype(np.float64), 1)
        assert math.isclose(t1, b, rel_tol=1e-9, abs_tol=1e-6)
        assert math.isclose(t0, a, rel_tol=1e-9, abs_tol=1e-3 * abs(b) + 1e-3)


def test_one_value_block_nan_bits(oracle_mod):
    blk = oracle_mod.encode_block(np.array([5], np.uint32))
    assert blk[0] == 4  # mag = (5 - INT64_MIN) mod 2^32 = 5 -> bit length 3, +1
    assert struct.unpack("<QQ", blk[1:17]) == (spec.X86_DEFAULT_NAN_BITS, spec.X86_DEFAULT_NAN_BITS)
    assert oracle_mod.decode_block(blk, 1, np.uint32)[0] == 5


def test_cvtt64_semantics(oracle_mod):
    imin = -(1 << 63)
    assert oracle_mod.cvtt64(float("nan")) == imin
    assert oracle_mod.cvtt64(2.0 ** 63) == imin
    assert oracle_mod.cvtt64(-(2.0 ** 63)) == imin
    assert oracle_mod.cvtt64(1e300) == imin
    assert oracle_mod.cvtt64(-1.5) == -1
    assert oracle_mod.cvtt64(2.0 ** 63 - 1024) == (1 << 63) - 1024
    for x in [0.0, 0.99, -0.99, 12345.7, -98765.2, 2.0 ** 52 + 0.5]:
        assert oracle_mod.cvtt64(x) == spec.cvtt64(x)


@pytest.mark.parametrize("name", CASES)
def test_block_sizes_and_residual_bounds(name):
    # A7: size = 1 + n*w/8 (raw) or 17 + 8*floor(E/8) + 9 with E = ceil(l*n/8);
    # every stored magnitude < 2^(l-1); bytes after the last field are zero.
    g = load_golden(name)
    vals, bs = g["values"], int(g["bs"])
    width = 8 * vals.dtype.itemsize
    stream, offs = g["stream"].tobytes(), [int(x) for x in g["offsets"]]
    for b in range(len(offs) - 1):
        n = min(bs, len(vals) - b * bs)
        blk = stream[offs[b]:offs[b + 1]]
        l = blk[0]
        if l >= width:
            assert len(blk) == 1 + n * width // 8
            continue
        e = (l * n + 7) // 8
        assert len(blk) == (17 + 8 * (e // 8) + 9 if l else 17)
        if l:
            code = int.from_bytes(blk[17:], "little")
            assert code >> (l * n) == 0
            mags = [(code >> (k * l)) & ((1 << (l - 1)) - 1) for k in range(n)]
            assert max(mags) < (1 << (l - 1))
            # l is minimal: the largest magnitude needs l-1 bits
            assert max(mags).bit_length() == l - 1


@pytest.mark.parametrize("name", ["lw1_u32", "lw1_u64"])
def test_reference_decoder_deviation_on_width_minus_one(oracle_mod, name):
    # The reference decodeArray8 copies an l == w-1 block as raw bytes
    # (piecewise_fix_integer_template.h:171-177); this codec decodes it.
    g = load_golden(name)
    vals = g["values"]
    blk = g["stream"].tobytes()
    assert blk[0] == 8 * vals.dtype.itemsize - 1
    good = oracle_mod.decode_block(blk, len(vals), vals.dtype)
    bad = oracle_mod.decode_block(blk, len(vals), vals.dtype, ref_compat=True)
    assert np.array_equal(good, vals)
    assert not np.array_equal(bad, vals)
    assert bad.tobytes() == (blk[1:] + bytes(len(vals) * vals.dtype.itemsize))[:len(vals) * vals.dtype.itemsize]


def test_threaded_encode_equals_serial(oracle_mod):
    from leco_fix import datagen
    y = datagen.sorted_pwl_u64(200_000, seed=3)
    s1, o1 = oracle_mod.encode_column(y, 4096, threads=1)
    s4, o4 = oracle_mod.encode_column(y, 4096, threads=4)
    assert np.array_equal(o1, o4) and s1.tobytes() == s4.tobytes()
    out = oracle_mod.decode_column(s4, o4, len(y), 4096, np.uint64, threads=4)
    assert np.array_equal(out, y)


def test_reference_cpu_workload(oracle_mod):
    # 10M uint32 values y_i = clamp(round(1000 i + N(0, 50^2))), block size
    # 10,000: full encode, full decode, 1M uniform lookups, round trip exact.
    from leco_fix import datagen
    n, bs = 10_000_000, 10_000
    y = datagen.linear_noise_u32(n, seed=1)
    stream, offs = oracle_mod.encode_column(y, bs, threads=8)
    assert len(offs) == 1001
    cr = len(stream) / (4 * n)
    # ~429 blocks at l = 9 (11,274 B each), ~570 constant blocks of 17 B
    assert 0.11 < cr < 0.13
    out = oracle_mod.decode_column(stream, offs, n, bs, np.uint32, threads=8)
    assert np.array_equal(out, y)
    out = oracle_mod.decode_column(stream, offs, n, bs, np.uint32, threads=8, refill=True)
    assert np.array_equal(out, y)
    idx = np.random.default_rng(1).integers(0, n, 1_000_000, dtype=np.uint64)
    got = oracle_mod.lookup_column(stream, offs, n, bs, idx, np.uint32, threads=8)
    assert np.array_equal(got, y[idx])
