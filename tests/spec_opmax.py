"""Executable specification of piecewise_fix_op_max, the uint32 IntegerCODEC
(reference headers/piecewise_fix_op_minimize_maxerror.h:45-197 with lr.h:89-111,
bit_write.h:186-247 write_delta_T, bit_write.h:395-410 write_delta_default,
bit_read.h:26-75 and 559-604 read_all_bit_fix_add, bit_read.h:1107-1132
read_bit_fix_T, and the fix-up loops of experiments/fix_int.cpp:110-117),
written independently of the C oracle:

- the fit and every fused multiply-add come from spec_leco (exact rationals,
  one correctly rounded conversion);
- theta1 is narrowed to float with struct's IEEE single packing;
- the payload is one Python big integer cut with int.to_bytes;
- int32 wrapping is explicit modular arithmetic.

Test infrastructure only; pure Python, for blocks of up to a few thousand values.
"""
from __future__ import annotations

import math
import struct

import spec_leco as sl

TWO32 = 1 << 32
X86_DEFAULT_NAN_F32_BITS = 0xFFC00000


def s32(v: int) -> int:
    """Wrap an integer to int32 (the reference's `int` results)."""
    v %= TWO32
    return v - TWO32 if v >= 1 << 31 else v


def to_f32_bits(t1_bits: int) -> int:
    """(float)theta1: IEEE narrowing, round to nearest even; the x86 default NaN keeps its sign."""
    x = sl.b2f(t1_bits)
    if math.isnan(x):
        return X86_DEFAULT_NAN_F32_BITS if t1_bits >> 63 else 0x7FC00000
    return struct.unpack("<I", struct.pack("<f", x))[0]


def f32(bits: int) -> float:
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def block_bytes(l: int, n: int) -> int:
    if l == 0:
        return 13
    if l >= 31:
        return 13 + 4 * n
    return 13 + (l * n + 7) // 8 + 1 + (32 - l) // 8


def encode_block(y: list[int]) -> tuple[bytes, list[int]]:
    """encodeArray8 of one block -> (bytes, fix) with fix[i] = +1 where the
    reference appends i to mul_add_diff_set, -1 for mul_add_diff_set_minus."""
    t0b, t1b = sl.fit(y)
    t0 = sl.b2f(t0b)
    t1f = to_f32_bits(t1b)
    t1 = f32(t1f)
    n = len(y)
    d1 = [s32(v - sl.cvtt64(sl.fma(t1, float(i), t0))) for i, v in enumerate(y)]
    t0 = t0 + float(s32(max(d1) + min(d1))) / 2.0
    fix, delta = [], []
    pred = t0
    for i, v in enumerate(y):
        pm = sl.cvtt64(sl.fma(t1, float(i), t0))
        pa = sl.cvtt64(pred)
        fix.append(1 if pm > pa else (-1 if pm < pa else 0))
        pred = pred + t1
        delta.append(s32(v - pm))
    max_err = max((abs(d) for d in delta if d != -(1 << 31)), default=0)
    l = max_err.bit_length() + 1 if max_err else 0
    out = bytes([l]) + struct.pack("<d", t0) + struct.pack("<I", t1f)
    if l >= 31:
        out += b"".join(v.to_bytes(4, "little") for v in y)
    elif l:
        acc = 0
        low = (1 << (l - 1)) - 1
        for k, d in enumerate(delta):
            mag = (-d) % TWO32 if d <= 0 else d
            acc |= ((mag & low) | (int(d > 0) << (l - 1))) << (k * l)
        e = (l * n + 7) // 8
        out += acc.to_bytes(e + 1, "little") + bytes((32 - l) // 8)
    assert len(out) == block_bytes(l, n)
    return out, fix


def _header(blk: bytes) -> tuple[int, float, float]:
    return blk[0], struct.unpack("<d", blk[1:9])[0], f32(struct.unpack("<I", blk[9:13])[0])


def decode_block(blk: bytes, n: int) -> list[int]:
    """decodeArray8: the decoder's output before the harness's fix-up."""
    l, t0, t1 = _header(blk)
    if l >= 31:
        return [int.from_bytes(blk[13 + 4 * i:17 + 4 * i], "little") for i in range(n)]
    acc = int.from_bytes(blk[13:], "little")
    out, pred = [], t0
    for i in range(n):
        v = sl.cvtt64(pred)
        if l:
            f = (acc >> (i * l)) & ((1 << l) - 1)
            mag = f & ((1 << (l - 1)) - 1)
            v = v + mag if f >> (l - 1) else v - mag
        out.append(v % TWO32)
        pred = pred + t1
    return out


def lookup_block(blk: bytes, j: int) -> int:
    """randomdecodeArray8(in, j)."""
    l, t0, t1 = _header(blk)
    if l >= 31:
        return int.from_bytes(blk[13 + 4 * j:17 + 4 * j], "little")
    p = sl.cvtt64(sl.fma(float(j), t1, t0))
    if not l:
        return p % TWO32
    f = (int.from_bytes(blk[13:], "little") >> (j * l)) & ((1 << l) - 1)
    mag = f & ((1 << (l - 1)) - 1)
    return (p + (mag if f >> (l - 1) else -mag)) % TWO32


def encode_column(y: list[int], bs: int) -> tuple[bytes, list[int], list[int], list[int]]:
    """The harness's block loop -> (stream, offsets, mul_add_diff_set, mul_add_diff_set_minus)."""
    stream, offs, plus, minus = b"", [0], [], []
    for b in range((len(y) + bs - 1) // bs):
        blk, fix = encode_block(y[b * bs:(b + 1) * bs])
        stream += blk
        offs.append(len(stream))
        plus += [b * bs + i for i, f in enumerate(fix) if f > 0]
        minus += [b * bs + i for i, f in enumerate(fix) if f < 0]
    return stream, offs, plus, minus
