"""Executable specification of the LeCo_fix block format, in exact arithmetic.

An independent restatement of Codecset::Leco_int<T> (reference
headers/piecewise_fix_integer_template.h:27-140,142-204,305-337,
headers/lr.h:66-85, headers/bit_write.h:257-323, headers/bit_read.h:222-275,
1170-1199), written differently from the C oracle on purpose, so that a slip
in one (an index, a shift, a dropped term) shows up as a mismatch:

- every fused multiply-add is computed with exact rationals and one
  correctly rounded conversion (Fraction -> float), not with libm's fma;
- the payload is built as one Python big integer and cut into bytes with
  int.to_bytes, not with a 128-bit shift register;
- int64/uint64 wrapping is explicit modular arithmetic.

Pure Python: meant for blocks of up to a few thousand values.
"""
from __future__ import annotations

import math
import struct
from fractions import Fraction

TWO63 = 1 << 63
TWO64 = 1 << 64
X86_DEFAULT_NAN_BITS = 0xFFF8000000000000


def f2b(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def b2f(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def fma(a: float, b: float, c: float) -> float:
    """IEEE fusedMultiplyAdd, round to nearest even, for finite inputs."""
    if math.isnan(a) or math.isnan(b) or math.isnan(c):
        return math.nan
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def cvtt64(x: float) -> int:
    """x86 CVTTSD2SI r64: truncate; NaN or outside [-2^63, 2^63) -> INT64_MIN."""
    if math.isnan(x) or not (-9223372036854775808.0 <= x < 9223372036854775808.0):
        return -TWO63
    return math.trunc(x)


def s64(v: int) -> int:
    """Wrap an integer to int64."""
    v %= TWO64
    return v - TWO64 if v >= TWO63 else v


def fit(y: list[int]) -> tuple[int, int]:
    """lr_int_T<T>::caltheta under the FMA recipe -> (theta0 bits, theta1 bits)."""
    sx = sy = sxx = sxy = 0.0
    for i, v in enumerate(y):
        xi, yi = float(i), float(v)          # int -> double rounds to nearest even
        sx = sx + xi
        sy = sy + yi
        sxx = fma(xi, xi, sxx)
        sxy = fma(xi, yi, sxy)
    md = float(len(y))
    ccc = fma(sxy, md, -(sx * sy))
    xxx = fma(sxx, md, -(sx * sx))
    if xxx == 0.0:
        return X86_DEFAULT_NAN_BITS, X86_DEFAULT_NAN_BITS
    t1 = ccc / xxx
    t0 = fma(-t1, sx, sy) / md
    return f2b(t0), f2b(t1)


def predict(t0: float, t1: float, i: int) -> int:
    return cvtt64(fma(t1, float(i), t0))


def encode_block(y: list[int], width: int) -> bytes:
    """Leco_int<T>::encodeArray8_int for one block of T = uint{width}."""
    y = [int(v) for v in y]
    t0b, t1b = fit(y)
    t0, t1 = b2f(t0b), b2f(t1b)
    d = [s64(v - predict(t0, t1, i)) for i, v in enumerate(y)]
    if not math.isnan(t0):
        t0 = t0 + float(s64(max(d) + min(d))) / 2.0
        t0b = f2b(t0)
    mask = (1 << width) - 1
    signs, mags = [], []
    for i, v in enumerate(y):
        p = predict(t0, t1, i)
        # uint64 compares after converting p to unsigned; uint32 compares as int64
        pos = v > (p % TWO64) if width == 64 else v > p
        mags.append(((v - p) if pos else (p - v)) & mask)
        signs.append(pos)
    mx = max(mags)
    l = min(mx.bit_length() + 1, width) if mx else 0
    if l == width:
        return bytes([l]) + b"".join(v.to_bytes(width // 8, "little") for v in y)
    out = bytes([l]) + struct.pack("<QQ", t0b, t1b)
    if l == 0:
        return out
    code = 0
    for k, (m, s) in enumerate(zip(mags, signs)):
        code |= ((m & ((1 << (l - 1)) - 1)) | (int(s) << (l - 1))) << (k * l)
    e = (l * len(y) + 7) // 8
    return out + code.to_bytes(e, "little") + bytes(8 * (e // 8) + 9 - e)


def decode_block(blk: bytes, n: int, width: int) -> list[int]:
    """Leco_int<T>::decodeArray8, except that l == width-1 blocks are decoded
    (the reference copies them as raw values, :171-177)."""
    l = blk[0]
    bpv = width // 8
    if l == 127:
        return [int.from_bytes(blk[1:1 + bpv], "little")]
    if l == 126:
        return [int.from_bytes(blk[1 + k * bpv:1 + (k + 1) * bpv], "little") for k in range(2)]
    if l >= width:
        return [int.from_bytes(blk[1 + k * bpv:1 + (k + 1) * bpv], "little") for k in range(n)]
    t0, t1 = struct.unpack("<dd", blk[1:17])
    mask = (1 << width) - 1
    if l == 0:
        return [predict(t0, t1, i) & mask for i in range(n)]
    code = int.from_bytes(blk[17:], "little")
    out = []
    for i in range(n):
        f = (code >> (i * l)) & ((1 << l) - 1)
        mag = f & ((1 << (l - 1)) - 1)
        p = predict(t0, t1, i)
        out.append((p + mag if f >> (l - 1) else p - mag) & mask)
    return out


def block_size_bytes(l: int, width: int, n: int) -> int:
    if l >= width:
        return 1 + n * width // 8
    if l == 0:
        return 17
    e = (l * n + 7) // 8
    return 17 + 8 * (e // 8) + 9


def encode_column(y: list[int], width: int, bs: int) -> tuple[bytes, list[int]]:
    """experiments/leco_int_fix_template.cpp:98-127 with a contiguous stream."""
    stream, offs = b"", [0]
    for s in range(0, len(y), bs):
        stream += encode_block(y[s:s + bs], width)
        offs.append(len(stream))
    return stream, offs
