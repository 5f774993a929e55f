"""Executable spec of the FOR_int<T> / Delta_int<T> fixed-partition block
formats (headers/FOR_integer_template.h:26-250, delta_integer_template.h:26-160,
bit_write.h:258-390), written independently of the C oracle: whole payloads
are one Python integer cut with to_bytes, arithmetic is on Python ints with
explicit masks.  Test infrastructure only."""
from __future__ import annotations


def _bitlen(v: int) -> int:
    return v.bit_length()


def _trunc_div(a: int, b: int) -> int:
    """C integer division (toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _pack(fields: list[int], l: int, nbytes: int) -> bytes:
    acc = 0
    for k, f in enumerate(fields):
        acc |= (f & ((1 << l) - 1)) << (k * l)
    return acc.to_bytes(nbytes, "little")


def for_encode(y: list[int], width: int) -> bytes:
    tb = width // 8
    mn, mx = min(y), max(y)
    l = min(_bitlen(mx - mn), width)
    if l == width:
        return bytes([l]) + b"".join(v.to_bytes(tb, "little") for v in y)
    out = bytes([l]) + mn.to_bytes(tb, "little") + mx.to_bytes(tb, "little")
    if l:
        e = (l * len(y) + 7) // 8
        pad = max(0, _trunc_div(32 - l, 8))   # write_FOR_int_T: ceil((32 - l) / 8) in int arithmetic
        out += _pack([v - mn for v in y], l, e + 1) + bytes(pad)
    return out


def for_decode(blk: bytes, n: int, width: int) -> list[int]:
    tb = width // 8
    l = blk[0]
    if l >= width:
        return [int.from_bytes(blk[1 + i * tb:1 + (i + 1) * tb], "little") for i in range(n)]
    base = int.from_bytes(blk[1:1 + tb], "little")
    if not l:
        return [base] * n
    acc = int.from_bytes(blk[1 + 2 * tb:], "little")
    return [(base + ((acc >> (i * l)) & ((1 << l) - 1))) & ((1 << width) - 1) for i in range(n)]


def delta_encode(y: list[int], width: int) -> bytes:
    tb = width // 8
    n = len(y)
    d = [(abs(y[i + 1] - y[i]), y[i + 1] > y[i]) for i in range(n - 1)]
    mxd = max((m for m, _ in d), default=0)
    l = min(_bitlen(mxd) + 1, width) if mxd else 0
    if l == width:
        return bytes([l]) + b"".join(v.to_bytes(tb, "little") for v in y)
    out = bytes([l]) + y[0].to_bytes(tb, "little")
    if l:
        fields = [(m & ((1 << (l - 1)) - 1)) | (int(s) << (l - 1)) for m, s in d] + [0]  # n fields, the last zero
        e = (l * n + 7) // 8
        out += _pack(fields, l, e + 1) + bytes(8 - e % 8)
    return out


def delta_decode(blk: bytes, n: int, width: int) -> list[int]:
    tb = width // 8
    l = blk[0]
    if l >= width:
        return [int.from_bytes(blk[1 + i * tb:1 + (i + 1) * tb], "little") for i in range(n)]
    v = int.from_bytes(blk[1:1 + tb], "little")
    out = [v]
    acc = int.from_bytes(blk[1 + tb:], "little") if l else 0
    for i in range(n - 1):
        if l:
            f = (acc >> (i * l)) & ((1 << l) - 1)
            mag = f & ((1 << (l - 1)) - 1)
            v = (v + mag if f >> (l - 1) else v - mag) & ((1 << width) - 1)
        out.append(v)
    return out
