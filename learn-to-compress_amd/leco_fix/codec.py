"""Leco_int: the reference codec surface, one block at a time on host buffers.

Mirrors Codecset::Leco_int<T> (headers/piecewise_fix_integer_template.h:14-338):
init (:22), encodeArray8_int (:27), decodeArray8 (:142), filter_range (:206),
filter_range_close (:251), randomdecodeArray8 (:305), plus the IntegerCODEC
names encodeArray8 and get_block_nums
(headers/codecs.h:35,60).  Each call stages the block into HBM and runs the
gfx950 kernels through the C ABI (leco_fix_*_block_*), synchronously.

Differences from the reference, all deliberate:
- errors raise LecoError (the reference has no error reporting);
- a block with bit width 8*sizeof(T)-1 decodes correctly (the reference's
  decodeArray8 copies it as raw bytes, :171-177);
- get_block_nums() returns the block count given to init(); Leco_int<T> has
  no such method, and the IntegerCODEC fixed-partition codecs return 1;
- filter_range / filter_range_close filter raw blocks by their values (the
  reference reads theta from the raw bytes, which is undefined), and
  filter_range_close returns every period's positions (the reference
  overwrites them at `out`, :274; its count is the same).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


class Leco_int:  # noqa: N801 - the reference's class name
    def __init__(self, dtype=np.uint64):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.uint32), np.dtype(np.uint64)):
            raise TypeError("Leco_int holds uint32 or uint64 values")
        self.width = 8 * self.dtype.itemsize
        self.blocks = 0
        self.block_size = 0
        self.mul_add_diff_set: list = []  # :18, never filled by the reference either
        sfx = "u32" if self.width == 32 else "u64"
        L = lib()
        self._enc = getattr(L, f"leco_fix_encode_block_{sfx}")
        self._dec = getattr(L, f"leco_fix_decode_block_{sfx}")
        self._ra = getattr(L, f"leco_fix_random_decode_{sfx}")
        self._flt = getattr(L, f"leco_fix_filter_block_{sfx}")
        self._fltc = getattr(L, f"leco_fix_filter_close_block_{sfx}")

    def init(self, blocks: int, block_size: int) -> None:
        self.blocks = blocks
        self.block_size = block_size

    def encodeArray8_int(self, data, length: int, res=None, nvalue: int = 0) -> bytes:
        """Encodes data[0:length] as one block; returns its bytes (and fills res if given)."""
        y = np.ascontiguousarray(np.asarray(data, dtype=self.dtype)[:length])
        cap = int(lib().leco_fix_stream_bound(length, length, self.width))
        buf = np.empty(cap, np.uint8)
        written = ctypes.c_uint64()
        check(self._enc(y.ctypes.data, length, buf.ctypes.data, cap, ctypes.byref(written)), "encodeArray8_int")
        out = buf[:written.value].tobytes()
        if res is not None:
            res[:len(out)] = out
        return out

    encodeArray8 = encodeArray8_int

    def decodeArray8(self, in_, length: int, out=None, nvalue: int = 0) -> np.ndarray:
        src = np.frombuffer(bytes(in_), np.uint8)
        dst = np.empty(max(length, 2), self.dtype)
        check(self._dec(src.ctypes.data, length, dst.ctypes.data), "decodeArray8")
        if out is not None:
            out[:length] = dst[:length]
            return out
        return dst[:length]

    def randomdecodeArray8(self, in_, to_find: int, out=None, nvalue: int = 0) -> int:
        src = np.frombuffer(bytes(in_), np.uint8)
        v = np.zeros(1, self.dtype)
        check(self._ra(src.ctypes.data, to_find, v.ctypes.data), "randomdecodeArray8")
        return int(v[0])

    def filter_range(self, in_, length: int, filt: int, out=None, block_id: int = 0) -> np.ndarray:
        """Positions block_id*block_size + i (uint32) of the block's values above filt (:206-249)."""
        src = np.frombuffer(bytes(in_), np.uint8)
        dst = np.empty(max(length, 1), np.uint32)
        cnt = ctypes.c_uint64()
        check(self._flt(src.ctypes.data, length, filt, dst.ctypes.data, len(dst), block_id, self.block_size,
                        ctypes.byref(cnt)), "filter_range")
        res = dst[:cnt.value]
        if out is not None:
            out[:len(res)] = res
        return res

    def filter_range_close(self, in_, length: int, out=None, block_id: int = 0, filter1: int = 0,
                           filter2: int = 0, base: int = 0) -> np.ndarray:
        """Positions of values in (filter1 + k*base, filter2 + k*base) within the block's model reach
        (:251-303), period by period."""
        src = np.frombuffer(bytes(in_), np.uint8)
        cap = length + 1
        while True:
            dst = np.empty(cap, np.uint32)
            cnt = ctypes.c_uint64()
            check(self._fltc(src.ctypes.data, length, dst.ctypes.data, cap, block_id, self.block_size, filter1,
                             filter2, base, ctypes.byref(cnt)), "filter_range_close")
            if cnt.value <= cap:
                break
            cap = cnt.value
        res = dst[:cnt.value]
        if out is not None:
            out[:len(res)] = res
        return res

    def get_block_nums(self) -> int:
        return self.blocks
