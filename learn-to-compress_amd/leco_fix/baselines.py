"""The fixed-partition baselines FOR_int<T> and Delta_int<T>
(headers/FOR_integer_template.h, delta_integer_template.h) on the gfx950
library: BaseColumn for HBM-resident columns, FOR_int / Delta_int for the
reference's one-block host surface.  Streams are the reference encoders' bytes
(Delta's last packed field, which the reference reads past its delta vector,
is zero); FOR decodes width-(w-1) blocks as packed and Delta decodes raw
blocks as raw, where the reference's decoders are wrong."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import LECO_FIX_CODEC_DELTA, LECO_FIX_CODEC_FOR, LECO_FIX_TAIL_PAD, check, encoded_size, lib
from .column import _STORAGE, _as_storage, _stream_handle, width_of

CODECS = {"for": LECO_FIX_CODEC_FOR, "delta": LECO_FIX_CODEC_DELTA}


class BaseColumn:
    """A FOR or Delta column in HBM: stream + block offsets."""

    def __init__(self, codec: str, stream, offsets, n: int, block_size: int, width: int, nbytes: int):
        self.codec, self.stream, self.offsets = codec, stream, offsets
        self.n, self.block_size, self.width, self.nbytes = n, block_size, width, nbytes

    @property
    def num_blocks(self) -> int:
        return (self.n + self.block_size - 1) // self.block_size

    @property
    def compression_ratio(self) -> float:
        return self.nbytes / (self.n * self.width // 8)

    def _sfx(self):
        return "u32" if self.width == 32 else "u64"

    @classmethod
    def encode(cls, codec: str, values: torch.Tensor, block_size: int, stream=None) -> "BaseColumn":
        values = _as_storage(values.contiguous())
        w = width_of(values)
        n = values.numel()
        L = lib()
        bound = int(L.leco_fix_base_stream_bound(CODECS[codec], n, block_size, w))
        st = torch.empty(bound, dtype=torch.uint8, device=values.device)
        offs = torch.empty((n + block_size - 1) // block_size + 1, dtype=torch.int64, device=values.device)
        fn = getattr(L, f"leco_fix_base_encode_column_{'u32' if w == 32 else 'u64'}")
        check(fn(CODECS[codec], values.data_ptr(), n, block_size, st.data_ptr(), bound, offs.data_ptr(),
                 _stream_handle(stream)), f"{codec} encode_column")
        return cls(codec, st, offs, n, block_size, w, encoded_size(int(offs[-1].item()), "base_encode_column"))

    def stream_bytes(self) -> bytes:
        return self.stream[:self.nbytes].cpu().numpy().tobytes()

    def decode(self, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        if out is None:
            out = torch.empty(self.n, dtype=_STORAGE[self.width], device=self.stream.device)
        fn = getattr(lib(), f"leco_fix_base_decode_column_{self._sfx()}")
        check(fn(CODECS[self.codec], self.stream.data_ptr(), self.offsets.data_ptr(), self.n, self.block_size,
                 out.data_ptr(), _stream_handle(stream)), f"{self.codec} decode_column")
        return out

    def lookup(self, idx: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        idx = idx.contiguous().to(torch.int64)
        if out is None:
            out = torch.empty(idx.numel(), dtype=_STORAGE[self.width], device=self.stream.device)
        fn = getattr(lib(), f"leco_fix_base_lookup_{self._sfx()}")
        check(fn(CODECS[self.codec], self.stream.data_ptr(), self.offsets.data_ptr(), self.n, self.block_size,
                 idx.data_ptr(), idx.numel(), out.data_ptr(), _stream_handle(stream)), f"{self.codec} lookup")
        return out

    def _for_filter(self, close: bool, f1: int, f2: int, base: int, capacity, stream):
        if self.codec != "for":
            raise ValueError("the reference defines filters for FOR_int only")
        L = lib()
        dev = self.stream.device
        cap = self.n if capacity is None else capacity
        out = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        res = torch.zeros(2, dtype=torch.int64, device=dev)
        wb = int(L.leco_fix_filter_workspace_bytes(self.n, self.block_size))
        work = torch.empty(wb, dtype=torch.uint8, device=dev)
        if close:
            rc = getattr(L, f"leco_fix_for_filter_range_close_{self._sfx()}")(
                self.stream.data_ptr(), self.offsets.data_ptr(), self.n, self.block_size, f1, f2, base,
                out.data_ptr(), cap, res.data_ptr(), work.data_ptr(), wb, _stream_handle(stream))
        else:
            rc = getattr(L, f"leco_fix_for_filter_range_{self._sfx()}")(
                self.stream.data_ptr(), self.offsets.data_ptr(), self.n, self.block_size, f1, out.data_ptr(), cap,
                res.data_ptr(), work.data_ptr(), wb, _stream_handle(stream))
        check(rc, "for filter")
        count = int(res[0].item())
        return out[:min(count, cap)], count

    def filter_range(self, filt: int, capacity: int | None = None, stream=None) -> torch.Tensor:
        """FOR_int<T>::filter_range over every block: positions (uint32 as int32) of values > filt."""
        return self._for_filter(False, filt, 0, 1, capacity, stream)[0]

    def filter_range_close(self, filter1: int, filter2: int, base: int, capacity: int | None = None,
                           stream=None) -> torch.Tensor:
        """FOR_int<T>::filter_range_close: positions of values with v % base in (filter1, filter2)."""
        return self._for_filter(True, filter1, filter2, base, capacity, stream)[0]


class _BaseHost:
    codec = ""

    def __init__(self, dtype=np.uint64):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.uint32), np.dtype(np.uint64)):
            raise TypeError(f"{type(self).__name__} holds uint32 or uint64 values")
        self.width = 8 * self.dtype.itemsize
        self.blocks = self.block_size = 0
        self.mul_add_diff_set: list = []
        self._sfx = "u32" if self.width == 32 else "u64"

    def init(self, blocks: int, block_size: int) -> None:
        self.blocks, self.block_size = blocks, block_size

    def encodeArray8_int(self, data, length: int, res=None, nvalue: int = 0) -> bytes:
        y = np.ascontiguousarray(np.asarray(data, dtype=self.dtype)[:length])
        cap = int(lib().leco_fix_base_stream_bound(CODECS[self.codec], length, length, self.width))
        buf = np.empty(cap, np.uint8)
        written = ctypes.c_uint64()
        check(getattr(lib(), f"leco_fix_base_encode_block_{self._sfx}")(
            CODECS[self.codec], y.ctypes.data, length, buf.ctypes.data, cap, ctypes.byref(written)),
            "encodeArray8_int")
        out = buf[:written.value].tobytes()
        if res is not None:
            res[:len(out)] = out
        return out

    def decodeArray8(self, in_, length: int, out=None, nvalue: int = 0) -> np.ndarray:
        src = np.frombuffer(bytes(in_), np.uint8)
        dst = np.empty(max(length, 2), self.dtype)
        check(getattr(lib(), f"leco_fix_base_decode_block_{self._sfx}")(
            CODECS[self.codec], src.ctypes.data, length, dst.ctypes.data), "decodeArray8")
        if out is not None:
            out[:length] = dst[:length]
            return out
        return dst[:length]

    def randomdecodeArray8(self, in_, to_find: int, out=None, nvalue: int = 0) -> int:
        src = np.frombuffer(bytes(in_), np.uint8)
        v = np.zeros(1, self.dtype)
        check(getattr(lib(), f"leco_fix_base_random_decode_{self._sfx}")(
            CODECS[self.codec], src.ctypes.data, to_find + 1, to_find, v.ctypes.data), "randomdecodeArray8")
        return int(v[0])

    def get_block_nums(self) -> int:
        return self.blocks


class FOR_int(_BaseHost):  # noqa: N801 - the reference's class name
    codec = "for"

    def filter_range(self, in_, length: int, filt: int, out=None, nvalue: int = 0) -> np.ndarray:
        src = np.frombuffer(bytes(in_), np.uint8)
        dst = np.empty(max(length, 1), np.uint32)
        cnt = ctypes.c_uint64()
        check(getattr(lib(), f"leco_fix_for_filter_block_{self._sfx}")(
            src.ctypes.data, length, filt, dst.ctypes.data, len(dst), nvalue, self.block_size, ctypes.byref(cnt)),
            "filter_range")
        return dst[:cnt.value]

    def filter_range_close(self, in_, length: int, out=None, nvalue: int = 0, filter1: int = 0, filter2: int = 0,
                           base_val: int = 1) -> np.ndarray:
        src = np.frombuffer(bytes(in_), np.uint8)
        dst = np.empty(max(length, 1), np.uint32)
        cnt = ctypes.c_uint64()
        check(getattr(lib(), f"leco_fix_for_filter_close_block_{self._sfx}")(
            src.ctypes.data, length, dst.ctypes.data, len(dst), nvalue, self.block_size, filter1, filter2, base_val,
            ctypes.byref(cnt)), "filter_range_close")
        return dst[:cnt.value]


class Delta_int(_BaseHost):  # noqa: N801 - the reference's class name
    codec = "delta"
