"""A compressed column resident in HBM, over the C ABI's column calls.

torch provides the device allocations and the stream; every byte of codec
work runs in the HIP kernels of libleco_fix.so.  Values are held in int32 /
int64 tensors and read as uint32 / uint64 (same bits).

Reference behaviour mirrored: experiments/leco_int_fix_template.cpp:98-213
partitions a column into fixed blocks (:98-106), encodes each block
(:111-127), decodes all blocks (:142-153) and looks values up by
block = idx / block_size, position = idx % block_size (:189-193).
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import LECO_FIX_TAIL_PAD, BlockParam, check, encoded_size, lib

_STORAGE = {32: torch.int32, 64: torch.int64}
DECODE_AUTO, DECODE_STAGED, DECODE_VALUES = 0, 1, 2
PARAM_BYTES = ctypes.sizeof(BlockParam)


def width_of(t: torch.Tensor) -> int:
    if t.dtype in (torch.int32, torch.uint32):
        return 32
    if t.dtype in (torch.int64, torch.uint64):
        return 64
    raise TypeError(f"columns hold 32- or 64-bit integers, not {t.dtype}")


def _stream_handle(stream) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def _as_storage(t: torch.Tensor) -> torch.Tensor:
    w = width_of(t)
    return t.view(_STORAGE[w]) if t.dtype != _STORAGE[w] else t


class Column:
    """One column: concatenated block streams + offsets + param table."""

    def __init__(self, stream: torch.Tensor, offsets: torch.Tensor, params: torch.Tensor,
                 n: int, block_size: int, width: int, nbytes: int):
        self.stream = stream          # uint8 [capacity], >= nbytes + tail pad
        self.offsets = offsets        # int64 [nb + 1] (uint64 bits)
        self.params = params          # uint8 [nb * 32] (leco_fix_block_param)
        self.n = n
        self.block_size = block_size
        self.width = width
        self.nbytes = nbytes
        self.base = 0                 # stream byte of block 0 (nonzero inside a batch arena)

    @property
    def num_blocks(self) -> int:
        return int(lib().leco_fix_block_count(self.n, self.block_size))

    @property
    def compression_ratio(self) -> float:
        return self.nbytes / (self.n * self.width // 8)

    @classmethod
    def encode(cls, values: torch.Tensor, block_size: int, stream=None, compact: bool = False) -> "Column":
        """Leco_int<T>::encodeArray8_int over every block of `values` (device tensor)."""
        if not values.is_cuda:
            raise ValueError("Column.encode takes a device tensor")
        values = _as_storage(values.contiguous())
        w = width_of(values)
        n = values.numel()
        L = lib()
        bound = int(L.leco_fix_stream_bound(n, block_size, w))
        nb = int(L.leco_fix_block_count(n, block_size))
        dev = values.device
        st = torch.empty(bound, dtype=torch.uint8, device=dev)
        offs = torch.empty(nb + 1, dtype=torch.int64, device=dev)
        prm = torch.empty(nb * PARAM_BYTES, dtype=torch.uint8, device=dev)
        fn = L.leco_fix_encode_column_u32 if w == 32 else L.leco_fix_encode_column_u64
        check(fn(values.data_ptr(), n, block_size, st.data_ptr(), bound, offs.data_ptr(), prm.data_ptr(),
                 _stream_handle(stream)), "encode_column")
        nbytes = encoded_size(int(offs[nb].item()), "encode_column")
        col = cls(st, offs, prm, n, block_size, w, nbytes)
        if compact:
            col.compact()
        return col

    def compact(self) -> None:
        """Shrinks the stream buffer to the encoded size plus the tail pad."""
        st = torch.zeros(self.nbytes + LECO_FIX_TAIL_PAD, dtype=torch.uint8, device=self.stream.device)
        st[:self.nbytes].copy_(self.stream[:self.nbytes])
        self.stream = st

    def stream_bytes(self) -> bytes:
        return self.stream[self.base:self.base + self.nbytes].cpu().numpy().tobytes()

    def decode(self, out: torch.Tensor | None = None, stream=None, shape: int = DECODE_AUTO) -> torch.Tensor:
        """Leco_int<T>::decodeArray8 over every block."""
        return self.decode_range(0, self.n, out=out, stream=stream, shape=shape)

    def decode_range(self, begin: int, end: int, out: torch.Tensor | None = None, stream=None,
                     shape: int = DECODE_AUTO) -> torch.Tensor:
        """Values [begin, end); shape picks the decode kernel (auto, staged or value-parallel)."""
        if out is None:
            out = torch.empty(end - begin, dtype=_STORAGE[self.width], device=self.stream.device)
        L = lib()
        fn = L.leco_fix_decode_range_ex_u32 if self.width == 32 else L.leco_fix_decode_range_ex_u64
        check(fn(self.stream.data_ptr(), self.offsets.data_ptr(), self.params.data_ptr(), self.n, self.block_size,
                 begin, end, shape, out.data_ptr(), _stream_handle(stream)), "decode_range")
        return out

    def lookup(self, idx: torch.Tensor, out: torch.Tensor | None = None, stream=None,
               grouped: bool = False) -> torch.Tensor:
        """Leco_int<T>::randomdecodeArray8 for every index in idx (int64 device tensor).

        grouped: group the queries by value range first (leco_fix_lookup_grouped_*,
        8 bytes of workspace per query), so that each block is read from HBM about
        once per batch: 2.4x less HBM traffic, but about 30% slower on MI355X (each
        scattered 8-byte result costs a 32-byte HBM write).  Off unless asked for."""
        idx = idx.contiguous()
        if idx.dtype not in (torch.int64, torch.uint64):
            idx = idx.to(torch.int64)
        if out is None:
            out = torch.empty(idx.numel(), dtype=_STORAGE[self.width], device=self.stream.device)
        L = lib()
        nq = idx.numel()
        if grouped:
            ws = int(L.leco_fix_lookup_grouped_workspace(nq))
            work = torch.empty(ws, dtype=torch.uint8, device=self.stream.device)
            if stream is not None:
                work.record_stream(stream)  # not reused before the lookups on `stream` are done
            fn = L.leco_fix_lookup_grouped_u32 if self.width == 32 else L.leco_fix_lookup_grouped_u64
            check(fn(self.stream.data_ptr(), self.params.data_ptr(), self.n, self.block_size, idx.data_ptr(),
                     nq, out.data_ptr(), work.data_ptr(), ws, _stream_handle(stream)), "lookup_grouped")
            return out
        fn = L.leco_fix_lookup_u32 if self.width == 32 else L.leco_fix_lookup_u64
        check(fn(self.stream.data_ptr(), self.params.data_ptr(), self.n, self.block_size, idx.data_ptr(),
                 nq, out.data_ptr(), _stream_handle(stream)), "lookup")
        return out

    def _filter(self, close: bool, filt, f1, f2, base: int, out, capacity, stream):
        L = lib()
        sfx = "u32" if self.width == 32 else "u64"
        dev = self.stream.device
        if out is None:
            cap = self.n if capacity is None else capacity
            out = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        else:
            cap = out.numel() if capacity is None else min(capacity, out.numel())
        result = torch.zeros(2, dtype=torch.int64, device=dev)
        wb = int(L.leco_fix_filter_workspace_bytes(self.n, self.block_size))
        work = torch.empty(wb, dtype=torch.uint8, device=dev)
        if close:
            st = _STORAGE[self.width]
            f1 = _as_storage(f1.to(device=dev).contiguous()) if f1.dtype != st else f1.contiguous()
            f2 = _as_storage(f2.to(device=dev).contiguous()) if f2.dtype != st else f2.contiguous()
            if f1.numel() != self.num_blocks or f2.numel() != self.num_blocks:
                raise ValueError("filter_range_close takes one starting filter pair per block")
            rc = getattr(L, f"leco_fix_filter_range_close_{sfx}")(
                self.stream.data_ptr(), self.offsets.data_ptr(), self.n, self.block_size, f1.data_ptr(),
                f2.data_ptr(), base, out.data_ptr(), cap, result.data_ptr(), work.data_ptr(), wb,
                _stream_handle(stream))
        else:
            rc = getattr(L, f"leco_fix_filter_range_{sfx}")(
                self.stream.data_ptr(), self.offsets.data_ptr(), self.n, self.block_size, int(filt), out.data_ptr(),
                cap, result.data_ptr(), work.data_ptr(), wb, _stream_handle(stream))
        check(rc, "filter_range_close" if close else "filter_range")
        return out, result

    def filter_range(self, filt: int, out: torch.Tensor | None = None, capacity: int | None = None,
                     stream=None) -> torch.Tensor:
        """Leco_int<T>::filter_range over every block: uint32 positions (as int32) of values above `filt`,
        from each block's model threshold on.  Synchronizes to size the result."""
        out, result = self._filter(False, filt, None, None, 0, out, capacity, stream)
        count = int(result[0].item())
        if count > out.numel():
            raise ValueError(f"filter_range: {count} matches, room for {out.numel()}")
        return out[:count]

    def filter_range_close(self, filter1: torch.Tensor, filter2: torch.Tensor, base: int,
                           out: torch.Tensor | None = None, capacity: int | None = None,
                           stream=None) -> torch.Tensor:
        """Leco_int<T>::filter_range_close over every block with per-block starting filters
        (see close_filters_from_zonemap); positions block by block, period by period."""
        from ._lib import LECO_FIX_ELIMIT, LecoError
        out, result = self._filter(True, 0, filter1, filter2, int(base), out, capacity, stream)
        count, flags = (int(v) for v in result.tolist())
        if flags & 1:
            raise LecoError(LECO_FIX_ELIMIT, "filter_range_close")
        if count > out.numel():
            out, result = self._filter(True, 0, filter1, filter2, int(base), None, count, stream)
        return out[:count]

    def rebuild_params(self, stream=None) -> torch.Tensor:
        """Param table recomputed from the block headers (leco_fix_build_params)."""
        prm = torch.empty_like(self.params)
        check(lib().leco_fix_build_params(self.stream.data_ptr(), self.offsets.data_ptr(), self.n,
                                          self.block_size, self.width, prm.data_ptr(), _stream_handle(stream)),
              "build_params")
        return prm


def close_filters_from_zonemap(block_min: torch.Tensor, filter1: int, filter2: int, base: int, width: int):
    """Per-block starting filters of the reference's close-filter driver
    (experiments/leco_int_fix_template.cpp:293-298): from each block's minimum
    m, k = ceil((double)(T)(m - f1) / base) periods on, one period earlier
    when m % base lies inside (f1, f2).  Float64 arithmetic and T wrap as the
    driver's C++ expressions; minimums, filters and base below 2^63, base > 0.
    Returns (f1[], f2[]) in the column's storage dtype."""
    if base <= 0:
        raise ValueError("base must be positive")
    mn = block_min.to(torch.int64)
    if width == 64 and bool((mn < 0).any()):
        raise ValueError("block minimums at or above 2^63 are not supported")
    wmask = (1 << width) - 1
    d = mn - filter1
    if width == 32:
        d = (d & wmask).to(torch.float64)
    else:
        d = d.to(torch.float64)
        d = torch.where(d < 0, d + 2.0 ** 64, d)               # (uint64)(m - f1)
    k = torch.ceil(d / float(base)) * float(base)
    t1 = _to_uint(k + float(filter1))
    t2 = _to_uint(k + float(filter2))
    r = mn % base
    inside = (r < filter2) & (r > filter1)
    t1 = torch.where(inside, t1 - base, t1)
    t2 = torch.where(inside, t2 - base, t2)
    if width == 32:
        return (t1 & wmask).to(torch.int64).to(torch.int32), (t2 & wmask).to(torch.int64).to(torch.int32)
    return t1, t2


def _to_uint(x: torch.Tensor) -> torch.Tensor:
    """double -> unsigned integer by truncation, as int64 bits (values in [2^63, 2^64) wrap like uint64)."""
    return torch.trunc(torch.where(x >= 2.0 ** 63, x - 2.0 ** 64, x)).to(torch.int64)
