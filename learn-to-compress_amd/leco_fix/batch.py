"""Many columns, one launch: the batch C ABI (leco_fix_batch_*) and the
block-size autotuner.

A Batch holds a device block table over every block of its columns, one
arena with all the columns' streams, the arena offsets and the param table.
Encode is fit + one ticket-ordered workgroup per block + fix-up; decode is one
persistent launch.  Both can be replayed from a captured hipGraph.  Each
column's bytes are exactly its plain-path stream, so Batch.column(c) is a
Column usable with every column call.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import AutotuneChoice, ColumnDesc, check, encoded_size, lib
from .column import PARAM_BYTES, Column, _as_storage, _stream_handle, width_of

AUTOTUNE_CANDIDATES = (256, 512, 1024, 2048, 4096, 8192, 16384)


class _DeviceArray:
    """__cuda_array_interface__ over memory the library owns (no copy)."""

    def __init__(self, ptr: int, count: int, typestr: str):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": typestr, "data": (ptr, False),
                                         "version": 3, "strides": None}


def _view(ptr: int, count: int, typestr: str) -> torch.Tensor:
    return torch.as_tensor(_DeviceArray(ptr, count, typestr), device="cuda")


def _descs(values, outs, block_sizes, tiles=None):
    arr = (ColumnDesc * len(block_sizes))()
    for c, bs in enumerate(block_sizes):
        v = values[c] if values is not None else None
        o = outs[c] if outs is not None else None
        ref = v if v is not None else o
        arr[c] = ColumnDesc(v.data_ptr() if v is not None else None, o.data_ptr() if o is not None else None,
                            ref.numel(), int(bs), width_of(ref), int(tiles[c]) if tiles is not None else 0, 0)
    return arr


def autotune_block_sizes(values, candidates=AUTOTUNE_CANDIDATES, windows: int = 8, window: int | None = None,
                         tolerance: float = 0.01, stream=None):
    """Block size per column (search_bsize::search_block_size, headers/search_blocksize.h:14-114):
    encodes `windows` aligned sample windows at every candidate in one batched
    GPU encode and keeps the smallest encoding; within `tolerance` of it the
    largest block size wins.  Returns (block sizes, bytes[ncols, ncand])."""
    values = [_as_storage(v.contiguous()) for v in values]
    window = window or max(candidates)
    cand = (ctypes.c_uint32 * len(candidates))(*candidates)
    descs = _descs(values, None, [candidates[0]] * len(values))
    out = (ctypes.c_uint32 * len(values))()
    nbytes = np.zeros((len(values), len(candidates)), np.uint64)
    check(lib().leco_fix_autotune_block_sizes(descs, len(values), cand, len(candidates), windows, window,
                                              tolerance, out, nbytes.ctypes.data, _stream_handle(stream)),
          "autotune_block_sizes")
    return list(out), nbytes


def autotune_columns(values, candidates=AUTOTUNE_CANDIDATES, windows: int = 8, window: int | None = None,
                     tolerance: float = 0.02, stream=None):
    """Block size and decode tile shape per column (leco_fix_autotune_columns):
    candidate sizes and a refinement around the best one (after
    search_bsize::search_block_size, headers/search_blocksize.h:14-114) are
    encoded on sample windows; among the sizes within `tolerance` of the
    fewest bytes, the size and tile shape of least measured batched-decode
    time per value win.  Returns (block sizes, tile positions, choices) where
    choices holds the sample bytes and decode ns per value of each pick."""
    values = [_as_storage(v.contiguous()) for v in values]
    window = window or max(candidates)
    cand = (ctypes.c_uint32 * len(candidates))(*candidates)
    descs = _descs(values, None, [candidates[0]] * len(values))
    out = (AutotuneChoice * len(values))()
    check(lib().leco_fix_autotune_columns(descs, len(values), cand, len(candidates), windows, window, tolerance,
                                          out, _stream_handle(stream)), "autotune_columns")
    return [c.block_size for c in out], [c.tile_positions for c in out], \
        [(c.bytes_per_value, c.decode_ns_per_value) for c in out]


class Batch:
    def __init__(self, values, block_sizes, outs=None, tile_positions=None):
        """values: device tensors (int32/int64 storage); outs: decode targets (default: new tensors);
        tile_positions: per column, the positions one decode workgroup takes for blocks below 2048
        values (0 or None: 8192)."""
        self.values = [_as_storage(v.contiguous()) for v in values]
        self.block_sizes = [int(b) for b in block_sizes]
        self.tile_positions = [int(t) for t in tile_positions] if tile_positions is not None else None
        self.outs = outs if outs is not None else [torch.empty_like(v) for v in self.values]
        descs = _descs(self.values, self.outs, self.block_sizes, self.tile_positions)
        h = ctypes.c_void_p()
        check(lib().leco_fix_batch_create(descs, len(self.values), ctypes.byref(h)), "batch_create")
        self._h = h
        total, cap = ctypes.c_uint64(), ctypes.c_uint64()
        arena, offs, prm = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().leco_fix_batch_info(h, ctypes.byref(total), ctypes.byref(cap), ctypes.byref(arena),
                                        ctypes.byref(offs), ctypes.byref(prm)), "batch_info")
        self.total_blocks, self.arena_capacity = total.value, cap.value
        self.arena = _view(arena.value, cap.value, "|u1")
        self.offsets = _view(offs.value, total.value + 1, "<i8")
        self.params = _view(prm.value, total.value * PARAM_BYTES, "|u1")
        self.first = []
        for c in range(len(self.values)):
            f, nb = ctypes.c_uint64(), ctypes.c_uint64()
            check(lib().leco_fix_batch_column(h, c, ctypes.byref(f), ctypes.byref(nb)), "batch_column")
            self.first.append((f.value, nb.value))

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            torch.cuda.synchronize()
            lib().leco_fix_batch_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def raw_bytes(self) -> int:
        return sum(v.numel() * v.element_size() for v in self.values)

    def encode(self, stream=None, graph: bool = False) -> None:
        s = _stream_handle(stream)
        rc = lib().leco_fix_batch_graph_launch(self._h, 0, s) if graph else lib().leco_fix_batch_encode(self._h, s)
        check(rc, "batch_encode")

    def decode(self, stream=None, graph: bool = False) -> list:
        s = _stream_handle(stream)
        rc = lib().leco_fix_batch_graph_launch(self._h, 1, s) if graph else lib().leco_fix_batch_decode(self._h, s)
        check(rc, "batch_decode")
        return self.outs

    def decode_per_column(self, stream=None) -> list:
        """The plain path: one decode launch per column, from a host loop."""
        check(lib().leco_fix_batch_decode_per_column(self._h, _stream_handle(stream)), "batch_decode_per_column")
        return self.outs

    def encode_per_column(self, stream=None) -> None:
        """The plain path: one column encode per column into a scratch stream."""
        check(lib().leco_fix_batch_encode_per_column(self._h, _stream_handle(stream)), "batch_encode_per_column")

    def set_decode_grid(self, workgroups: int) -> None:
        check(lib().leco_fix_batch_set_decode_grid(self._h, workgroups), "batch_set_decode_grid")

    def compressed_bytes(self) -> int:
        return encoded_size(int(self.offsets[self.total_blocks].item()), "batch_encode")

    def column(self, c: int) -> Column:
        """Column c as a Column (arena, offsets + first, params + first)."""
        f, nb = self.first[c]
        v = self.values[c]
        offs = self.offsets[f:f + nb + 1]
        lo, hi = int(offs[0].item()), int(offs[nb].item())
        col = Column(self.arena, offs, self.params[f * PARAM_BYTES:(f + nb) * PARAM_BYTES], v.numel(),
                     self.block_sizes[c], width_of(v), hi - lo)
        col.base = lo
        return col

    def column_bytes(self, c: int) -> bytes:
        col = self.column(c)
        return self.arena[col.base:col.base + col.nbytes].cpu().numpy().tobytes()
