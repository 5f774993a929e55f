"""piecewise_fix_op_max, the reference's uint32 IntegerCODEC
(headers/piecewise_fix_op_minimize_maxerror.h:18-212), on the gfx950 library.

OpMaxColumn: an HBM-resident column (stream, offsets, params and the int8
fix-up flags).  piecewise_fix_op_max: the reference's one-block host surface
(init / encodeArray8 / decodeArray8 / randomdecodeArray8 / get_block_nums)
with the IntegerCODEC lists mul_add_diff_set and mul_add_diff_set_minus,
filled by encodeArray8 exactly as the reference fills them.

The codec's decoder predicts with an additive chain that can truncate one
away from its encoder's prediction; those positions are the lists (flags +1 /
-1 here).  decode(fixed=True) applies them, giving the input back;
decode(fixed=False) is the reference decoder's raw output, and apply_lists()
is the reference harness's correction (experiments/fix_int.cpp:110-117)."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import BlockParam, LECO_FIX_TAIL_PAD, check, encoded_size, lib
from .column import _stream_handle

PARAM_BYTES = ctypes.sizeof(BlockParam)


def _u32(values: torch.Tensor) -> torch.Tensor:
    if values.dtype not in (torch.int32, torch.uint32):
        raise TypeError(f"piecewise_fix_op_max holds uint32 values (int32 storage), not {values.dtype}")
    return values.contiguous().view(torch.int32)


class OpMaxColumn:
    """An op_max column in HBM."""

    def __init__(self, stream, offsets, params, fix, n: int, block_size: int, nbytes: int):
        self.stream, self.offsets, self.params, self.fix = stream, offsets, params, fix
        self.n, self.block_size, self.nbytes = n, block_size, nbytes

    @property
    def num_blocks(self) -> int:
        return (self.n + self.block_size - 1) // self.block_size

    @property
    def compression_ratio(self) -> float:
        return self.nbytes / (4 * self.n)

    @classmethod
    def encode(cls, values: torch.Tensor, block_size: int, with_fix: bool = True, stream=None) -> "OpMaxColumn":
        values = _u32(values)
        n = values.numel()
        dev = values.device
        L = lib()
        bound = int(L.leco_fix_opmax_stream_bound(n, block_size))
        nb = (n + block_size - 1) // block_size
        st = torch.empty(bound, dtype=torch.uint8, device=dev)
        offs = torch.empty(nb + 1, dtype=torch.int64, device=dev)
        prm = torch.empty(nb * PARAM_BYTES, dtype=torch.uint8, device=dev)
        fix = torch.empty(n, dtype=torch.int8, device=dev) if with_fix else None
        check(L.leco_fix_opmax_encode_column(values.data_ptr(), n, block_size, st.data_ptr(), bound, offs.data_ptr(),
                                             prm.data_ptr(), fix.data_ptr() if fix is not None else None,
                                             _stream_handle(stream)), "opmax encode_column")
        return cls(st, offs, prm, fix, n, block_size, encoded_size(int(offs[-1].item()), "opmax_encode_column"))

    @classmethod
    def from_stream(cls, stream_bytes: bytes, offsets, n: int, block_size: int, fix=None,
                    device="cuda") -> "OpMaxColumn":
        """A column from stored block bytes (e.g. another encoder's), params rebuilt on the GPU."""
        raw = np.frombuffer(bytes(stream_bytes) + bytes(LECO_FIX_TAIL_PAD), np.uint8)
        st = torch.from_numpy(raw.copy()).to(device)
        offs = torch.as_tensor(np.asarray(offsets, np.int64)).to(device)
        nb = (n + block_size - 1) // block_size
        prm = torch.empty(nb * PARAM_BYTES, dtype=torch.uint8, device=device)
        check(lib().leco_fix_opmax_build_params(st.data_ptr(), offs.data_ptr(), n, block_size, prm.data_ptr(),
                                                _stream_handle(None)), "opmax build_params")
        fx = None if fix is None else torch.as_tensor(np.asarray(fix, np.int8)).to(device)
        return cls(st, offs, prm, fx, n, block_size, len(stream_bytes))

    def stream_bytes(self) -> bytes:
        return self.stream[:self.nbytes].cpu().numpy().tobytes()

    def decode(self, fixed: bool = True, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """fixed: the input values (fix-up flags applied); else the reference decoder's output."""
        if fixed and self.fix is None:
            raise ValueError("this column carries no fix-up flags")
        if out is None:
            out = torch.empty(self.n, dtype=torch.int32, device=self.stream.device)
        check(lib().leco_fix_opmax_decode_column(self.stream.data_ptr(), self.offsets.data_ptr(), self.n,
                                                 self.block_size, self.fix.data_ptr() if fixed else None,
                                                 out.data_ptr(), _stream_handle(stream)), "opmax decode_column")
        return out

    def decode_lists(self, plus: torch.Tensor, minus: torch.Tensor, out: torch.Tensor | None = None,
                     stream=None) -> torch.Tensor:
        """The input values from the decoder's output and the reference's sparse
        fix-up lists (fix_lists()), raw blocks left as stored."""
        plus = plus.contiguous().to(torch.int32)
        minus = minus.contiguous().to(torch.int32)
        if out is None:
            out = torch.empty(self.n, dtype=torch.int32, device=self.stream.device)
        check(lib().leco_fix_opmax_decode_column_lists(self.stream.data_ptr(), self.offsets.data_ptr(), self.n,
                                                       self.block_size, plus.data_ptr(), plus.numel(),
                                                       minus.data_ptr(), minus.numel(), out.data_ptr(),
                                                       _stream_handle(stream)), "opmax decode_column_lists")
        return out

    def lookup(self, idx: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        idx = idx.contiguous().to(torch.int64)
        if out is None:
            out = torch.empty(idx.numel(), dtype=torch.int32, device=self.stream.device)
        check(lib().leco_fix_opmax_lookup(self.stream.data_ptr(), self.params.data_ptr(), self.n, self.block_size,
                                          idx.data_ptr(), idx.numel(), out.data_ptr(), _stream_handle(stream)),
              "opmax lookup")
        return out

    def fix_lists(self, capacity: int | None = None, stream=None) -> tuple[torch.Tensor, torch.Tensor]:
        """(mul_add_diff_set, mul_add_diff_set_minus) as ascending int32 tensors."""
        if self.fix is None:
            raise ValueError("this column carries no fix-up flags")
        L = lib()
        dev = self.stream.device
        cap = self.n if capacity is None else capacity
        plus = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        minus = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        counts = torch.zeros(2, dtype=torch.int64, device=dev)
        wb = int(L.leco_fix_opmax_lists_workspace_bytes(self.n))
        work = torch.empty(max(wb, 1), dtype=torch.uint8, device=dev)
        check(L.leco_fix_opmax_fix_lists(self.fix.data_ptr(), self.n, plus.data_ptr(), minus.data_ptr(), cap,
                                         counts.data_ptr(), work.data_ptr(), wb, _stream_handle(stream)),
              "opmax fix_lists")
        np_, nm = (int(v) for v in counts.tolist())
        return plus[:min(np_, cap)], minus[:min(nm, cap)]

    def apply_lists(self, values: torch.Tensor, plus: torch.Tensor, minus: torch.Tensor, stream=None) -> torch.Tensor:
        """The reference harness's correction loops, in place on values."""
        plus = plus.contiguous().to(torch.int32)
        minus = minus.contiguous().to(torch.int32)
        check(lib().leco_fix_opmax_apply_lists(values.data_ptr(), values.numel(), plus.data_ptr(), plus.numel(),
                                               minus.data_ptr(), minus.numel(), _stream_handle(stream)),
              "opmax apply_lists")
        return values


class piecewise_fix_op_max:  # noqa: N801 -- the reference's class name
    """Codecset::piecewise_fix_op_max's IntegerCODEC surface, one block at a time."""

    def __init__(self):
        self.block_num = self.block_size = 0
        self.mul_add_diff_set: list[int] = []
        self.mul_add_diff_set_minus: list[int] = []

    def init(self, blocks: int, blocksize: int, extra: int = 0) -> None:
        self.block_num, self.block_size = blocks, blocksize

    def name(self) -> str:
        return "piecewise_fix_op_max"

    def get_block_nums(self) -> int:
        return 1

    def encodeArray8(self, data, length: int, res=None, nvalue: int = 0) -> bytes:
        """Block bytes; appends the block's fix-up positions (i + nvalue * block_size) to the lists."""
        y = np.ascontiguousarray(np.asarray(data, dtype=np.uint32)[:length])
        cap = int(lib().leco_fix_opmax_stream_bound(length, max(length, 1)))
        buf = np.empty(cap, np.uint8)
        fix = np.empty(max(length, 1), np.int8)
        written = ctypes.c_uint64()
        check(lib().leco_fix_opmax_encode_block(y.ctypes.data, length, buf.ctypes.data, cap, ctypes.byref(written),
                                                fix.ctypes.data), "encodeArray8")
        base = nvalue * self.block_size
        self.mul_add_diff_set += (np.flatnonzero(fix[:length] > 0) + base).tolist()
        self.mul_add_diff_set_minus += (np.flatnonzero(fix[:length] < 0) + base).tolist()
        out = buf[:written.value].tobytes()
        if res is not None:
            res[:len(out)] = out
        return out

    def decodeArray8(self, in_, length: int, out=None, nvalue: int = 0) -> np.ndarray:
        """The reference decoder's output (before the lists are applied)."""
        src = np.frombuffer(bytes(in_), np.uint8)
        dst = np.empty(max(length, 1), np.uint32)
        check(lib().leco_fix_opmax_decode_block(src.ctypes.data, length, dst.ctypes.data), "decodeArray8")
        if out is not None:
            out[:length] = dst[:length]
            return out
        return dst[:length]

    def randomdecodeArray8(self, in_, l: int, out=None, nvalue: int = 0) -> int:
        """Value l of the block (exact; reads the block only up to that value)."""
        src = np.frombuffer(bytes(in_), np.uint8)
        v = ctypes.c_uint32()
        check(lib().leco_fix_opmax_random_decode(src.ctypes.data, l + 1, l, ctypes.byref(v)), "randomdecodeArray8")
        return int(v.value)
