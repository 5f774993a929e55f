"""Files: the reference's SOSD binary input and persisted column files
(C ABI leco_fix_sosd_* / leco_fix_column_* in include/leco_fix.h)."""
from __future__ import annotations

import ctypes
import os

import torch

from ._lib import ColumnFileInfo, check, lib
from .column import PARAM_BYTES, Column, _STORAGE, _as_storage, _stream_handle, width_of
from ._lib import LECO_FIX_TAIL_PAD


def _path(p) -> bytes:
    return os.fsencode(p)


def sosd_info(path, width: int) -> int:
    """Value count of an SOSD binary file (u64 count + raw values)."""
    c = ctypes.c_uint64()
    check(lib().leco_fix_sosd_info(_path(path), width, ctypes.byref(c)), "sosd_info")
    return c.value


def load_sosd(path, width: int, device="cuda", stream=None) -> torch.Tensor:
    """experiments/leco_int_fix_template.cpp:21-40 (load_data_binary), straight into HBM."""
    n = sosd_info(path, width)
    out = torch.empty(max(n, 1), dtype=_STORAGE[width], device=device)
    c = ctypes.c_uint64()
    check(lib().leco_fix_sosd_load(_path(path), width, out.data_ptr(), n, ctypes.byref(c), _stream_handle(stream)),
          "sosd_load")
    return out[:n]


def save_sosd(path, values: torch.Tensor, stream=None) -> None:
    values = _as_storage(values.contiguous())
    check(lib().leco_fix_sosd_save(_path(path), width_of(values), values.data_ptr(), values.numel(),
                                   _stream_handle(stream)), "sosd_save")


def text_info(path, width: int) -> int:
    """Value count of a text input file, parsed as the reference's drivers
    read it (leco_fix_text_info; host only)."""
    c = ctypes.c_uint64()
    check(lib().leco_fix_text_info(_path(path), width, ctypes.byref(c)), "text_info")
    return c.value


def load_text(path, width: int, device="cuda", stream=None) -> torch.Tensor:
    """experiments/leco_int_fix_template.cpp:43-58 (load_data) and
    experiments/fix_int.cpp:30-44: whitespace-separated decimal values, into HBM."""
    n = text_info(path, width)
    out = torch.empty(max(n, 1), dtype=_STORAGE[width], device=device)
    c = ctypes.c_uint64()
    check(lib().leco_fix_text_load(_path(path), width, out.data_ptr(), n, ctypes.byref(c), _stream_handle(stream)),
          "text_load")
    return out[:c.value]


def column_file_info(path) -> dict:
    info = ColumnFileInfo()
    check(lib().leco_fix_column_file_info(_path(path), ctypes.byref(info)), "column_file_info")
    return {"n": info.n, "num_blocks": info.num_blocks, "stream_bytes": info.stream_bytes,
            "block_size": info.block_size, "width": info.width_bits}


def save_column(path, col: Column, stream=None) -> None:
    check(lib().leco_fix_column_save(_path(path), col.stream.data_ptr(), col.offsets.data_ptr(), col.n,
                                     col.block_size, col.width, _stream_handle(stream)), "column_save")


def load_column(path, device="cuda", stream=None) -> Column:
    """A column file back into HBM: stream, offsets and the rebuilt param table."""
    info = column_file_info(path)
    st = torch.empty(info["stream_bytes"] + LECO_FIX_TAIL_PAD, dtype=torch.uint8, device=device)
    offs = torch.empty(info["num_blocks"] + 1, dtype=torch.int64, device=device)
    prm = torch.empty(info["num_blocks"] * PARAM_BYTES, dtype=torch.uint8, device=device)
    check(lib().leco_fix_column_load(_path(path), st.data_ptr(), st.numel(), offs.data_ptr(), prm.data_ptr(),
                                     _stream_handle(stream)), "column_load")
    return Column(st, offs, prm, info["n"], info["block_size"], info["width"], info["stream_bytes"])
