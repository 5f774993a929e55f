"""Loads libleco_fix.so (the gfx950 HIP library) and declares its C ABI.

There is no CPU fallback: if the library is not built, loading fails with an
error that says how to build it.
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libleco_fix.so")

c_u8p = ctypes.c_void_p
c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_ptr = ctypes.c_void_p
c_int = ctypes.c_int

LECO_FIX_OK = 0
LECO_FIX_EINVAL = 1
LECO_FIX_ELIMIT = 4
LECO_FIX_ESCRATCH = 5
LECO_FIX_EDEVICE = 6
LECO_FIX_OFFSETS_POISON = (1 << 64) - 1
LECO_FIX_NOALLOC_SCRATCH_BYTES = 64 << 20
LECO_FIX_TAIL_PAD = 64


class LecoError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {strerror(code)} (code {code})")
        self.code = code


class BlockParam(ctypes.Structure):
    """leco_fix_block_param (include/leco_fix.h)."""
    _fields_ = [("theta0", ctypes.c_double), ("theta1", ctypes.c_double),
                ("payload_offset", ctypes.c_uint64), ("bits", ctypes.c_uint32), ("length", ctypes.c_uint32)]


# name -> (restype, argtypes); every function declared in include/leco_fix.h
SIGNATURES = {
    "leco_fix_strerror": (ctypes.c_char_p, [c_int]),
    "leco_fix_block_count": (c_u64, [c_u64, c_u32]),
    "leco_fix_stream_bound": (c_u64, [c_u64, c_u32, c_u32]),
    "leco_fix_block_bytes": (c_u64, [c_u32, c_u32, c_u32]),
    "leco_fix_build_params": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u32, c_ptr, c_ptr]),
}
for _w in ("u32", "u64"):
    SIGNATURES.update({
        f"leco_fix_encode_column_{_w}": (c_int, [c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr, c_ptr]),
        f"leco_fix_decode_column_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_ptr]),
        f"leco_fix_decode_range_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u64, c_u64, c_ptr, c_ptr]),
        f"leco_fix_decode_range_ex_{_w}": (c_int, [c_ptr, c_ptr, c_ptr, c_u64, c_u32, c_u64, c_u64, c_int, c_ptr,
                                                   c_ptr]),
        f"leco_fix_lookup_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr]),
        f"leco_fix_lookup_grouped_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr, c_u64,
                                                  c_ptr]),
        f"leco_fix_encode_block_{_w}": (c_int, [c_ptr, c_u64, c_ptr, c_u64, ctypes.POINTER(c_u64)]),
        f"leco_fix_decode_block_{_w}": (c_int, [c_ptr, c_u64, c_ptr]),
        f"leco_fix_random_decode_{_w}": (c_int, [c_ptr, c_u32, c_ptr]),
        f"leco_fix_filter_range_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u32 if _w == "u32" else c_u64, c_ptr,
                                                c_u64, c_ptr, c_ptr, c_u64, c_ptr]),
        f"leco_fix_filter_range_close_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_ptr, ctypes.c_int32,
                                                      c_ptr, c_u64, c_ptr, c_ptr, c_u64, c_ptr]),
        f"leco_fix_filter_block_{_w}": (c_int, [c_ptr, c_u64, c_u32 if _w == "u32" else c_u64, c_ptr, c_u64,
                                                ctypes.c_int32, c_u32, ctypes.POINTER(c_u64)]),
        f"leco_fix_filter_close_block_{_w}": (c_int, [c_ptr, c_u64, c_ptr, c_u64, ctypes.c_int32, c_u32,
                                                      c_u32 if _w == "u32" else c_u64,
                                                      c_u32 if _w == "u32" else c_u64, ctypes.c_int32,
                                                      ctypes.POINTER(c_u64)]),
    })
SIGNATURES["leco_fix_filter_workspace_bytes"] = (c_u64, [c_u64, c_u32])
SIGNATURES["leco_fix_lookup_grouped_workspace"] = (c_u64, [c_u64])
SIGNATURES["leco_fix_encode_result"] = (c_int, [c_ptr, c_u64, ctypes.POINTER(c_u64), c_ptr])
SIGNATURES["leco_fix_test_lookback_hook"] = (c_int, [c_u32, c_u64])
SIGNATURES["leco_fix_filter_range_noalloc"] = (c_int, [c_u32, c_ptr, c_ptr, c_u64, c_u32, c_u64, c_ptr, c_u64, c_ptr,
                                                     c_ptr])


LECO_FIX_CODEC_FOR, LECO_FIX_CODEC_DELTA = 1, 2
SIGNATURES["leco_fix_base_stream_bound"] = (c_u64, [c_int, c_u64, c_u32, c_u32])
for _w in ("u32", "u64"):
    _T = c_u32 if _w == "u32" else c_u64
    SIGNATURES.update({
        f"leco_fix_base_encode_column_{_w}": (c_int, [c_int, c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr]),
        f"leco_fix_base_decode_column_{_w}": (c_int, [c_int, c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_ptr]),
        f"leco_fix_base_lookup_{_w}": (c_int, [c_int, c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr]),
        f"leco_fix_for_filter_range_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, _T, c_ptr, c_u64, c_ptr, c_ptr, c_u64,
                                                    c_ptr]),
        f"leco_fix_for_filter_range_close_{_w}": (c_int, [c_ptr, c_ptr, c_u64, c_u32, _T, _T, ctypes.c_int32, c_ptr,
                                                          c_u64, c_ptr, c_ptr, c_u64, c_ptr]),
        f"leco_fix_base_encode_block_{_w}": (c_int, [c_int, c_ptr, c_u64, c_ptr, c_u64, ctypes.POINTER(c_u64)]),
        f"leco_fix_base_decode_block_{_w}": (c_int, [c_int, c_ptr, c_u64, c_ptr]),
        f"leco_fix_base_random_decode_{_w}": (c_int, [c_int, c_ptr, c_u64, c_u32, c_ptr]),
        f"leco_fix_for_filter_block_{_w}": (c_int, [c_ptr, c_u64, _T, c_ptr, c_u64, ctypes.c_int32, c_u32,
                                                    ctypes.POINTER(c_u64)]),
        f"leco_fix_for_filter_close_block_{_w}": (c_int, [c_ptr, c_u64, c_ptr, c_u64, ctypes.c_int32, c_u32, _T, _T,
                                                          ctypes.c_int32, ctypes.POINTER(c_u64)]),
    })


class ColumnFileInfo(ctypes.Structure):
    """leco_fix_file_info (include/leco_fix.h)."""
    _fields_ = [("n", ctypes.c_uint64), ("num_blocks", ctypes.c_uint64), ("stream_bytes", ctypes.c_uint64),
                ("block_size", ctypes.c_uint32), ("width_bits", ctypes.c_uint32)]


SIGNATURES.update({
    "leco_fix_sosd_info": (c_int, [ctypes.c_char_p, c_u32, ctypes.POINTER(c_u64)]),
    "leco_fix_sosd_load": (c_int, [ctypes.c_char_p, c_u32, c_ptr, c_u64, ctypes.POINTER(c_u64), c_ptr]),
    "leco_fix_sosd_save": (c_int, [ctypes.c_char_p, c_u32, c_ptr, c_u64, c_ptr]),
    "leco_fix_text_info": (c_int, [ctypes.c_char_p, c_u32, ctypes.POINTER(c_u64)]),
    "leco_fix_text_load": (c_int, [ctypes.c_char_p, c_u32, c_ptr, c_u64, ctypes.POINTER(c_u64), c_ptr]),
    "leco_fix_column_file_info": (c_int, [ctypes.c_char_p, ctypes.POINTER(ColumnFileInfo)]),
    "leco_fix_column_save": (c_int, [ctypes.c_char_p, c_ptr, c_ptr, c_u64, c_u32, c_u32, c_ptr]),
    "leco_fix_column_load": (c_int, [ctypes.c_char_p, c_ptr, c_u64, c_ptr, c_ptr, c_ptr]),
})



class AutotuneChoice(ctypes.Structure):
    """leco_fix_autotune_choice (include/leco_fix.h)."""
    _fields_ = [("block_size", ctypes.c_uint32), ("tile_positions", ctypes.c_uint32),
                ("bytes_per_value", ctypes.c_double), ("decode_ns_per_value", ctypes.c_double)]


class ColumnDesc(ctypes.Structure):
    """leco_fix_column_desc (include/leco_fix.h)."""
    _fields_ = [("d_in", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("n", ctypes.c_uint64),
                ("block_size", ctypes.c_uint32), ("width_bits", ctypes.c_uint32),
                ("tile_positions", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


_pp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES.update({
    "leco_fix_batch_create": (c_int, [c_ptr, c_u32, _pp]),
    "leco_fix_batch_destroy": (c_int, [c_ptr]),
    "leco_fix_batch_encode": (c_int, [c_ptr, c_ptr]),
    "leco_fix_batch_decode": (c_int, [c_ptr, c_ptr]),
    "leco_fix_batch_graph_launch": (c_int, [c_ptr, c_int, c_ptr]),
    "leco_fix_batch_set_output": (c_int, [c_ptr, c_u32, c_ptr]),
    "leco_fix_batch_decode_per_column": (c_int, [c_ptr, c_ptr]),
    "leco_fix_batch_encode_per_column": (c_int, [c_ptr, c_ptr]),
    "leco_fix_batch_set_decode_grid": (c_int, [c_ptr, c_u32]),
    "leco_fix_batch_info": (c_int, [c_ptr, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64), _pp, _pp, _pp]),
    "leco_fix_batch_column": (c_int, [c_ptr, c_u32, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)]),
    "leco_fix_autotune_block_sizes": (c_int, [c_ptr, c_u32, c_ptr, c_u32, c_u32, c_u32, ctypes.c_double, c_ptr,
                                              c_ptr, c_ptr]),
})

SIGNATURES.update({
    "leco_fix_server_create": (c_int, [c_ptr, _pp]),
    "leco_fix_server_destroy": (c_int, [c_ptr]),
    "leco_fix_server_attach_column": (c_int, [c_ptr, c_u32, c_ptr, c_ptr, c_ptr, c_u64, c_u32, c_u32]),
    "leco_fix_server_run": (c_int, [c_ptr, c_ptr, c_u64, c_int, c_ptr, c_ptr]),
    "leco_fix_server_column": (c_int, [c_ptr, c_u32, _pp, _pp, _pp, ctypes.POINTER(c_u64), ctypes.POINTER(c_u32),
                                       ctypes.POINTER(c_u32), ctypes.POINTER(c_u64)]),
})

SIGNATURES["leco_fix_autotune_columns"] = (c_int, [c_ptr, c_u32, c_ptr, c_u32, c_u32, c_u32, ctypes.c_double, c_ptr,
                                                 c_ptr])
SIGNATURES["leco_fix_server_slowest_iterations"] = (c_int, [c_ptr, c_ptr, c_u32, ctypes.POINTER(c_u32)])

SIGNATURES.update({
    "leco_fix_opmax_stream_bound": (c_u64, [c_u64, c_u32]),
    "leco_fix_opmax_encode_column": (c_int, [c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr, c_ptr, c_ptr]),
    "leco_fix_opmax_decode_column": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_ptr, c_ptr]),
    "leco_fix_opmax_decode_column_lists": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_u64, c_ptr,
                                                   c_ptr]),
    "leco_fix_opmax_lookup": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_u64, c_ptr, c_ptr]),
    "leco_fix_opmax_build_params": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_ptr, c_ptr]),
    "leco_fix_opmax_lists_workspace_bytes": (c_u64, [c_u64]),
    "leco_fix_opmax_fix_lists": (c_int, [c_ptr, c_u64, c_ptr, c_ptr, c_u64, c_ptr, c_ptr, c_u64, c_ptr]),
    "leco_fix_opmax_apply_lists": (c_int, [c_ptr, c_u64, c_ptr, c_u64, c_ptr, c_u64, c_ptr]),
    "leco_fix_opmax_encode_block": (c_int, [c_ptr, c_u64, c_ptr, c_u64, ctypes.POINTER(c_u64), c_ptr]),
    "leco_fix_opmax_decode_block": (c_int, [c_ptr, c_u64, c_ptr]),
    "leco_fix_opmax_random_decode": (c_int, [c_ptr, c_u64, c_u32, c_ptr]),
})

_lib = None


def _load_torch_runtime_first() -> None:
    """torch ships its own HIP runtime (torch/lib/libamdhip64.so, soname
    libamdhip64.so.7).  The library links libamdhip64.so.7 by soname, so when
    torch is imported first the library binds to torch's runtime and the
    process has one.  Loaded the other way round, /opt/rocm's runtime comes in
    first and torch then loads its own copy as well: two HIP runtimes in one
    process, where device pointers and kernels of one are invalid in the
    other (torch reports no GPU, or the library's launches fail).  So torch,
    when it is installed, is imported before the library is loaded."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python learn-to-compress_amd/build.py` "
                              "(or __graft_entry__.build()) first")
        _load_torch_runtime_first()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def strerror(code: int) -> str:
    try:
        return lib().leco_fix_strerror(code).decode()
    except ImportError:
        return "library not loaded"


def check(code: int, what: str) -> None:
    if code != LECO_FIX_OK:
        raise LecoError(code, what)


def encoded_size(end: int, what: str) -> int:
    """The encoded size d_offsets[nb] read back as an int64; raises
    LECO_FIX_EDEVICE if the encoder poisoned it (LECO_FIX_OFFSETS_POISON)."""
    if end & LECO_FIX_OFFSETS_POISON == LECO_FIX_OFFSETS_POISON:
        raise LecoError(LECO_FIX_EDEVICE, what)
    return end
