"""ctypes wrapper of the CPU oracle (oracle/leco_fix_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker of the GPU library; nothing in learn-to-compress_amd/ imports it.
Each function mirrors the reference call it restates (see the C file).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_leco.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle_leco.so"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.leco_oracle_cvtt64.argtypes = [ctypes.c_double]
        L.leco_oracle_cvtt64.restype = ctypes.c_int64
        L.leco_oracle_fit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.leco_oracle_fit.restype = None
        L.leco_oracle_payload_bytes.argtypes = [ctypes.c_int, ctypes.c_uint32]
        L.leco_oracle_payload_bytes.restype = ctypes.c_uint64
        L.leco_oracle_block_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
        L.leco_oracle_block_bytes.restype = ctypes.c_uint64
        L.leco_oracle_stream_bound.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int]
        L.leco_oracle_stream_bound.restype = ctypes.c_uint64
        L.leco_oracle_encode_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]
        L.leco_oracle_encode_block.restype = ctypes.c_uint64
        L.leco_oracle_decode_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                               ctypes.c_void_p, ctypes.c_int]
        L.leco_oracle_decode_block.restype = None
        L.leco_oracle_lookup_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]
        L.leco_oracle_lookup_block.restype = ctypes.c_uint64
        L.leco_oracle_encode_column.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.leco_oracle_encode_column.restype = ctypes.c_uint64
        L.leco_oracle_decode_column.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                                ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
        L.leco_oracle_decode_column.restype = None
        L.leco_oracle_lookup_column.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                                ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                                                ctypes.c_void_p, ctypes.c_int]
        L.leco_oracle_lookup_column.restype = None
        L.leco_oracle_filter_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                               ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
        L.leco_oracle_filter_block.restype = ctypes.c_uint64
        L.leco_oracle_filter_close_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                                     ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32,
                                                     ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                                                     ctypes.POINTER(ctypes.c_uint64)]
        L.leco_oracle_filter_close_block.restype = ctypes.c_int
        L.leco_oracle_filter_column.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                                ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.leco_oracle_filter_column.restype = None
        L.leco_oracle_filter_close_column.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                      ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                                      ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.leco_oracle_filter_close_column.restype = ctypes.c_int
        for name, res, args in (
                ("leco_oracle_for_block_bytes", ctypes.c_uint64, [ctypes.c_int, ctypes.c_int, ctypes.c_uint32]),
                ("leco_oracle_delta_block_bytes", ctypes.c_uint64, [ctypes.c_int, ctypes.c_int, ctypes.c_uint32]),
                ("leco_oracle_for_encode_block", ctypes.c_uint64,
                 [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]),
                ("leco_oracle_delta_encode_block", ctypes.c_uint64,
                 [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]),
                ("leco_oracle_for_decode_block", None,
                 [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]),
                ("leco_oracle_delta_decode_block", None,
                 [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]),
                ("leco_oracle_for_lookup_block", ctypes.c_uint64, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]),
                ("leco_oracle_delta_lookup_block", ctypes.c_uint64, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]),
                ("leco_oracle_base_encode_column", ctypes.c_uint64,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p]),
                ("leco_oracle_base_decode_column", None,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                  ctypes.c_void_p]),
                ("leco_oracle_base_lookup_column", None,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
                ("leco_oracle_for_filter_column", ctypes.c_uint64,
                 [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                  ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint64]),
                ("leco_oracle_opmax_block_bytes", ctypes.c_uint64, [ctypes.c_int, ctypes.c_uint32]),
                ("leco_oracle_opmax_encode_block", ctypes.c_uint64,
                 [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
                ("leco_oracle_opmax_decode_block", None, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
                ("leco_oracle_opmax_lookup_block", ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_uint32]),
                ("leco_oracle_opmax_encode_column", ctypes.c_uint64,
                 [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]),
                ("leco_oracle_opmax_decode_column", None,
                 [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p]),
                ("leco_oracle_decode_block_refill", None,
                 [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]),
                ("leco_oracle_decode_column_refill", None,
                 [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_int])):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _width(a: np.ndarray) -> int:
    if a.dtype == np.uint32:
        return 32
    if a.dtype == np.uint64:
        return 64
    raise TypeError(f"oracle takes uint32 or uint64 arrays, not {a.dtype}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def cvtt64(x: float) -> int:
    return lib().leco_oracle_cvtt64(x)


def fit(y: np.ndarray) -> tuple[float, float]:
    """lr_int_T<T>::caltheta (lr.h:66-85) -> (theta0, theta1)."""
    y = np.ascontiguousarray(y)
    t0, t1 = ctypes.c_double(), ctypes.c_double()
    lib().leco_oracle_fit(_ptr(y), _width(y), len(y), ctypes.byref(t0), ctypes.byref(t1))
    return t0.value, t1.value


def block_bytes(l: int, width: int, n: int) -> int:
    return lib().leco_oracle_block_bytes(l, width, n)


def stream_bound(n: int, bs: int, width: int) -> int:
    return lib().leco_oracle_stream_bound(n, bs, width)


def encode_block(y: np.ndarray) -> bytes:
    """Leco_int<T>::encodeArray8_int for one block."""
    y = np.ascontiguousarray(y)
    w = _width(y)
    out = np.zeros(max(block_bytes(w, w, len(y)), block_bytes(w - 1, w, len(y))) + 64, np.uint8)
    nb = lib().leco_oracle_encode_block(_ptr(y), w, len(y), _ptr(out))
    return out[:nb].tobytes()


def decode_block(blk: bytes, n: int, dtype, ref_compat: bool = False) -> np.ndarray:
    """Leco_int<T>::decodeArray8 for one block of n values."""
    dtype = np.dtype(dtype)
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    out = np.zeros(max(n, 2), dtype)
    lib().leco_oracle_decode_block(_ptr(src), 8 * dtype.itemsize, n, _ptr(out), int(ref_compat))
    return out[:n]


def lookup_block(blk: bytes, j: int, dtype) -> int:
    """Leco_int<T>::randomdecodeArray8(blk, j)."""
    dtype = np.dtype(dtype)
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    return int(lib().leco_oracle_lookup_block(_ptr(src), 8 * dtype.itemsize, j))


def encode_column(y: np.ndarray, bs: int, threads: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Column of blocks -> (stream bytes as uint8 array, offsets[nb+1])."""
    y = np.ascontiguousarray(y)
    w = _width(y)
    n = len(y)
    nb = (n + bs - 1) // bs
    stream = np.zeros(stream_bound(n, bs, w) + 64, np.uint8)
    offs = np.zeros(nb + 1, np.uint64)
    total = lib().leco_oracle_encode_column(_ptr(y), w, n, bs, _ptr(stream), _ptr(offs), threads)
    return stream[:total].copy(), offs


def decode_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int, dtype, threads: int = 1,
                  out: np.ndarray | None = None, refill: bool = False) -> np.ndarray:
    """Every block decoded; refill=True uses the loop shape of the reference's
    read_all_bit_fix (64-bit words into a 128-bit register) instead of the
    byte-window field reads of the checker."""
    dtype = np.dtype(dtype)
    padded = stream if len(stream) >= int(offs[-1]) + 64 else np.concatenate([stream, np.zeros(64, np.uint8)])
    if out is None:
        out = np.empty(n, dtype)
    fn = lib().leco_oracle_decode_column_refill if refill else lib().leco_oracle_decode_column
    fn(_ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs, _ptr(out), threads)
    return out


def lookup_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int, idx: np.ndarray, dtype,
                  threads: int = 1) -> np.ndarray:
    dtype = np.dtype(dtype)
    idx = np.ascontiguousarray(idx, np.uint64)
    padded = np.concatenate([stream, np.zeros(64, np.uint8)])
    out = np.empty(len(idx), dtype)
    lib().leco_oracle_lookup_column(_ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs, _ptr(idx), len(idx),
                                    _ptr(out), threads)
    return out


def _padded(stream: np.ndarray, offs: np.ndarray) -> np.ndarray:
    stream = np.ascontiguousarray(stream, np.uint8)
    if len(stream) >= int(offs[-1]) + 64:
        return stream
    return np.concatenate([stream, np.zeros(64, np.uint8)])


def filter_block(blk: bytes, n: int, dtype, filt: int, block_start: int = 0) -> np.ndarray:
    """Leco_int<T>::filter_range(blk, n, filt, out, block_id) positions (block_start = block_id * bs)."""
    dtype = np.dtype(dtype)
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    out = np.zeros(max(n, 1), np.uint32)
    c = lib().leco_oracle_filter_block(_ptr(src), 8 * dtype.itemsize, n, filt, block_start, _ptr(out), len(out))
    return out[:c].copy()


def filter_close_block(blk: bytes, n: int, dtype, f1: int, f2: int, base: int, block_start: int = 0) -> np.ndarray:
    """Leco_int<T>::filter_range_close(blk, n, out, block_id, f1, f2, base) positions, every period's in order."""
    dtype = np.dtype(dtype)
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    cap = 4 * n + 16
    while True:
        out = np.zeros(cap, np.uint32)
        cnt = ctypes.c_uint64()
        st = lib().leco_oracle_filter_close_block(_ptr(src), 8 * dtype.itemsize, n, f1, f2, base, block_start,
                                                  _ptr(out), cap, ctypes.byref(cnt))
        if st:
            raise RuntimeError("filter_close_block: too many periods in one block")
        if cnt.value <= cap:
            return out[:cnt.value].copy()
        cap = cnt.value


def filter_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int, dtype, filt: int) -> np.ndarray:
    """filter_range over every block of a column (positions in block order)."""
    dtype = np.dtype(dtype)
    padded = _padded(stream, offs)
    cnt = ctypes.c_uint64()
    out = np.zeros(n, np.uint32)
    lib().leco_oracle_filter_column(_ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs, filt, _ptr(out), n,
                                    ctypes.byref(cnt))
    return out[:cnt.value].copy()


def filter_close_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int, dtype, f1: np.ndarray,
                        f2: np.ndarray, base: int) -> np.ndarray:
    """filter_range_close over every block with per-block starting filters f1[b], f2[b]."""
    dtype = np.dtype(dtype)
    padded = _padded(stream, offs)
    f1 = np.ascontiguousarray(f1, np.uint64)
    f2 = np.ascontiguousarray(f2, np.uint64)
    cap = n + len(offs)
    while True:
        out = np.zeros(cap, np.uint32)
        cnt = ctypes.c_uint64()
        st = lib().leco_oracle_filter_close_column(_ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs, _ptr(f1),
                                                   _ptr(f2), base, _ptr(out), cap, ctypes.byref(cnt))
        if st:
            raise RuntimeError("filter_close_column: too many periods in one block")
        if cnt.value <= cap:
            return out[:cnt.value].copy()
        cap = cnt.value


# ---- FOR_int / Delta_int baselines (codec "for" or "delta") ----------------------------

_CODEC = {"for": 1, "delta": 2}


def base_block_bytes(codec: str, l: int, width: int, n: int) -> int:
    fn = lib().leco_oracle_for_block_bytes if codec == "for" else lib().leco_oracle_delta_block_bytes
    return fn(l, width, n)


def base_encode_block(codec: str, y: np.ndarray) -> bytes:
    y = np.ascontiguousarray(y)
    w = _width(y)
    out = np.zeros(len(y) * 8 + 64, np.uint8)
    fn = lib().leco_oracle_for_encode_block if codec == "for" else lib().leco_oracle_delta_encode_block
    nb = fn(_ptr(y), w, len(y), _ptr(out))
    return out[:nb].tobytes()


def base_decode_block(codec: str, blk: bytes, n: int, dtype, ref_compat: bool = False) -> np.ndarray:
    dtype = np.dtype(dtype)
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    out = np.zeros(max(n, 2), dtype)
    if codec == "for":
        lib().leco_oracle_for_decode_block(_ptr(src), 8 * dtype.itemsize, n, _ptr(out), int(ref_compat))
    else:
        lib().leco_oracle_delta_decode_block(_ptr(src), 8 * dtype.itemsize, n, _ptr(out))
    return out[:n]


def base_lookup_block(codec: str, blk: bytes, j: int, dtype) -> int:
    dtype = np.dtype(dtype)
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    fn = lib().leco_oracle_for_lookup_block if codec == "for" else lib().leco_oracle_delta_lookup_block
    return int(fn(_ptr(src), 8 * dtype.itemsize, j))


def base_encode_column(codec: str, y: np.ndarray, bs: int) -> tuple[np.ndarray, np.ndarray]:
    y = np.ascontiguousarray(y)
    w = _width(y)
    n = len(y)
    nb = (n + bs - 1) // bs
    stream = np.zeros(n * (w // 8) + nb * 48 + 128, np.uint8)
    offs = np.zeros(nb + 1, np.uint64)
    total = lib().leco_oracle_base_encode_column(_CODEC[codec], _ptr(y), w, n, bs, _ptr(stream), _ptr(offs))
    return stream[:total].copy(), offs


def base_decode_column(codec: str, stream: np.ndarray, offs: np.ndarray, n: int, bs: int, dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    out = np.empty(n, dtype)
    padded = _padded(stream, offs)
    lib().leco_oracle_base_decode_column(_CODEC[codec], _ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs,
                                         _ptr(out))
    return out


def base_lookup_column(codec: str, stream: np.ndarray, offs: np.ndarray, n: int, bs: int, idx: np.ndarray,
                       dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    idx = np.ascontiguousarray(idx, np.uint64)
    out = np.empty(len(idx), dtype)
    padded = _padded(stream, offs)
    lib().leco_oracle_base_lookup_column(_CODEC[codec], _ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs,
                                         _ptr(idx), len(idx), _ptr(out))
    return out


def for_filter_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int, dtype, filt: int = 0,
                      close: bool = False, f1: int = 0, f2: int = 0, base: int = 1) -> np.ndarray:
    """FOR_int<T>::filter_range (close=False: values > filt) or filter_range_close
    (v % base in (f1, f2)) over every block."""
    dtype = np.dtype(dtype)
    out = np.zeros(n, np.uint32)
    padded = _padded(stream, offs)
    c = lib().leco_oracle_for_filter_column(_ptr(padded), _ptr(offs), 8 * dtype.itemsize, n, bs, int(close),
                                            f1 if close else filt, f2, base, _ptr(out), n)
    return out[:c].copy()


# ---- piecewise_fix_op_max (uint32) ------------------------------------------------------
# fix arrays: int8 per value, +1 where the reference lists the position in
# mul_add_diff_set, -1 in mul_add_diff_set_minus, 0 elsewhere.


def opmax_block_bytes(l: int, n: int) -> int:
    return lib().leco_oracle_opmax_block_bytes(l, n)


def opmax_encode_block(y: np.ndarray) -> tuple[bytes, np.ndarray]:
    y = np.ascontiguousarray(y, np.uint32)
    out = np.zeros(len(y) * 4 + 64, np.uint8)
    fix = np.zeros(max(len(y), 1), np.int8)
    nb = lib().leco_oracle_opmax_encode_block(_ptr(y), len(y), _ptr(out), _ptr(fix))
    return out[:nb].tobytes(), fix[:len(y)].copy()


def opmax_decode_block(blk: bytes, n: int) -> np.ndarray:
    """decodeArray8's output (before the fix-up lists are applied)."""
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    out = np.zeros(max(n, 1), np.uint32)
    lib().leco_oracle_opmax_decode_block(_ptr(src), n, _ptr(out))
    return out[:n]


def opmax_lookup_block(blk: bytes, j: int) -> int:
    src = np.frombuffer(bytes(blk) + bytes(64), np.uint8)
    return int(lib().leco_oracle_opmax_lookup_block(_ptr(src), j))


def opmax_encode_column(y: np.ndarray, bs: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """-> (stream, offsets, fix)."""
    y = np.ascontiguousarray(y, np.uint32)
    n = len(y)
    nb = (n + bs - 1) // bs
    stream = np.zeros(n * 4 + nb * 20 + 128, np.uint8)
    offs = np.zeros(nb + 1, np.uint64)
    fix = np.zeros(max(n, 1), np.int8)
    total = lib().leco_oracle_opmax_encode_column(_ptr(y), n, bs, _ptr(stream), _ptr(offs), _ptr(fix))
    return stream[:total].copy(), offs, fix[:n].copy()


def opmax_decode_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int) -> np.ndarray:
    out = np.empty(max(n, 1), np.uint32)
    padded = _padded(stream, offs)
    lib().leco_oracle_opmax_decode_column(_ptr(padded), _ptr(offs), n, bs, _ptr(out))
    return out[:n]


def opmax_lookup_column(stream: np.ndarray, offs: np.ndarray, n: int, bs: int, idx: np.ndarray) -> np.ndarray:
    padded = _padded(stream, offs)
    idx = np.asarray(idx, np.uint64)
    out = np.empty(len(idx), np.uint32)
    for q, i in enumerate(idx):
        b, j = divmod(int(i), bs)
        out[q] = lib().leco_oracle_opmax_lookup_block(_ptr(padded) + int(offs[b]), j)
    return out


def opmax_fix_lists(fix: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The reference's (mul_add_diff_set, mul_add_diff_set_minus) from a fix array."""
    return np.flatnonzero(fix > 0).astype(np.int32), np.flatnonzero(fix < 0).astype(np.int32)


def opmax_apply_lists(decoded: np.ndarray, plus: np.ndarray, minus: np.ndarray) -> np.ndarray:
    """The harness's fix-up loops (experiments/fix_int.cpp:110-117)."""
    out = decoded.copy()
    np.add.at(out, plus, np.uint32(1))
    np.subtract.at(out, minus, np.uint32(1))
    return out
