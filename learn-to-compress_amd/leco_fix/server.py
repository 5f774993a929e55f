"""Request server over resident compressed columns (leco_fix_server_* C ABI).

The server loop itself is native (learn-to-compress_amd/csrc/leco_serve.hip):
continuous batching of point lookups and range decodes into shared launches,
memory admission from pools carved out of free HBM minus a reserve, and
column insert (GPU encode) / evict while traffic runs.  This module builds
request traces as numpy structured arrays laid out like leco_fix_request and
reads back per-request results and run statistics.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib

REQ_LOOKUP, REQ_RANGE, REQ_INSERT, REQ_EVICT = 0, 1, 2, 3
SERVE_BATCHED, SERVE_PLAIN = 0, 1

REQUEST_DTYPE = np.dtype([("kind", "<u4"), ("column", "<u4"), ("count", "<u8"), ("begin", "<u8"),
                          ("h_idx", "<u8"), ("h_out", "<u8"), ("d_values", "<u8"), ("block_size", "<u4"),
                          ("width_bits", "<u4"), ("arrival_s", "<f8")], align=True)
RESULT_DTYPE = np.dtype([("latency_s", "<f8"), ("queued_s", "<f8"), ("batch", "<u8"), ("status", "<i4"),
                         ("reserved", "<i4")], align=True)
STATS_FIELDS = ["wall_s", "batches", "completed", "failed", "max_batch_requests", "values_out", "bytes_out",
                "encoded_values", "encoded_bytes", "encode_gpu_s", "peak_result_bytes", "peak_column_bytes"]
STATS_DTYPE = np.dtype([(f, "<f8" if f.endswith("_s") else "<u8") for f in STATS_FIELDS], align=True)
ITERATION_FIELDS = ["between_us", "arrivals_us", "completions_us", "admission_us", "copies_us", "launch_us",
                    "rest_us", "kernel_us", "event_us"]
ITERATION_DTYPE = np.dtype([("start_s", "<f8")] + [(f, "<f4") for f in ITERATION_FIELDS] +
                           [("admitted", "<u4"), ("queued", "<u4"), ("reserved", "<u4")], align=True)
TRACE_CAP = 16
assert REQUEST_DTYPE.itemsize == 64 and RESULT_DTYPE.itemsize == 32 and STATS_DTYPE.itemsize == 96
assert ITERATION_DTYPE.itemsize == 56


class ServerConfig(ctypes.Structure):
    """leco_fix_server_config (include/leco_fix.h)."""
    _fields_ = [("slots", ctypes.c_uint32), ("max_batch_requests", ctypes.c_uint32),
                ("max_batch_values", ctypes.c_uint64), ("reserve_bytes", ctypes.c_uint64),
                ("result_pool_bytes", ctypes.c_uint64), ("column_pool_bytes", ctypes.c_uint64)]


class Server:
    def __init__(self, slots: int = 4, max_batch_requests: int = 8192, max_batch_values: int = 1 << 26,
                 reserve_bytes: int = 64 << 30, result_pool_bytes: int = 0, column_pool_bytes: int = 0):
        cfg = ServerConfig(slots, max_batch_requests, max_batch_values, reserve_bytes, result_pool_bytes,
                           column_pool_bytes)
        h = ctypes.c_void_p()
        check(lib().leco_fix_server_create(ctypes.byref(cfg), ctypes.byref(h)), "server_create")
        self._h = h
        self._keep = []  # columns attached from Python stay alive with the server

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().leco_fix_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except TypeError:  # interpreter shutdown: the module's globals are already gone
            pass

    def attach(self, slot: int, col) -> None:
        """Serves a Column (or batch column view) in `slot`."""
        check(lib().leco_fix_server_attach_column(self._h, slot, col.stream.data_ptr(), col.offsets.data_ptr(),
                                                  col.params.data_ptr(), col.n, col.block_size, col.width),
              "server_attach_column")
        self._keep.append(col)

    def column_info(self, slot: int) -> dict:
        st, offs, prm = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        n, bs, w, nbytes = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().leco_fix_server_column(self._h, slot, ctypes.byref(st), ctypes.byref(offs), ctypes.byref(prm),
                                           ctypes.byref(n), ctypes.byref(bs), ctypes.byref(w), ctypes.byref(nbytes)),
              "server_column")
        return {"stream": st.value, "offsets": offs.value, "params": prm.value, "n": n.value,
                "block_size": bs.value, "width": w.value, "stream_bytes": nbytes.value}

    def slowest_iterations(self) -> list[dict]:
        """The last batched run's slowest admission-loop iterations, slowest
        first, with the microseconds each phase took."""
        out = np.zeros(TRACE_CAP, ITERATION_DTYPE)
        cnt = ctypes.c_uint32()
        check(lib().leco_fix_server_slowest_iterations(self._h, out.ctypes.data, TRACE_CAP, ctypes.byref(cnt)),
              "server_slowest_iterations")
        rows = []
        for r in out[:cnt.value]:
            row = {"start_s": round(float(r["start_s"]), 6)}
            row.update({f: round(float(r[f]), 1) for f in ITERATION_FIELDS})
            row.update({"admitted": int(r["admitted"]), "queued": int(r["queued"])})
            rows.append(row)
        return rows

    def run(self, reqs: np.ndarray, mode: int = SERVE_BATCHED):
        """Serves the trace in real time; returns (results, stats dict)."""
        reqs = np.ascontiguousarray(reqs, REQUEST_DTYPE)
        res = np.zeros(len(reqs), RESULT_DTYPE)
        stats = np.zeros(1, STATS_DTYPE)
        check(lib().leco_fix_server_run(self._h, reqs.ctypes.data, len(reqs), mode, res.ctypes.data,
                                        stats.ctypes.data), "server_run")
        return res, {f: stats[f][0].item() for f in STATS_FIELDS}


def make_trace(rng: np.random.Generator, columns: list[tuple[int, int, int]], rate: float, nreq: int,
               lookup_frac: float = 0.8, max_lookup: int = 4096, range_lo: float = 1e2, range_hi: float = 1e7,
               keep_outputs: bool = False):
    """Poisson arrivals at `rate` req/s over `columns` = [(slot, n, width), ...]:
    lookup_frac point lookups of U[1, max_lookup] uniform random indices,
    the rest range decodes of log-uniform length in [range_lo, range_hi].
    Returns (requests, index buffer, output buffers or None)."""
    reqs = np.zeros(nreq, REQUEST_DTYPE)
    reqs["arrival_s"] = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs["arrival_s"] -= reqs["arrival_s"][0]
    which = rng.integers(0, len(columns), nreq)
    slots = np.array([c[0] for c in columns], np.uint32)[which]
    ns = np.array([c[1] for c in columns], np.uint64)[which]
    is_lookup = rng.random(nreq) < lookup_frac
    counts = np.where(is_lookup, rng.integers(1, max_lookup + 1, nreq),
                      np.exp(rng.uniform(np.log(range_lo), np.log(range_hi), nreq)).astype(np.int64))
    counts = np.minimum(counts.astype(np.uint64), ns)
    reqs["kind"] = np.where(is_lookup, REQ_LOOKUP, REQ_RANGE)
    reqs["column"] = slots
    reqs["count"] = counts
    span = ns - np.where(is_lookup, 0, counts)
    reqs["begin"] = np.where(is_lookup, 0, (rng.random(nreq) * (span + 1)).astype(np.uint64))
    nidx = int(counts[is_lookup].sum())
    idx = np.empty(nidx, np.uint64)
    pos = 0
    for i in np.nonzero(is_lookup)[0]:
        c = int(counts[i])
        idx[pos:pos + c] = (rng.random(c) * float(ns[i])).astype(np.uint64)
        reqs["h_idx"][i] = idx.ctypes.data + 8 * pos
        pos += c
    outs = None
    if keep_outputs:
        widths = np.array([c[2] for c in columns])[which]
        outs = [np.zeros(int(counts[i]), np.uint32 if widths[i] == 32 else np.uint64) for i in range(nreq)]
        for i, o in enumerate(outs):
            reqs["h_out"][i] = o.ctypes.data
    return reqs, idx, outs


def device_outputs(reqs: np.ndarray, select: np.ndarray, width: int):
    """One HBM buffer for the results of the selected lookup / range requests
    (all on columns of `width` bits): their d_out (the d_values field) points
    at consecutive slices of it, so the server writes those results straight
    into HBM.  Returns (buffer, start of each request's slice)."""
    import torch
    counts = np.where(select, reqs["count"], 0).astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    total = int(counts.sum())
    buf = torch.empty(max(total, 1), dtype=torch.int64 if width == 64 else torch.int32, device="cuda")
    esize = width // 8
    sel = np.asarray(select, bool)
    reqs["d_values"][sel] = (buf.data_ptr() + esize * starts[sel]).astype(np.uint64)
    return buf, starts


def check_device_outputs(reqs: np.ndarray, idx: np.ndarray, select: np.ndarray, buf, starts, values) -> int:
    """Mismatching values among the selected requests' device results:
    lookups against values[idx], ranges against values[begin:begin + count]
    of the column (`values`: slot -> device tensor of the raw column),
    compared on the device, one synchronization at the end."""
    import torch
    idx_dev = torch.from_numpy(idx.view(np.int64)).to("cuda") if len(idx) else None
    base = idx.ctypes.data
    bad = torch.zeros((), dtype=torch.int64, device="cuda")
    for i in np.nonzero(select)[0]:
        r = reqs[i]
        c, s = int(r["count"]), int(starts[i])
        got = buf[s:s + c]
        v = values[int(r["column"])]
        if r["kind"] == REQ_LOOKUP:
            q = (int(r["h_idx"]) - base) // 8
            want = v[idx_dev[q:q + c]]
        else:
            b = int(r["begin"])
            want = v[b:b + c]
        bad += (got != want).sum()
    return int(bad.item())
