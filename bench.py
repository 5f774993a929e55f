#!/usr/bin/env python3
"""Full-decode throughput of the gfx950 LeCo_fix codec on one MI355X.

Workload (one step = one full decode of the column, already resident in HBM):
500M sorted piecewise-linear uint64 keys (segment lengths geometric with mean
2^20, slopes log-uniform in [1, 2^20], noise sigma log-uniform in [1, 2^12]),
block size 4096 -> 122,071 blocks.  The column is generated in HBM (seeded),
encoded once by the GPU encoder, verified (decode == input, lookups ==
input[idx]) and then decoded `--steps` times between HIP events recorded on the
stream the kernels run on.

Also reported: GPU encode GB/s, 64M uniform random lookups/s, the latency
of a single random access (p50 / p99, one index per launch), the FOR / Delta
baselines, filters, files, piecewise_fix_op_max, the batched, served (p50 /
p99 per request kind) and churn workloads, and the CPU oracle's figures on a
bounded sample of the same column (decode on 1 thread and on all cores; the
other rows on 1 thread).

    python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "learn-to-compress_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(torch, fn, reps: int, warmup: int):
    """Average ms per call of fn(), HIP events on torch's current stream (the
    stream every launch goes to)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / reps, wall


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), as lscpu reports it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def config1_section(torch, cpu: bool) -> dict:
    """The reference's own workload (experiments/leco_int_fix_template.cpp:98-213):
    10M uint32 y_i = clamp(round(1000 i + N(0, 50^2)), 0, 2^32-1), seed 1,
    block size 10,000 (1000 blocks).  On the GPU: compression ratio, encode,
    full decode over 10 repeats, 1M uniform lookups, round trip and the stream
    against the oracle's bytes.  On the host (cpu=True): the oracle's encode,
    the refill decode loop (the shape of read_all_bit_fix, bit_read.h:222-275)
    over 10 repeats, and 1M lookups, on 1 thread and on all cores."""
    import numpy as np

    from leco_fix import Column, datagen
    from leco_fix._lib import lib

    sys.path.insert(0, ROOT)
    from oracle import oracle

    n, bs = 10_000_000, 10_000
    y = datagen.linear_noise_u32(n, seed=1)
    raw = 4 * n
    dev = torch.from_numpy(y.view(np.int32)).to("cuda")
    col = Column.encode(dev, bs)
    out = torch.empty_like(dev)
    col.decode(out=out)
    ok = bool(torch.equal(out, dev))
    idx_h = np.random.default_rng(1).integers(0, n, 1_000_000, dtype=np.uint64)
    idx = torch.from_numpy(idx_h.view(np.int64)).to("cuda")
    got = col.lookup(idx)
    ok = ok and bool(torch.equal(got, dev[idx]))
    nthr = min(16, os.cpu_count() or 1)
    want, offs = oracle.encode_column(y, bs, threads=nthr)
    same_stream = col.stream_bytes() == want.tobytes()

    def enc():
        lib().leco_fix_encode_column_u32(dev.data_ptr(), n, bs, col.stream.data_ptr(), col.stream.numel(),
                                         col.offsets.data_ptr(), col.params.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream)
    ms_enc, _ = timed(torch, enc, 10, 2)
    ms_dec, _ = timed(torch, lambda: col.decode(out=out), 10, 2)
    ms_lk, _ = timed(torch, lambda: col.lookup(idx, out=got), 10, 2)
    col.decode(out=out)
    ok = ok and bool(torch.equal(out, dev))
    res = {"workload": "10M uint32, y_i = clamp(round(1000 i + N(0, 50^2))), seed 1, block size 10000",
           "compression_ratio": round(col.nbytes / raw, 5), "compressed_bytes": col.nbytes,
           "round_trip_exact": ok, "stream_equals_oracle": same_stream,
           "gpu": {"encode_gbs": round(raw / (ms_enc * 1e-3) / 1e9, 1), "encode_ms": round(ms_enc, 4),
                   "decode_gbs": round(raw / (ms_dec * 1e-3) / 1e9, 1), "decode_ms": round(ms_dec, 4),
                   "lookups": len(idx_h), "lookup_batch_ms": round(ms_lk, 4),
                   "lookups_per_s": round(len(idx_h) / (ms_lk * 1e-3), 0),
                   "ns_per_lookup_in_batch": round(ms_lk * 1e6 / len(idx_h), 4)}}
    if cpu:
        host = {"threads_all": nthr, "cpu_model": cpu_model()}
        for tag, thr in (("1_thread", 1), ("all_cores", nthr)):
            t0 = time.perf_counter()
            st, of = oracle.encode_column(y, bs, threads=thr)
            t_enc = time.perf_counter() - t0
            dec = np.empty_like(y)
            oracle.decode_column(st, of, n, bs, np.uint32, threads=thr, out=dec, refill=True)  # warm-up
            t0 = time.perf_counter()
            for _ in range(10):
                oracle.decode_column(st, of, n, bs, np.uint32, threads=thr, out=dec, refill=True)
            t_dec = (time.perf_counter() - t0) / 10
            t0 = time.perf_counter()
            lk = oracle.lookup_column(st, of, n, bs, idx_h, np.uint32, threads=thr)
            t_lk = time.perf_counter() - t0
            host[tag] = {"encode_gbs": round(raw / t_enc / 1e9, 3), "decode_gbs": round(raw / t_dec / 1e9, 3),
                         "decode_ns_per_value": round(t_dec * 1e9 / n, 3),
                         "ns_per_lookup": round(t_lk * 1e9 / len(idx_h), 2),
                         "round_trip_exact": bool(np.array_equal(dec, y)) and bool(np.array_equal(lk, y[idx_h])),
                         "stream_equals_gpu": st.tobytes() == want.tobytes()}
        res["cpu"] = host
    del dev, out, idx, got, col
    torch.cuda.empty_cache()
    return res


def dropin_section(vals, bs: int, blocks: int, col_encode_ms: float, col_decode_ms: float) -> dict:
    """The per-block host-buffer calls a reference driver reaches through the
    drop-in Leco_int<T> header (encodeArray8_int / decodeArray8 /
    randomdecodeArray8, one call per block or lookup, each staging its block
    through the GPU and synchronizing), timed over the column's first `blocks`
    blocks, against the column calls on the whole column."""
    import ctypes

    import numpy as np

    from leco_fix._lib import lib

    L = lib()
    y = vals[:blocks * bs].cpu().numpy().view(np.uint64)
    cap = int(L.leco_fix_stream_bound(bs, bs, 64))
    enc = [np.empty(cap, np.uint8) for _ in range(blocks)]
    sizes = []
    written = ctypes.c_uint64()
    L.leco_fix_encode_block_u64(y.ctypes.data, bs, enc[0].ctypes.data, cap, ctypes.byref(written))  # warm-up
    t0 = time.perf_counter()
    for b in range(blocks):
        L.leco_fix_encode_block_u64(y[b * bs:].ctypes.data, bs, enc[b].ctypes.data, cap, ctypes.byref(written))
        sizes.append(written.value)
    t_enc = (time.perf_counter() - t0) / blocks
    out = np.empty(bs, np.uint64)
    ok = True
    t0 = time.perf_counter()
    for b in range(blocks):
        L.leco_fix_decode_block_u64(enc[b].ctypes.data, bs, out.ctypes.data)
    t_dec = (time.perf_counter() - t0) / blocks
    ok = ok and bool(np.array_equal(out, y[(blocks - 1) * bs:blocks * bs]))
    rng = np.random.default_rng(3)
    qs = rng.integers(0, blocks * bs, 2000)
    v = ctypes.c_uint64()
    t0 = time.perf_counter()
    for q in qs:
        L.leco_fix_random_decode_u64(enc[int(q) // bs].ctypes.data, int(q) % bs, ctypes.byref(v))
    t_ra = (time.perf_counter() - t0) / len(qs)
    ok = ok and v.value == int(y[int(qs[-1])])
    nb = (vals.numel() + bs - 1) // bs
    return {"blocks_timed": blocks, "encode_block_us": round(t_enc * 1e6, 1), "decode_block_us": round(t_dec * 1e6, 1),
            "random_decode_us": round(t_ra * 1e6, 1), "verified": ok,
            "column_by_blocks_encode_s": round(t_enc * nb, 2), "column_by_blocks_decode_s": round(t_dec * nb, 2),
            "column_call_encode_ms": round(col_encode_ms, 3), "column_call_decode_ms": round(col_decode_ms, 4),
            "note": "host buffers: each call copies its block to HBM, runs the kernels and synchronizes; the "
                    "column calls above work on HBM-resident data"}


def cpu_baseline(vals_dev, bs: int, sample: int, repeat: int):
    """Decode of the first `sample` values of the column on the host: the
    oracle's refill loop (the loop shape of the reference's read_all_bit_fix,
    bit_read.h:222-275) on 1 thread and on all cores, and the byte-window
    decode of the checker on 1 thread."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import oracle

    y = vals_dev[:sample].cpu().numpy().view(np.uint64)
    stream, offs = oracle.encode_column(y, bs, threads=min(16, os.cpu_count() or 1))
    padded = np.concatenate([stream, np.zeros(64, np.uint8)])
    out = np.empty(len(y), np.uint64)
    oracle.decode_column(padded, offs, len(y), bs, np.uint64, threads=1, out=out, refill=True)  # warm-up
    t0 = time.perf_counter()
    for _ in range(repeat):
        oracle.decode_column(padded, offs, len(y), bs, np.uint64, threads=1, out=out, refill=True)
    t1 = time.perf_counter()
    assert np.array_equal(out, y)
    gbs1 = len(y) * 8 * repeat / (t1 - t0) / 1e9
    nthr = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    for _ in range(repeat):
        oracle.decode_column(padded, offs, len(y), bs, np.uint64, threads=nthr, out=out, refill=True)
    t1 = time.perf_counter()
    gbsn = len(y) * 8 * repeat / (t1 - t0) / 1e9
    assert np.array_equal(out, y)
    t0 = time.perf_counter()
    oracle.decode_column(padded, offs, len(y), bs, np.uint64, threads=1, out=out)
    gbs_window = len(y) * 8 / (time.perf_counter() - t0) / 1e9
    # filter_range at the sample's median, 1 thread
    f = int(y[len(y) // 2])
    oracle.filter_column(padded, offs, len(y), bs, np.uint64, f)
    t0 = time.perf_counter()
    for _ in range(max(1, repeat // 2)):
        got = oracle.filter_column(padded, offs, len(y), bs, np.uint64, f)
    t1 = time.perf_counter()
    assert len(got) == int(np.count_nonzero(y > np.uint64(f)))
    gbsf = len(y) * 8 * max(1, repeat // 2) / (t1 - t0) / 1e9
    return gbs1, gbsn, nthr, gbsf, gbs_window


def cpu_rows(vals_dev, bs: int, sample: int) -> dict:
    """The other rows' CPU figures, 1 thread, on the same sample: the C
    oracle's encode, 1M random lookups, filter_range_close with the driver's
    zonemap filters (period 2^20), and FOR / Delta decode."""
    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    from leco_fix.column import close_filters_from_zonemap
    from oracle import oracle

    y = vals_dev[:sample].cpu().numpy().view(np.uint64)
    n = len(y)
    out = {}
    t0 = time.perf_counter()
    stream, offs = oracle.encode_column(y, bs, threads=1)
    out["encode_gbs"] = round(n * 8 / (time.perf_counter() - t0) / 1e9, 3)
    idx = np.random.default_rng(7).integers(0, n, 1_000_000).astype(np.uint64)
    t0 = time.perf_counter()
    got = oracle.lookup_column(stream, offs, n, bs, idx, np.uint64, threads=1)
    out["lookups_per_s"] = round(len(idx) / (time.perf_counter() - t0), 0)
    assert np.array_equal(got, y[idx])
    base = 1 << 20
    full = n // bs * bs
    mins = torch.from_numpy(y[:full].view(np.int64).reshape(-1, bs).min(axis=1))
    if full < n:
        mins = torch.cat([mins, torch.tensor([int(y[full:].min())], dtype=torch.int64)])
    f1, f2 = close_filters_from_zonemap(mins, base // 4, base // 2, base, 64)
    t0 = time.perf_counter()
    pos = oracle.filter_close_column(stream, offs, n, bs, np.uint64, f1.numpy().view(np.uint64),
                                     f2.numpy().view(np.uint64), base)
    out["filter_range_close_gbs"] = round(n * 8 / (time.perf_counter() - t0) / 1e9, 3)
    r = y % np.uint64(base)
    assert len(pos) == int(np.count_nonzero((r > base // 4) & (r < base // 2)))
    for codec in ("for", "delta"):
        st, of = oracle.base_encode_column(codec, y, bs)
        t0 = time.perf_counter()
        dec = oracle.base_decode_column(codec, st, of, n, bs, np.uint64)
        out[f"{codec}_decode_gbs"] = round(n * 8 / (time.perf_counter() - t0) / 1e9, 3)
        assert np.array_equal(dec, y)
    return out


def baselines_section(torch, vals, bs: int, raw: int, leco_bytes: int, leco_ms: float, steps: int) -> dict:
    """FOR_int and Delta_int (the reference's fixed-partition baselines) on the
    same column and block size: compression ratio, encode / decode GB/s,
    lookups (Delta's cost per lookup grows with the position in the block, as
    read_Delta_int's does), verified round trips."""
    from leco_fix.baselines import BaseColumn

    n = vals.numel()
    out = {"leco": {"compression_ratio": round(leco_bytes / raw, 5), "decode_gbs": round(raw / (leco_ms * 1e-3) / 1e9, 1)}}
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    for codec, nq in (("for", 64_000_000), ("delta", 4_000_000)):
        col = BaseColumn.encode(codec, vals, bs)
        dec = torch.empty_like(vals)
        ms_enc, _ = timed(torch, lambda: BaseColumn.encode(codec, vals, bs), 3, 1)
        ms_dec, _ = timed(torch, lambda: col.decode(out=dec), steps, 2)
        ok = bool(torch.equal(dec, vals))
        idx = torch.randint(0, n, (nq,), device="cuda", generator=g)
        res = torch.empty(nq, dtype=vals.dtype, device="cuda")
        ms_lk, _ = timed(torch, lambda: col.lookup(idx, out=res), 3, 1)
        ok = ok and bool(torch.equal(res, vals[idx]))
        out[codec] = {"compression_ratio": round(col.nbytes / raw, 5), "compressed_bytes": col.nbytes,
                      "encode_gbs": round(raw / (ms_enc * 1e-3) / 1e9, 1), "encode_ms_incl_alloc": round(ms_enc, 3),
                      "decode_gbs": round(raw / (ms_dec * 1e-3) / 1e9, 1), "decode_ms": round(ms_dec, 4),
                      "lookups": nq, "lookups_per_s": round(nq / (ms_lk * 1e-3), 0), "verified": ok}
        if codec == "for":
            fout = torch.empty(n, dtype=torch.int32, device="cuda")
            f = int(vals[n // 2].item())
            ms_f, _ = timed(torch, lambda: col.filter_range(f, capacity=n), 3, 1)
            out[codec]["filter_range_ms_incl_alloc"] = round(ms_f, 3)
            del fout
        del col, dec, idx, res
        torch.cuda.empty_cache()
    return out


def opmax_section(torch, n: int, steps: int, cpu_sample: int) -> dict:
    """piecewise_fix_op_max, the uint32 IntegerCODEC, on the shape of its
    published runs (benchmark/fix_int.sh: 200M uint32 keys in 100,000 blocks,
    bs 2000, and in 1,000,000 blocks, bs 200): synthetic linear keys with
    noise.  Encode (fit, pack, fix-up flags), decode with the fix-ups applied
    (the harness's output), the decoder alone, lookups, list compaction; the C
    oracle's decode + list correction on a sample as the CPU figure."""
    import numpy as np

    from leco_fix.opmax import OpMaxColumn

    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    i = torch.arange(n, device="cuda", dtype=torch.float64)
    y = torch.round(1e6 + 21.3 * i + torch.randint(-3, 4, (n,), device="cuda", generator=g)).to(torch.int64)
    y = y.to(torch.int32)
    del i
    torch.cuda.empty_cache()
    raw = 4 * n
    out = {"workload": f"{n // 1_000_000}M uint32 linear keys with noise (y = round(1e6 + 21.3 i + U[-3,3]))",
           "reference_published": {"source": "benchmark/fix_int_benchmark_intel.log:4 (linear_200M_uint32, "
                                             "100000 blocks, 1 CPU thread)",
                                   "da_ns": 1.75531, "ra_ns": 28.285, "cr_percent": 4.48606,
                                   "compression_gbs": 0.931192}}
    for bs in (2000, 200):
        col = OpMaxColumn.encode(y, bs)
        plus, minus = col.fix_lists()  # the reference's lists, kept with the column
        dec = torch.empty_like(y)
        ms_enc, _ = timed(torch, lambda: OpMaxColumn.encode(y, bs), 3, 1)
        ms_dec, _ = timed(torch, lambda: col.decode_lists(plus, minus, out=dec), steps, 2)
        ok = bool(torch.equal(dec, y))
        ms_dense, _ = timed(torch, lambda: col.decode(fixed=True, out=dec), steps, 2)
        ok = ok and bool(torch.equal(dec, y))
        ms_raw, _ = timed(torch, lambda: col.decode(fixed=False, out=dec), steps, 2)
        nq = 16_000_000
        idx = torch.randint(0, n, (nq,), device="cuda", generator=g)
        res = torch.empty(nq, dtype=torch.int32, device="cuda")
        ms_lk, _ = timed(torch, lambda: col.lookup(idx, out=res), 3, 1)
        ok = ok and bool(torch.equal(res, y[idx]))
        ms_lists, _ = timed(torch, lambda: col.fix_lists(), 3, 1)
        alg = col.nbytes + raw  # stream read, values written (the lists are sparse)
        out[f"bs{bs}"] = {
            "blocks": col.num_blocks, "compression_ratio": round(col.compression_ratio, 5),
            "compressed_bytes": col.nbytes, "fixups": [len(plus), len(minus)],
            "encode_gbs": round(raw / (ms_enc * 1e-3) / 1e9, 1), "encode_ms_incl_alloc": round(ms_enc, 3),
            "decode_gbs": round(raw / (ms_dec * 1e-3) / 1e9, 1), "decode_ms": round(ms_dec, 4),
            "decode_hbm_gbs": round(alg / (ms_dec * 1e-3) / 1e9, 1),
            "decode_hbm_frac": round(alg / (ms_dec * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "decode_ns_per_value": round(ms_dec * 1e6 / n, 5),
            "decode_dense_flags_gbs": round(raw / (ms_dense * 1e-3) / 1e9, 1),
            "decoder_only_gbs": round(raw / (ms_raw * 1e-3) / 1e9, 1),
            "lookups": nq, "lookups_per_s": round(nq / (ms_lk * 1e-3), 0),
            "fix_lists_ms_incl_alloc": round(ms_lists, 3), "verified": ok}
        del col, dec, idx, res, plus, minus
        torch.cuda.empty_cache()
    if cpu_sample > 0:
        sys.path.insert(0, ROOT)
        from oracle import oracle
        m = min(cpu_sample, n)
        yh = y[:m].cpu().numpy().view(np.uint32)
        stream, offs, fix = oracle.opmax_encode_column(yh, 2000)
        plus, minus = oracle.opmax_fix_lists(fix)
        t0 = time.perf_counter()
        d = oracle.opmax_decode_column(stream, offs, m, 2000)
        d = oracle.opmax_apply_lists(d, plus, minus)
        dt = time.perf_counter() - t0
        out["cpu_oracle"] = {"sample": f"first {m} values, bs 2000: decode + list correction, C oracle, 1 thread",
                             "gbs": round(4 * m / dt / 1e9, 3), "ns_per_value": round(dt * 1e9 / m, 3),
                             "verified": bool(np.array_equal(d, yh))}
    del y
    torch.cuda.empty_cache()
    return out


def io_section(torch, keys, bs: int) -> dict:
    """SOSD binary -> HBM -> encode -> column file -> HBM, in a temporary
    directory (the host page cache serves the reads)."""
    import tempfile

    from leco_fix import Column, io

    raw = keys.numel() * 8
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
        sosd, colf = os.path.join(d, "keys.sosd"), os.path.join(d, "keys.leco")
        t0 = time.perf_counter()
        io.save_sosd(sosd, keys)
        t1 = time.perf_counter()
        loaded = io.load_sosd(sosd, 64)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        col = Column.encode(loaded, bs)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        io.save_column(colf, col)
        t4 = time.perf_counter()
        back = io.load_column(colf)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        ok = bool(torch.equal(loaded, keys)) and back.stream_bytes() == col.stream_bytes()
        comp = col.nbytes
    return {"keys": keys.numel(), "raw_bytes": raw, "column_bytes": comp, "verified": ok,
            "sosd_save_gbs": round(raw / (t1 - t0) / 1e9, 2), "sosd_load_gbs": round(raw / (t2 - t1) / 1e9, 2),
            "encode_gbs": round(raw / (t3 - t2) / 1e9, 1),
            "column_save_gbs": round(comp / (t4 - t3) / 1e9, 2),
            "column_load_gbs": round(comp / (t5 - t4) / 1e9, 2),
            "column_load_raw_equiv_gbs": round(raw / (t5 - t4) / 1e9, 2),
            "note": "wall clock; file bytes through the host page cache and PCIe; encode includes allocation"}


def filter_section(torch, vals, col, reps: int) -> dict:
    """filter_range at the median key and filter_range_close with the reference
    driver's zonemap-derived per-block filters, on the resident column; one
    launch each, preallocated output and workspace, HIP events."""
    from leco_fix._lib import lib
    from leco_fix.column import _stream_handle, close_filters_from_zonemap

    L = lib()
    n, bs = col.n, col.block_size
    out = torch.empty(n + col.num_blocks, dtype=torch.int32, device="cuda")
    res = torch.zeros(2, dtype=torch.int64, device="cuda")
    wb = int(L.leco_fix_filter_workspace_bytes(n, bs))
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    fmed = int(vals[n // 2].item())

    def gt():
        L.leco_fix_filter_range_u64(col.stream.data_ptr(), col.offsets.data_ptr(), n, bs, fmed, out.data_ptr(),
                                    out.numel(), res.data_ptr(), work.data_ptr(), wb,
                                    _stream_handle(None))
    ms_gt, _ = timed(torch, gt, reps, 1)
    cnt = int(res[0].item())
    expect = int((vals > fmed).sum().item())
    ok_gt = cnt == expect and (cnt == 0 or int(out[0].item()) == n - cnt)

    # period 2^20 over keys up to ~2^45: a few hundred periods per block
    base = 1 << 20
    f1, f2 = base // 4, base // 2
    nb = col.num_blocks
    full = (n // bs) * bs
    mins = vals[:full].view(-1, bs).amin(dim=1)
    if full < n:
        mins = torch.cat([mins, vals[full:].amin().view(1)])
    t1, t2 = close_filters_from_zonemap(mins, f1, f2, base, 64)

    def close():
        L.leco_fix_filter_range_close_u64(col.stream.data_ptr(), col.offsets.data_ptr(), n, bs, t1.data_ptr(),
                                          t2.data_ptr(), base, out.data_ptr(), out.numel(), res.data_ptr(),
                                          work.data_ptr(), wb, _stream_handle(None))
    ms_close, _ = timed(torch, close, reps, 1)
    ccnt, flags = (int(v) for v in res.tolist())
    r = vals % base
    truecount = int(((r > f1) & (r < f2)).sum().item())
    del r
    raw = n * 8
    comp = col.nbytes
    return {
        "filter_range": {"filter": fmed, "matches": cnt, "ms": round(ms_gt, 4),
                         "gbs": round(raw / (ms_gt * 1e-3) / 1e9, 1),
                         "algorithmic_gbs": round((comp + 4 * cnt) / (ms_gt * 1e-3) / 1e9, 1),
                         "verified": ok_gt},
        "filter_range_close": {"filter1": f1, "filter2": f2, "base": base, "matches": ccnt,
                               "driver_truecount": truecount, "flags": flags, "ms": round(ms_close, 4),
                               "gbs": round(raw / (ms_close * 1e-3) / 1e9, 1),
                               "algorithmic_gbs": round((comp + 4 * ccnt) / (ms_close * 1e-3) / 1e9, 1)},
        "note": "gbs = uncompressed bytes covered / time; algorithmic_gbs = (compressed stream + 4 B per match) / "
                "time, an upper bound on the bytes the scanned windows need",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=500_000_000)
    ap.add_argument("--block-size", type=int, default=4096)
    ap.add_argument("--lookups", type=int, default=64_000_000)
    ap.add_argument("--encode-reps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=64_000_000)
    ap.add_argument("--cpu-repeat", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the reference's own workload (10M uint32, block size 10000)")
    ap.add_argument("--copy-reps", type=int, default=3,
                    help="plain device copies of the decoded column (a known-size HBM reference kernel)")
    ap.add_argument("--no-filter", action="store_true", help="skip the filter_range / filter_range_close section")
    ap.add_argument("--no-baselines", action="store_true", help="skip the FOR_int / Delta_int section")
    ap.add_argument("--opmax-n", type=int, default=200_000_000,
                    help="values of the uint32 piecewise_fix_op_max section (0 = skip)")
    ap.add_argument("--io-values", type=int, default=128_000_000,
                    help="keys written / read through SOSD and column files (0 = skip)")
    ap.add_argument("--filter-reps", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched many-column section")
    ap.add_argument("--batched-cols", type=int, default=800)
    ap.add_argument("--batched-max", type=float, default=1e8, help="largest column length (log-uniform from 1e3)")
    ap.add_argument("--batched-small-cols", type=int, default=20000,
                    help="second batched workload: this many columns of 1e3..1e5 values (0 = skip)")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-server section")
    ap.add_argument("--serve-nreq", type=int, default=60_000)
    ap.add_argument("--no-churn", action="store_true", help="skip the insert/evict-under-traffic section")
    ap.add_argument("--churn-cols", type=int, default=2000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the codec has no cross-GPU path in this build)", file=sys.stderr)
        return 2

    import torch

    from leco_fix import Column, datagen

    n, bs = args.n, args.block_size
    log(f"generate sorted_pwl_u64 n={n} bs={bs} in HBM")
    vals = datagen.sorted_pwl_u64(n, seed=2, device="cuda")
    torch.cuda.synchronize()

    log("encode + verify")
    col = Column.encode(vals, bs)
    out = torch.empty_like(vals)
    col.decode(out=out)
    verified = bool(torch.equal(out, vals))
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    idx = torch.randint(0, n, (args.lookups,), device="cuda", generator=g, dtype=torch.int64)
    lk = col.lookup(idx)
    verified = verified and bool(torch.equal(lk, vals[idx]))
    if not verified:
        log("VERIFICATION FAILED")

    raw_bytes = n * 8
    comp = col.nbytes

    log(f"decode: warmup {args.warmup}, steps {args.steps}")
    ms_dec, wall_dec = timed(torch, lambda: col.decode(out=out), args.steps, args.warmup)

    from leco_fix.column import DECODE_VALUES
    ms_dec_values, _ = timed(torch, lambda: col.decode(out=out, shape=DECODE_VALUES), args.steps, 1)
    verified = verified and bool(torch.equal(out, vals))

    log("encode timing")
    st_buf, off_buf, prm_buf = col.stream, col.offsets, col.params

    def enc():
        from leco_fix._lib import lib
        lib().leco_fix_encode_column_u64(vals.data_ptr(), n, bs, st_buf.data_ptr(), st_buf.numel(),
                                         off_buf.data_ptr(), prm_buf.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream)
    ms_enc, _ = timed(torch, enc, args.encode_reps, 1)
    # the encoder rewrote the same bytes; check once more
    col.decode(out=out)
    verified = verified and bool(torch.equal(out, vals))

    log("lookup timing")
    ms_lk, _ = timed(torch, lambda: col.lookup(idx, out=lk), 5, 1)
    # the grouped variant (queries sorted into value ranges first): fewer HBM bytes, same results
    lkg = col.lookup(idx, grouped=True)
    verified = verified and bool(torch.equal(lkg, lk))
    ms_lkg, _ = timed(torch, lambda: col.lookup(idx, out=lkg, grouped=True), 5, 1)
    del lkg
    # random-access latency: one index per call, launch to result on the host (plain path)
    one = lk[:1].clone()
    for q in range(50):  # warm-up
        col.lookup(idx[q:q + 1], out=one)
    torch.cuda.synchronize()
    lat_us = []
    for q in range(2000):
        t0 = time.perf_counter()
        col.lookup(idx[q:q + 1], out=one)
        torch.cuda.synchronize()
        lat_us.append((time.perf_counter() - t0) * 1e6)
    lat_us.sort()
    ra_p50, ra_p99 = lat_us[len(lat_us) // 2], lat_us[int(len(lat_us) * 0.99)]

    log("drop-in per-block calls")
    dropin = dropin_section(vals, bs, 256, ms_enc, ms_dec)
    log(f"dropin: {dropin}")

    config1 = None
    if not args.no_config1:
        log("config 1: 10M uint32, block size 10000, GPU and host")
        config1 = config1_section(torch, not args.no_cpu)
        log(f"config1: {config1}")

    base = None
    if not args.no_baselines:
        log("baselines: FOR_int / Delta_int on the same column")
        base = baselines_section(torch, vals, bs, raw_bytes, comp, ms_dec, args.steps)
        log(f"baselines: {base}")

    opmax = None
    if args.opmax_n > 0:
        log("piecewise_fix_op_max (uint32)")
        opmax = opmax_section(torch, args.opmax_n, args.steps, 0 if args.no_cpu else 20_000_000)
        log(f"opmax: {opmax}")

    files = None
    if args.io_values > 0:
        log("file timing")
        files = io_section(torch, vals[:args.io_values], bs)
        log(f"files: {files}")

    filt = None
    if not args.no_filter:
        log("filter timing")
        filt = filter_section(torch, vals, col, args.filter_reps)
        log(f"filter: {filt}")

    copy_gbs = fill_gbs = None
    if args.copy_reps > 0:
        log("device copy and fill references")
        dst = torch.empty_like(out)
        ms_cp, _ = timed(torch, lambda: dst.copy_(out), args.copy_reps, 1)
        copy_gbs = 2 * raw_bytes / (ms_cp * 1e-3) / 1e9
        # the runtime writing the decoded column's N x 8 bytes and reading nothing: the rate the
        # decode's output alone could be written at
        ms_fill, _ = timed(torch, lambda: dst.fill_(7), args.copy_reps, 1)
        fill_gbs = raw_bytes / (ms_fill * 1e-3) / 1e9
        del dst

    del lk, idx
    torch.cuda.empty_cache()
    from leco_fix import workloads

    batched = serving = churn = None
    if not args.no_batched:
        cols = workloads.batched_columns(args.batched_cols, 1e3, args.batched_max)
        batched = workloads.run_batched(cols, log=log)
        torch.cuda.empty_cache()
        # the same columns with the bytes-only tuner, for comparison
        batched["bytes_only_tuner"] = workloads.run_batched(cols, tuner="bytes", plain=False, log=log)
        del cols
        torch.cuda.empty_cache()
        log(f"batched: {batched}")
        if args.batched_small_cols > 0:
            # many short columns: where one launch per column from a host loop is launch-bound
            cols = workloads.batched_columns(args.batched_small_cols, 1e3, 1e5, seed=7)
            small = workloads.run_batched(cols, log=log)
            small["workload"] = f"{args.batched_small_cols} columns, lengths log-uniform in [1e3, 1e5]"
            batched["small_columns"] = small
            del cols
            torch.cuda.empty_cache()
            log(f"batched small columns: {small}")
    if not args.no_serving:
        serving = workloads.run_serving(col, vals, nreq=args.serve_nreq, log=log)
        log(f"serving: {serving}")
    if not args.no_churn:
        churn = workloads.run_churn(ncols=args.churn_cols, log=log)
        torch.cuda.empty_cache()
        log(f"churn: {churn}")

    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        log(f"cpu baseline: oracle decode of the first {sample} values, {args.cpu_repeat} repeats")
        gbs1, gbsn, nthr, gbsf, gbsw = cpu_baseline(vals, bs, sample, args.cpu_repeat)
        cpu = {"value": round(gbs1, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} values of the column ({(sample + bs - 1) // bs} blocks), "
                         f"decoded {args.cpu_repeat}x, 1 thread, by the oracle's refill loop (the loop shape of "
                         "read_all_bit_fix, bit_read.h:222-275: 64-bit words into a 128-bit register, "
                         "fma + CVTTSD2SI per value)",
               "all_cores_value": round(gbsn, 3), "all_cores": nthr,
               "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
               "checker_byte_window_decode_value": round(gbsw, 3),
               "filter_range_value": round(gbsf, 3)}
        log("cpu baseline: encode, lookups, filter_range_close, FOR / Delta decode on the same sample, 1 thread")
        cpu["rows_1_thread"] = cpu_rows(vals, bs, sample)

    traffic, traffic_src = None, None
    pmc_file = os.path.join(ROOT, "profiles", f"decode_pwl{n // 1_000_000}M_u64_bs{bs}_pmc.json")
    if os.path.exists(pmc_file):
        with open(pmc_file) as f:
            k = json.load(f)["kernels"].get("k_decode")
        if k:
            traffic = int(k["read_bytes"] + k["write_bytes"])
            traffic_src = os.path.relpath(pmc_file, ROOT) + (" (rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B_sum, "
                                                             "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum; x 32 B)")

    dec_gbs = raw_bytes / (ms_dec * 1e-3) / 1e9
    achieved = (comp + raw_bytes) / (ms_dec * 1e-3) / 1e9
    line = {
        "metric": "decode GB/s (uncompressed)",
        "value": round(dec_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_dec, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded, generated in HBM)",
        "config": {
            "workload": "full decode, 500M sorted piecewise-linear uint64 keys, block size 4096",
            "n": n, "block_size": bs, "blocks": col.num_blocks,
            "compressed_bytes": comp, "compression_ratio": round(comp / raw_bytes, 5),
            "verified_bit_exact": verified,
            "encode_gbs": round(raw_bytes / (ms_enc * 1e-3) / 1e9, 2),
            "encode_ms": round(ms_enc, 3),
            "lookups": args.lookups,
            "lookups_per_s": round(args.lookups / (ms_lk * 1e-3), 0),
            "lookups_grouped_per_s": round(args.lookups / (ms_lkg * 1e-3), 0),
            "lookup_batch_ms": round(ms_lk, 3),
            "random_access_single_p50_us": round(ra_p50, 1),
            "random_access_single_p99_us": round(ra_p99, 1),
            "decode_wall_ms_per_step": round(wall_dec * 1e3 / args.steps, 4),
            "decode_value_parallel_ms": round(ms_dec_values, 4),
            "device_copy_gbs": round(copy_gbs, 1) if copy_gbs else None,
            "device_fill_gbs": round(fill_gbs, 1) if fill_gbs else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "k_decode<uint64_t,64>",
            "bytes_per_launch": comp + raw_bytes,
        },
        "cpu_baseline": cpu,
        "config1": config1,
        "dropin_per_block": dropin,
        "baselines": base,
        "opmax": opmax,
        "filter": filt,
        "files": files,
        "batched": batched,
        "serving": serving,
        "churn": churn,
    }
    print(json.dumps(line), flush=True)
    return 0 if verified else 1


if __name__ == "__main__":
    sys.exit(main())
