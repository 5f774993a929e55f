"""Drivers for the batched, served and churn measurements (used by bench.py
and, at small scale, by the GPU tests).

Column generators run in HBM with torch (seeded): a 40+ GB batch cannot be
built on the host in bench time.  Distributions:
- sorted_uniform: uniform draws, sorted;
- sorted_zipf: cumulative sum of Zipf(1.2)-distributed gaps, drawn with the
  continuous Pareto inverse CDF floor(U^(-1/0.2)) and capped at 2^20;
- linear_sigma: slope*i + N(0, sigma), slope and sigma log-uniform;
- stepwise: runs of Poisson(500)+1 equal values, jumps uniform in [1, 2^16];
- drift: a*i + b*i^2 + N(0, 2^10) (variance 2^10), a ~ U[1, 2^16],
  b ~ logU[1e-6, 1e-3], clamped at 0.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import server as srvmod
from .batch import AUTOTUNE_CANDIDATES, Batch, autotune_block_sizes, autotune_columns

U32_MAX = (1 << 32) - 1
DISTS = ("sorted_uniform", "sorted_zipf", "linear_sigma", "stepwise")


def _gen(seed: int) -> torch.Generator:
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return g


def _finish(x: torch.Tensor, width: int) -> torch.Tensor:
    hi = float(U32_MAX) if width == 32 else 9.2e18
    v = torch.clamp(torch.round(x), 0.0, hi)
    if width == 32:
        return v.to(torch.int64).to(torch.int32)  # values < 2^32 keep their bits in int32 storage
    return v.to(torch.int64)


def gen_column(dist: str, n: int, width: int, seed: int) -> torch.Tensor:
    """One column in HBM (int32 storage for uint32, int64 for uint64)."""
    g = _gen(seed)
    rng = np.random.default_rng(seed)
    if dist == "sorted_uniform":
        hi = U32_MAX if width == 32 else (1 << 62)
        v = torch.randint(0, hi, (n,), device="cuda", generator=g, dtype=torch.int64)
        v, _ = torch.sort(v)
    elif dist == "sorted_zipf":
        u = torch.rand(n, device="cuda", generator=g, dtype=torch.float64).clamp_min(1e-300)
        gaps = torch.clamp(torch.floor(u.pow(-1.0 / 0.2)), 1.0, float(1 << 20)).to(torch.int64)
        v = torch.cumsum(gaps, 0)
        if width == 32:
            v = torch.clamp(v, max=U32_MAX)
    elif dist == "linear_sigma":
        hi = float(U32_MAX) if width == 32 else float(1 << 52)
        slope = math.exp(rng.uniform(0.0, math.log(max(hi / max(n, 1), 2.0))))
        sigma = math.exp(rng.uniform(0.0, math.log(1 << 10)))
        i = torch.arange(n, device="cuda", dtype=torch.float64)
        x = torch.clamp(slope * i + sigma * torch.randn(n, device="cuda", generator=g, dtype=torch.float64), max=hi)
        return _finish(x, width)
    elif dist == "stepwise":
        runs = torch.poisson(torch.full((n // 300 + 2,), 500.0, device="cuda", dtype=torch.float64), generator=g)
        runs = runs.to(torch.int64) + 1
        jumps = torch.randint(1, 1 << 16, (runs.numel(),), device="cuda", generator=g, dtype=torch.int64)
        levels = torch.cumsum(jumps, 0)
        v = torch.repeat_interleave(levels, runs)[:n]
        if v.numel() < n:
            v = torch.cat([v, v[-1:].expand(n - v.numel())])
        if width == 32:
            v = torch.clamp(v, max=U32_MAX)
    elif dist == "drift":
        a = rng.uniform(1.0, float(1 << 16))
        b = math.exp(rng.uniform(math.log(1e-6), math.log(1e-3)))
        i = torch.arange(n, device="cuda", dtype=torch.float64)
        x = a * i + b * i * i + 32.0 * torch.randn(n, device="cuda", generator=g, dtype=torch.float64)
        return _finish(x, width)
    else:
        raise ValueError(dist)
    return v.to(torch.int32) if width == 32 else v


def log_uniform_lengths(rng: np.random.Generator, count: int, lo: float, hi: float) -> np.ndarray:
    return np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(np.int64)


def batched_columns(ncols: int = 800, lo: float = 1e3, hi: float = 1e8, seed: int = 3):
    """ncols columns, lengths log-uniform in [lo, hi], half uint32 / half
    uint64, the four distributions in equal shares."""
    rng = np.random.default_rng(seed)
    lengths = log_uniform_lengths(rng, ncols, lo, hi)
    cols = []
    for c in range(ncols):
        width = 32 if c % 2 else 64
        dist = DISTS[(c // 2) % 4]
        cols.append((gen_column(dist, int(lengths[c]), width, seed * 100003 + c), dist, width))
    return cols


def _timed(fn, reps: int, warmup: int = 1) -> float:
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def run_batched(cols, reps: int = 5, tuner: str = "shape", plain: bool = True, log=print) -> dict:
    """Batched (one persistent launch, graph replay) vs plain per-column launches.
    tuner "shape": block size and decode tile per column by measured decode
    cost within 2% of the fewest bytes (autotune_columns); "bytes": the
    fewest encoded bytes, the largest size within 1% (autotune_block_sizes)."""
    values = [c[0] for c in cols]
    raw = sum(v.numel() * v.element_size() for v in values)
    log(f"batched: {len(values)} columns, {raw / 1e9:.2f} GB raw; autotune ({tuner})")
    t0 = time.perf_counter()
    if tuner == "shape":
        bss, tiles, _ = autotune_columns(values)
    else:
        bss, _ = autotune_block_sizes(values)
        tiles = [0] * len(bss)
    t_tune = time.perf_counter() - t0
    batch = Batch(values, bss, tile_positions=tiles)
    log("batched: encode")
    batch.encode(graph=True)
    torch.cuda.synchronize()
    batch.decode(graph=True)
    torch.cuda.synchronize()
    ok = all(torch.equal(o, v) for o, v in zip(batch.outs, values))
    comp = batch.compressed_bytes()
    t_benc = _timed(lambda: batch.encode(graph=True), reps)
    t_bdec = _timed(lambda: batch.decode(graph=True), reps)
    # the same launch as a persistent grid (8 workgroups per CU striding over the tiles)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    batch.set_decode_grid(8 * cus)
    t_pers = _timed(lambda: batch.decode(graph=True), reps)
    batch.set_decode_grid(0)
    hist = {int(b): bss.count(b) for b in sorted(set(bss))}
    shapes = {}
    for b, t in zip(bss, tiles):
        k = "staged" if b >= 2048 else f"tile{t or 8192}"
        shapes[k] = shapes.get(k, 0) + 1
    out = {
        "columns": len(values), "raw_bytes": raw, "compressed_bytes": comp, "compression_ratio": round(comp / raw, 5),
        "tuner": tuner, "block_sizes": hist, "decode_shapes": shapes, "autotune_s": round(t_tune, 3),
        "batched_decode_gbs": round(raw / t_bdec / 1e9, 1),
        "batched_decode_algorithmic_gbs": round((raw + comp) / t_bdec / 1e9, 1),
        "batched_decode_hbm_frac": round((raw + comp) / t_bdec / 1e9 / 8000.0, 4),
        "persistent_decode_gbs": round(raw / t_pers / 1e9, 1),
        "batched_encode_gbs": round(raw / t_benc / 1e9, 1),
        "batched_decode_ms": round(t_bdec * 1e3, 3),
    }
    if plain:
        # plain path: one decode launch / one column encode per column, from a host loop
        for o in batch.outs:
            o.zero_()
        t_pdec = _timed(batch.decode_per_column, reps)
        ok = ok and all(torch.equal(o, v) for o, v in zip(batch.outs, values))
        t_penc = _timed(batch.encode_per_column, max(1, reps // 2))
        out.update({"plain_decode_gbs": round(raw / t_pdec / 1e9, 1), "decode_speedup": round(t_pdec / t_bdec, 3),
                    "plain_encode_gbs": round(raw / t_penc / 1e9, 1), "encode_speedup": round(t_penc / t_benc, 3),
                    "plain_decode_ms": round(t_pdec * 1e3, 3)})
    out["verified_bit_exact"] = ok
    batch.close()
    return out


def _latency_stats(res, reqs, kinds=(srvmod.REQ_LOOKUP, srvmod.REQ_RANGE), split: bool = True) -> dict:
    """p50 / p99 latency of the completed requests of `kinds`; with split,
    also of the point-lookup (random-access) and range requests apart."""
    sel = np.isin(reqs["kind"], kinds) & (res["status"] == 0)
    lat = res["latency_s"][sel]
    out = {"p50_us": None, "p99_us": None}
    if lat.size:
        out = {"p50_us": round(float(np.percentile(lat, 50)) * 1e6, 1),
               "p99_us": round(float(np.percentile(lat, 99)) * 1e6, 1)}
    if split:
        for name, kind in (("lookup", srvmod.REQ_LOOKUP), ("range", srvmod.REQ_RANGE)):
            part = _latency_stats(res, reqs, (kind,), split=False)
            out[f"{name}_p50_us"], out[f"{name}_p99_us"] = part["p50_us"], part["p99_us"]
    return out


def _verify_sample(rng, reqs, per_kind: int):
    """A seeded sample of per_kind lookups and per_kind ranges of a trace."""
    sel = np.zeros(len(reqs), bool)
    for kind in (srvmod.REQ_LOOKUP, srvmod.REQ_RANGE):
        ids = np.nonzero(reqs["kind"] == kind)[0]
        sel[rng.choice(ids, min(per_kind, len(ids)), replace=False)] = True
    return sel


def run_serving(col, values=None, rates=(1e5, 1e6, 1e7), nreq: int = 60_000, plain_nreq: int = 6_000,
                seed: int = 4, reserve_bytes: int = 64 << 30, verify_per_kind: int = 200, log=print) -> dict:
    """Poisson traffic on one resident column: 80% lookups of U[1, 4096]
    indices, 20% ranges of log-uniform length in [1e2, 1e7].  One untimed
    run of 2000 requests first, so that the measured runs see a warm server
    (the first launch of a kernel loads its code object, which otherwise
    stalls the first run's queue for milliseconds).  Every run checks that
    all requests completed with status 0; with `values` (the raw column in
    HBM), a seeded sample of verify_per_kind lookups and ranges of each
    measured batched run writes its results to HBM (d_out), and they are
    compared with the column on the device."""
    srv = srvmod.Server(slots=4, reserve_bytes=reserve_bytes)
    srv.attach(0, col)
    specs = [(0, col.n, col.width)]
    out = {"rates": []}
    warm, warm_idx, _ = srvmod.make_trace(np.random.default_rng(seed + 1), specs, rates[0], 2000)
    srv.run(warm, srvmod.SERVE_BATCHED)  # the requests point into warm_idx, alive until here
    srv.run(warm[:200], srvmod.SERVE_PLAIN)
    del warm, warm_idx
    verified, checked, statuses_ok = True, 0, True
    for rate in rates:
        rng = np.random.default_rng(seed)
        reqs, idx, _ = srvmod.make_trace(rng, specs, rate, nreq)
        sel = buf = None
        if values is not None:
            sel = _verify_sample(np.random.default_rng(seed + 7), reqs, verify_per_kind)
            buf, starts = srvmod.device_outputs(reqs, sel, col.width)
        log(f"serving: batched at {rate:.0e} req/s, {nreq} requests")
        res, st = srv.run(reqs, srvmod.SERVE_BATCHED)
        slow = srv.slowest_iterations()[:3]
        statuses_ok = statuses_ok and bool((res["status"] == 0).all())
        if sel is not None:
            bad = srvmod.check_device_outputs(reqs, idx, sel, buf, starts, {0: values})
            verified = verified and bad == 0
            checked += int(sel.sum())
            del buf
        row = {"offered_req_s": rate, "requests": nreq, "completed": st["completed"],
               "throughput_req_s": round(st["completed"] / st["wall_s"], 1),
               "gbs": round(st["bytes_out"] / st["wall_s"] / 1e9, 1), "batches": st["batches"],
               "max_batch_requests": st["max_batch_requests"], **_latency_stats(res, reqs),
               "slowest_loop_iterations": slow}
        out["rates"].append(row)
        del idx
    # the same kind of trace, one launch + sync per request
    rng = np.random.default_rng(seed)
    rate = rates[0]
    reqs, idx, _ = srvmod.make_trace(rng, specs, rate, plain_nreq)
    log(f"serving: plain per-request path at {rate:.0e} req/s, {plain_nreq} requests")
    res, st = srv.run(reqs, srvmod.SERVE_PLAIN)
    statuses_ok = statuses_ok and bool((res["status"] == 0).all())
    plain = {"offered_req_s": rate, "requests": plain_nreq,
             "throughput_req_s": round(st["completed"] / st["wall_s"], 1),
             "gbs": round(st["bytes_out"] / st["wall_s"] / 1e9, 1), **_latency_stats(res, reqs)}
    res2, st2 = srv.run(reqs, srvmod.SERVE_BATCHED)
    statuses_ok = statuses_ok and bool((res2["status"] == 0).all())
    same = {"throughput_req_s": round(st2["completed"] / st2["wall_s"], 1),
            "gbs": round(st2["bytes_out"] / st2["wall_s"] / 1e9, 1), **_latency_stats(res2, reqs)}
    out["plain_same_trace"] = plain
    out["batched_same_trace"] = same
    out["gbs_ratio_same_trace"] = round(same["gbs"] / plain["gbs"], 2) if plain["gbs"] else None
    # saturation: every request arrives at t = 0, so each path runs at its maximum rate
    sat = reqs.copy()
    sat["arrival_s"] = 0.0
    log(f"serving: saturation, {plain_nreq} requests at t=0, plain then batched")
    res_p, st_p = srv.run(sat, srvmod.SERVE_PLAIN)
    res_b, st_b = srv.run(sat, srvmod.SERVE_BATCHED)
    statuses_ok = statuses_ok and bool((res_p["status"] == 0).all()) and bool((res_b["status"] == 0).all())
    out["saturation"] = {
        "requests": plain_nreq,
        "plain_req_s": round(st_p["completed"] / st_p["wall_s"], 1),
        "plain_gbs": round(st_p["bytes_out"] / st_p["wall_s"] / 1e9, 1),
        "batched_req_s": round(st_b["completed"] / st_b["wall_s"], 1),
        "batched_gbs": round(st_b["bytes_out"] / st_b["wall_s"] / 1e9, 1),
        "gbs_ratio": round((st_b["bytes_out"] / st_b["wall_s"]) / (st_p["bytes_out"] / st_p["wall_s"]), 2),
    }
    out["all_status_ok"] = statuses_ok
    out["verified"] = verified if values is not None else None
    out["verified_requests"] = checked
    srv.close()
    return out


def run_churn(ncols: int = 2000, sources: int = 16, resident: int = 48, insert_rate: float = 2000.0,
              req_rate: float = 1e5, lo: float = 1e4, hi: float = 1e8, seed: int = 5,
              reserve_bytes: int = 64 << 30, verify_per_kind: int = 200, log=print) -> dict:
    """Columns inserted (GPU encode) and evicted while lookup/range traffic
    runs.  Each of the `ncols` inserts draws its own length, log-uniform in
    [lo, hi]; its values are a window of that length, at a random offset, of
    one of `sources` uint64 drift columns of `hi` values (a*i + b*i^2 +
    N(0, 2^10), a ~ U[1, 2^16], b ~ logU[1e-6, 1e-3], each source its own
    a, b and noise) generated in HBM before the run: a bounded pool of
    sources * hi * 8 bytes.  A window of a drift column is a drift column (the
    same a and b, shifted).  At most `resident` columns live; traffic (80%
    lookups of 1-4096 indices, 20% ranges of log-uniform length in
    [1e2, 1e7]) goes to live columns; a seeded sample of verify_per_kind
    lookups and ranges writes to HBM and is compared with its window."""
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(rng, ncols, lo, hi)
    src_of = np.arange(ncols) % sources
    offs = (rng.random(ncols) * (int(hi) - lens + 1)).astype(np.int64)
    log(f"churn: {sources} pool sources of {int(hi)} values ({sources * int(hi) * 8 / 1e9:.1f} GB), "
        f"{ncols} inserts totalling {lens.sum() * 8 / 1e9:.1f} GB raw; block sizes by autotune")
    pool = [gen_column("drift", int(hi), 64, seed * 1000 + s) for s in range(sources)]
    bss, _ = autotune_block_sizes(pool, candidates=(1024, 2048, 4096, 8192), window=8192)
    R = srvmod.REQUEST_DTYPE
    rows = []
    t = 0.0
    dt_ins = 1.0 / insert_rate
    n_req_per_insert = max(1, int(req_rate / insert_rate))
    idx_bufs = []
    live = []
    for i in range(ncols):
        s, L = int(src_of[i]), int(lens[i])
        r = np.zeros(1, R)[0]
        r["kind"], r["column"], r["count"] = srvmod.REQ_INSERT, i, L
        r["d_values"] = pool[s].data_ptr() + 8 * int(offs[i])
        r["block_size"], r["width_bits"], r["arrival_s"] = bss[s], 64, t
        rows.append(r)
        live.append((i, L))
        if len(live) > resident:
            ev, _ = live.pop(0)
            r = np.zeros(1, R)[0]
            r["kind"], r["column"], r["arrival_s"] = srvmod.REQ_EVICT, ev, t
            rows.append(r)
        for k in range(n_req_per_insert):
            c, n = live[int(rng.integers(0, len(live)))]
            r = np.zeros(1, R)[0]
            r["column"], r["arrival_s"] = c, t + dt_ins * (k + 1) / (n_req_per_insert + 1)
            if rng.random() < 0.8:
                cnt = int(rng.integers(1, 4097))
                ii = (rng.random(cnt) * n).astype(np.uint64)
                idx_bufs.append(ii)
                r["kind"], r["count"], r["h_idx"] = srvmod.REQ_LOOKUP, cnt, ii.ctypes.data
            else:
                cnt = min(n, int(math.exp(rng.uniform(math.log(1e2), math.log(1e7)))))
                r["kind"], r["count"] = srvmod.REQ_RANGE, cnt
                r["begin"] = int(rng.integers(0, n - cnt + 1))
            rows.append(r)
        t += dt_ins
    reqs = np.array(rows, R)
    # the sampled requests' lookup indices, gathered into one buffer for the check
    sel = _verify_sample(np.random.default_rng(seed + 3), reqs, verify_per_kind) if verify_per_kind else None
    buf = starts = None
    if sel is not None:
        buf, starts = srvmod.device_outputs(reqs, sel, 64)
    srv = srvmod.Server(slots=4, reserve_bytes=reserve_bytes)
    log(f"churn: {len(reqs)} requests ({ncols} inserts) over {t:.2f} s")
    res, st = srv.run(reqs, srvmod.SERVE_BATCHED)
    srv.close()
    verified = None
    if sel is not None:
        cols = {int(c): None for c in reqs["column"][sel]}
        values = {c: pool[int(src_of[c])][int(offs[c]):int(offs[c]) + int(lens[c])] for c in cols}
        allidx = np.concatenate(idx_bufs) if idx_bufs else np.zeros(0, np.uint64)
        # check_device_outputs reads lookup indices at h_idx relative to one buffer: rebase them
        chk = reqs.copy()
        at = 0
        for i in np.nonzero(chk["kind"] == srvmod.REQ_LOOKUP)[0]:
            chk["h_idx"][i] = allidx.ctypes.data + 8 * at
            at += int(chk["count"][i])
        bad = srvmod.check_device_outputs(chk, allidx, sel, buf, starts, values)
        verified = bad == 0 and bool((res["status"][sel] == 0).all())
        del buf
    ins = reqs["kind"] == srvmod.REQ_INSERT
    ok_ins = int((res["status"][ins] == 0).sum())
    traffic = _latency_stats(res, reqs)
    edges = [10 ** k for k in range(int(math.log10(lo)), int(math.ceil(math.log10(hi))) + 1)]
    hist = {f"[{edges[k]:.0e},{edges[k + 1]:.0e})": int(((lens >= edges[k]) & (lens < edges[k + 1])).sum())
            for k in range(len(edges) - 1)}
    hist[f"[{edges[-2]:.0e},{edges[-1]:.0e})"] += int((lens >= edges[-1]).sum())
    del pool
    return {
        "inserts": ncols, "inserts_ok": ok_ins, "requests": int((~ins & (reqs["kind"] <= 1)).sum()),
        "failed": st["failed"], "wall_s": round(st["wall_s"], 3),
        "insert_lengths": {"distribution": f"log-uniform in [{lo:.0e}, {hi:.0e}], one draw per insert",
                           "histogram": hist, "min": int(lens.min()), "max": int(lens.max()),
                           "pool_sources": sources, "pool_gb": round(sources * int(hi) * 8 / 1e9, 2)},
        "realized_raw_gb": round(float(lens.sum()) * 8 / 1e9, 2),
        "encode_gbs": round(st["encoded_bytes"] / st["encode_gpu_s"] / 1e9, 1) if st["encode_gpu_s"] else None,
        "encoded_gb": round(st["encoded_bytes"] / 1e9, 2),
        "decode_gbs": round(st["bytes_out"] / st["wall_s"] / 1e9, 1),
        "insert_p99_ms": round(float(np.percentile(res["latency_s"][ins], 99)) * 1e3, 2),
        **traffic, "peak_column_pool_gb": round(st["peak_column_bytes"] / 1e9, 2),
        "verified": verified, "verified_requests": int(sel.sum()) if sel is not None else 0,
    }
