"""GPU: the request server.  Every served result must equal the column's
values (ground truth), in the batched mix and when the same request runs
alone on the plain per-request path; columns inserted while traffic runs
must encode byte-identically to the oracle; memory admission must hold
requests back, not fail them, when the result pool is small."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GB = 1 << 30


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def to_dev(torch, a):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


def expected(req, values, idx):
    v = values[int(req["column"])]
    if req["kind"] == 0:
        start = (int(req["h_idx"]) - idx.ctypes.data) // 8
        return v[idx[start:start + int(req["count"])]]
    b = int(req["begin"])
    return v[b:b + int(req["count"])]


@pytest.fixture(scope="module")
def served_columns():
    import torch
    from leco_fix import Column, datagen
    a = datagen.sorted_pwl_u64(2_000_000, seed=21)
    b = datagen.linear_noise_u32(1_000_000, seed=22)
    cols = {0: Column.encode(to_dev(torch, a), 4096), 1: Column.encode(to_dev(torch, b), 1000)}
    return cols, {0: a, 1: b}


def test_served_results_equal_alone_and_truth(served_columns):
    from leco_fix.server import SERVE_BATCHED, SERVE_PLAIN, Server, make_trace
    cols, values = served_columns
    srv = Server(slots=4, reserve_bytes=0, result_pool_bytes=1 * GB, column_pool_bytes=64 << 20)
    for s, c in cols.items():
        srv.attach(s, c)
    rng = np.random.default_rng(3)
    specs = [(0, len(values[0]), 64), (1, len(values[1]), 32)]
    reqs, idx, outs = make_trace(rng, specs, rate=2e5, nreq=3000, range_hi=2e5, keep_outputs=True)
    reqs["arrival_s"][:800] = 0.0  # a burst, so batches hold many requests
    say("batched run")
    res, stats = srv.run(reqs)
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    assert stats["max_batch_requests"] > 1 and stats["batches"] < len(reqs)
    mixed = [o.copy() for o in outs]
    for i in range(len(reqs)):
        assert np.array_equal(mixed[i], expected(reqs[i], values, idx)), i
    say("plain run of the same requests, one at a time")
    for o in outs:
        o[:] = 0
    res2, stats2 = srv.run(reqs, SERVE_PLAIN)
    assert stats2["completed"] == len(reqs) and stats2["batches"] == len(reqs)
    for i in range(len(reqs)):
        assert np.array_equal(outs[i], mixed[i]), i
    assert (res["status"] == 0).all() and (res2["status"] == 0).all()
    srv.close()


def test_lookup_batch_with_split_index_copies(served_columns):
    """One burst of lookups with odd counts: the batch's index bytes (well
    above the size where the host copy is split over threads) are cut into
    parts inside requests and inside 8-byte indices."""
    from leco_fix.server import REQUEST_DTYPE, Server
    cols, values = served_columns
    srv = Server(slots=2, reserve_bytes=0, result_pool_bytes=1 * GB, column_pool_bytes=64 << 20)
    srv.attach(0, cols[0])
    srv.attach(1, cols[1])
    rng = np.random.default_rng(17)
    counts = rng.integers(1, 20_001, 96) | 1
    col = np.arange(96) % 2
    idx = [rng.integers(0, len(values[c]), int(k)).astype(np.uint64) for c, k in zip(col, counts)]
    outs = [np.zeros(int(k), np.uint64 if c == 0 else np.uint32) for c, k in zip(col, counts)]
    reqs = np.zeros(96, REQUEST_DTYPE)
    reqs["kind"], reqs["column"], reqs["count"] = 0, col, counts
    for i in range(96):
        reqs["h_idx"][i], reqs["h_out"][i] = idx[i].ctypes.data, outs[i].ctypes.data
    say(f"burst of 96 lookups, {int(counts.sum()) * 8 >> 10} KiB of indices")
    res, stats = srv.run(reqs)
    assert (res["status"] == 0).all() and stats["max_batch_requests"] > 1
    for i in range(96):
        assert np.array_equal(outs[i], values[int(col[i])][idx[i]]), i
    srv.close()


def test_admission_holds_requests_when_results_pool_is_small(served_columns):
    from leco_fix.server import REQUEST_DTYPE, Server
    cols, values = served_columns
    pool = 4 << 20  # 4 MiB of results: at most a few ranges in flight
    srv = Server(slots=4, reserve_bytes=0, result_pool_bytes=pool, column_pool_bytes=64 << 20)
    srv.attach(0, cols[0])
    n = len(values[0])
    reqs = np.zeros(41, REQUEST_DTYPE)
    reqs["kind"] = 1
    reqs["column"] = 0
    reqs["count"] = 300_000          # 2.4 MB each: one fits in the pool at a time
    reqs["begin"] = np.linspace(0, n - 300_000, 41).astype(np.uint64)
    reqs["count"][-1] = 600_000      # 4.8 MB: larger than the whole pool
    reqs["begin"][-1] = 0
    outs = [np.zeros(int(c), np.uint64) for c in reqs["count"]]
    for i, o in enumerate(outs):
        reqs["h_out"][i] = o.ctypes.data
    res, stats = srv.run(reqs)
    assert (res["status"][:-1] == 0).all() and res["status"][-1] == 3  # ENOMEM for the oversized one
    assert stats["peak_result_bytes"] <= pool
    for i in range(40):
        b = int(reqs["begin"][i])
        assert np.array_equal(outs[i], values[0][b:b + 300_000]), i
    srv.close()


def test_churn_insert_evict_while_serving(oracle_mod):
    import torch
    from leco_fix import datagen
    from leco_fix.batch import _view
    from leco_fix.server import REQ_EVICT, REQ_INSERT, REQ_LOOKUP, REQ_RANGE, REQUEST_DTYPE, Server
    srv = Server(slots=3, reserve_bytes=0, result_pool_bytes=256 << 20, column_pool_bytes=1 * GB)
    rng = np.random.default_rng(9)
    data = {s: datagen.drift_u64(int(rng.integers(20_000, 400_000)), seed=100 + s) for s in range(4)}
    dev = {s: to_dev(torch, v) for s, v in data.items()}
    bss = {0: 1024, 1: 2048, 2: 4096, 3: 8192}
    rows, outs, t = [], [], 0.0
    idx_keep = []

    def add(kind, col, count=0, begin=0, **kw):
        nonlocal t
        r = np.zeros(1, REQUEST_DTYPE)[0]
        r["kind"], r["column"], r["count"], r["begin"], r["arrival_s"] = kind, col, count, begin, t
        for k, v in kw.items():
            r[k] = v
        rows.append(r)
        t += 2e-5

    # insert 0 and 1, traffic on both; evict 0 and insert 2; traffic on 1 and 2
    for c in (0, 1):
        add(REQ_INSERT, c, len(data[c]), d_values=dev[c].data_ptr(), block_size=bss[c], width_bits=64)
    for k in range(60):
        c = k % 2 if k <= 30 else 1 + k % 2
        n = len(data[c])
        if k % 3:
            ii = rng.integers(0, n, 700, dtype=np.uint64)
            idx_keep.append(ii)
            o = np.zeros(700, np.uint64)
            outs.append((len(rows), o, data[c][ii]))
            add(REQ_LOOKUP, c, 700, h_idx=ii.ctypes.data, h_out=o.ctypes.data)
        else:
            b = int(rng.integers(0, n - 5000))
            o = np.zeros(5000, np.uint64)
            outs.append((len(rows), o, data[c][b:b + 5000]))
            add(REQ_RANGE, c, 5000, b, h_out=o.ctypes.data)
        if k == 30:
            add(REQ_EVICT, 0)
            add(REQ_INSERT, 2, len(data[2]), d_values=dev[2].data_ptr(), block_size=bss[2], width_bits=64)
    late = len(rows)
    add(REQ_LOOKUP, 0, 10, h_idx=idx_keep[0].ctypes.data)   # slot 0 is gone by now: must fail
    reqs = np.array(rows, REQUEST_DTYPE)
    say(f"churn trace of {len(reqs)} requests")
    res, stats = srv.run(reqs)
    for i, o, want in outs:
        assert res["status"][i] == 0, (i, res["status"][i])
        assert np.array_equal(o, want), i
    assert res["status"][late] != 0
    assert (res["status"][np.array(reqs["kind"]) >= REQ_INSERT] == 0).all()
    assert stats["encoded_values"] == sum(len(data[s]) for s in (0, 1, 2))
    # the inserted columns are byte-identical to the oracle
    for s in (1, 2):
        info = srv.column_info(s)
        got = _view(info["stream"], info["stream_bytes"], "|u1").cpu().numpy().tobytes()
        want, _ = oracle_mod.encode_column(data[s], bss[s])
        assert got == want.tobytes(), s
    srv.close()


def test_config4_shape_served_results_equal_the_column():
    """The serving workload at its full shape: the resident 500M-key column
    (sorted piecewise-linear uint64, block size 4096), Poisson arrivals at
    10^6 req/s, 80% lookups of 1-4096 random indices and 20% ranges of
    log-uniform length in [10^2, 10^7].  Every result is written to HBM by
    the server (d_out) and compared with the input column on the device; a
    subset of requests, run again alone on the plain path, gives the same
    values as in the mix."""
    import torch
    from leco_fix import Column, datagen
    from leco_fix.server import (REQ_LOOKUP, REQ_RANGE, SERVE_PLAIN, Server, check_device_outputs,
                                 device_outputs, make_trace)
    n = 500_000_000
    say("config-4 shape: generate and encode the 500M-key column")
    vals = datagen.sorted_pwl_u64(n, seed=2, device="cuda")
    col = Column.encode(vals, 4096)
    srv = Server(slots=4, reserve_bytes=0, result_pool_bytes=1 * GB, column_pool_bytes=64 << 20)
    srv.attach(0, col)
    rng = np.random.default_rng(44)
    reqs, idx, _ = make_trace(rng, [(0, n, 64)], rate=1e6, nreq=3000, max_lookup=4096, range_hi=1e7)
    kinds = reqs["kind"]
    assert (kinds == REQ_RANGE).sum() > 400 and reqs["count"][kinds == REQ_RANGE].max() > 1_000_000
    everything = np.ones(len(reqs), bool)
    buf, starts = device_outputs(reqs, everything, 64)
    say(f"config-4 shape: batched run of {len(reqs)} requests, {int(reqs['count'].sum())} values out")
    res, stats = srv.run(reqs)
    assert (res["status"] == 0).all() and stats["completed"] == len(reqs)
    assert stats["max_batch_requests"] > 1
    bad = check_device_outputs(reqs, idx, everything, buf, starts, {0: vals})
    assert bad == 0
    # alone = in the mix: 24 lookups and 24 ranges (the longest included), one request at a time
    lk = np.nonzero(kinds == REQ_LOOKUP)[0][:24]
    rg = np.nonzero(kinds == REQ_RANGE)[0]
    rg = np.unique(np.concatenate([rg[:23], [rg[np.argmax(reqs['count'][rg])]]]))
    sub = np.sort(np.concatenate([lk, rg]))
    alone = reqs[sub].copy()
    alone["arrival_s"] = np.arange(len(sub)) * 1e-4
    abuf, astarts = device_outputs(alone, np.ones(len(sub), bool), 64)
    say(f"config-4 shape: {len(sub)} requests alone on the plain path")
    res2, _ = srv.run(alone, SERVE_PLAIN)
    assert (res2["status"] == 0).all()
    for k, i in enumerate(sub):
        c = int(reqs["count"][i])
        assert torch.equal(abuf[int(astarts[k]):int(astarts[k]) + c], buf[int(starts[i]):int(starts[i]) + c]), i
    srv.close()
    del buf, abuf, col, vals
    torch.cuda.empty_cache()
