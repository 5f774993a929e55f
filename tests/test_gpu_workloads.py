"""GPU: the bench's batched / served / churn drivers at small scale (their
results must come back verified and complete)."""
import sys

import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def test_batched_driver_small():
    from leco_fix import workloads
    cols = workloads.batched_columns(24, 1e3, 2e5, seed=7)
    out = workloads.run_batched(cols, reps=2, log=say)
    assert out["verified_bit_exact"] and out["columns"] == 24
    assert all(256 * 3 // 4 <= b <= 16384 * 3 // 2 for b in out["block_sizes"])
    assert out["batched_decode_gbs"] > 0 and out["plain_decode_gbs"] > 0


def test_generators_match_their_specs():
    import torch
    from leco_fix import workloads
    for dist in workloads.DISTS + ("drift",):
        for width in (32, 64):
            v = workloads.gen_column(dist, 100_000, width, 3)
            assert v.numel() == 100_000 and v.dtype == (torch.int32 if width == 32 else torch.int64)
            if dist.startswith("sorted") or dist == "stepwise":
                u = v.to(torch.int64) & 0xFFFFFFFF if width == 32 else v
                assert bool((u[1:] >= u[:-1]).all()), dist


def test_serving_and_churn_drivers_small():
    import torch
    from leco_fix import Column, workloads
    vals = workloads.gen_column("sorted_uniform", 3_000_000, 64, 9)
    col = Column.encode(vals, 4096)
    out = workloads.run_serving(col, vals, rates=(1e5, 1e6), nreq=3000, plain_nreq=400, reserve_bytes=8 << 30,
                                log=say)
    for row in out["rates"]:
        assert row["completed"] == 3000 and row["p99_us"] is not None
    assert out["all_status_ok"] and out["verified"] and out["verified_requests"] == 2 * 400
    assert out["plain_same_trace"]["throughput_req_s"] > 0
    ch = workloads.run_churn(ncols=40, sources=4, resident=6, insert_rate=2000, req_rate=5e4, lo=1e4, hi=2e6,
                             reserve_bytes=8 << 30, log=say)
    assert ch["inserts_ok"] == 40 and ch["failed"] == 0 and ch["encode_gbs"] > 0
    assert ch["verified"] and ch["verified_requests"] > 0
    assert sum(ch["insert_lengths"]["histogram"].values()) == 40
    assert abs(ch["encoded_gb"] - ch["realized_raw_gb"]) <= 0.01
