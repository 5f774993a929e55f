"""GPU: the FOR_int / Delta_int baselines against the golden fixtures written
by tests/spec_for_delta.py and against the oracle (streams byte for byte,
decode, lookups, FOR filters), the host one-block surface, and the 500M-key
column through round trips."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "base_cases.json")) as _f:
    BASE = json.load(_f)


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def to_dev(torch, a):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


@pytest.mark.parametrize("case", [c["name"] for c in BASE])
def test_base_golden_on_gpu(case):
    import torch
    from leco_fix.baselines import BaseColumn
    with np.load(os.path.join(GOLDEN, f"{case}.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    codec = case.split("_")[1]
    vals, bs = g["values"], int(g["bs"])
    say(f"golden {case}")
    col = BaseColumn.encode(codec, to_dev(torch, vals), bs)
    assert np.array_equal(host(col.offsets, np.uint64), g["offsets"])
    assert col.stream_bytes() == g["stream"].tobytes()
    assert np.array_equal(host(col.decode(), vals.dtype), vals)
    idx = np.arange(len(vals), dtype=np.int64)
    assert np.array_equal(host(col.lookup(torch.from_numpy(idx).cuda()), vals.dtype), vals)


def mixed(rng, n, width):
    from leco_fix import datagen
    top = (1 << width) - 1
    k = n // 4
    parts = [datagen.sorted_uniform(k, width, int(rng.integers(1 << 20))).astype(np.uint64),
             np.full(k, int(rng.integers(0, top, dtype=np.uint64)), np.uint64),
             rng.integers(0, top, k, dtype=np.uint64, endpoint=True),
             rng.integers(0, 1 << 20, n - 3 * k, dtype=np.uint64)]
    return np.concatenate(parts).astype(np.uint32 if width == 32 else np.uint64)


@pytest.mark.parametrize("codec", ["for", "delta"])
@pytest.mark.parametrize("bs", [1, 5, 256, 1000, 4096, 10000])
def test_base_columns_match_oracle(oracle_mod, codec, bs):
    import torch
    from leco_fix.baselines import BaseColumn
    rng = np.random.default_rng(bs + (7 if codec == "for" else 13))
    for width in (32, 64):
        n = min(80_000, max(40, 9 * bs))
        vals = mixed(rng, n, width)
        say(f"{codec} width={width} bs={bs} n={n}")
        st, off = oracle_mod.base_encode_column(codec, vals, bs)
        col = BaseColumn.encode(codec, to_dev(torch, vals), bs)
        assert np.array_equal(host(col.offsets, np.uint64), off)
        assert col.stream_bytes() == st.tobytes()
        assert np.array_equal(host(col.decode(), vals.dtype), vals)
        idx = rng.integers(0, n, 20_000, dtype=np.uint64)
        got = host(col.lookup(torch.from_numpy(idx.view(np.int64)).cuda()), vals.dtype)
        assert np.array_equal(got, oracle_mod.base_lookup_column(codec, st, off, n, bs, idx, vals.dtype))
        assert np.array_equal(got, vals[idx])
        if codec == "for":
            v = vals.astype(np.uint64)
            for f in (0, int(np.median(v)), (1 << width) - 1):
                want = oracle_mod.for_filter_column(st, off, n, bs, vals.dtype, filt=f)
                assert np.array_equal(host(col.filter_range(f), np.uint32), want), (width, f)
            for base, f1, f2 in ((1000, 100, 700), ((1 << 31) - 1, 5, 1 << 30)):
                want = oracle_mod.for_filter_column(st, off, n, bs, vals.dtype, close=True, f1=f1, f2=f2, base=base)
                assert np.array_equal(host(col.filter_range_close(f1, f2, base), np.uint32), want), (width, base)


@pytest.mark.parametrize("codec", ["for", "delta"])
@pytest.mark.parametrize("width,bs,n", [(64, 100, 1_000_003), (32, 256, 2_000_000), (64, 1000, 1_500_001),
                                        (32, 4096, 3_000_017), (64, 7, 300_001), (32, 2500, 999_999)])
def test_base_encode_many_units(oracle_mod, codec, width, bs, n):
    # blocks of up to 4096 values are encoded in runs of blocks per workgroup, one
    # look-back per run: many runs, a ragged last run and a ragged last block
    import torch
    from leco_fix.baselines import BaseColumn
    rng = np.random.default_rng(bs * 7 + width)
    vals = mixed(rng, n, width)
    say(f"{codec} width={width} bs={bs} n={n} (many encode units)")
    st, off = oracle_mod.base_encode_column(codec, vals, bs)
    col = BaseColumn.encode(codec, to_dev(torch, vals), bs)
    assert np.array_equal(host(col.offsets, np.uint64), off)
    assert col.stream_bytes() == st.tobytes()
    assert np.array_equal(host(col.decode(), vals.dtype), vals)


def test_base_host_surface(oracle_mod):
    from leco_fix.baselines import Delta_int, FOR_int
    rng = np.random.default_rng(3)
    for cls, codec in ((FOR_int, "for"), (Delta_int, "delta")):
        for dt in (np.uint32, np.uint64):
            c = cls(dt)
            c.init(4, 500)
            width = 8 * np.dtype(dt).itemsize
            for y in (np.sort(rng.integers(0, 1 << (width - 3), 500, dtype=np.uint64)).astype(dt),
                      rng.integers(0, (1 << width) - 1, 500, dtype=np.uint64).astype(dt),
                      np.array([0, 1 << (width - 2), 5, (1 << (width - 2)) + 7] * 125, dtype=np.uint64).astype(dt)):
                say(f"host {codec} width={width} l")
                blk = c.encodeArray8_int(y, len(y))
                assert blk == oracle_mod.base_encode_block(codec, y)
                assert np.array_equal(c.decodeArray8(blk, len(y)), y)
                for j in (0, 1, 250, 499):
                    assert c.randomdecodeArray8(blk, j) == int(y[j])
                if codec == "for":
                    f = int(y[200])
                    want = [2 * 500 + i for i in range(500) if int(y[i]) > f]
                    assert c.filter_range(blk, 500, f, nvalue=2).tolist() == want
                    got = c.filter_range_close(blk, 500, nvalue=1, filter1=10, filter2=600, base_val=1000)
                    assert got.tolist() == [500 + i for i in range(500) if 10 < int(y[i]) % 1000 < 600]


@pytest.mark.parametrize("bs", [8193, 30000])
def test_for_host_filter_large_block(bs):
    # one FOR block split into several filter tiles on the GPU; brute force is the reference
    from leco_fix.baselines import FOR_int
    rng = np.random.default_rng(bs)
    for dt in (np.uint32, np.uint64):
        width = 8 * np.dtype(dt).itemsize
        c = FOR_int(dt)
        c.init(1, bs)
        y = rng.integers(0, 1 << (width - 6), bs, dtype=np.uint64).astype(dt)
        blk = c.encodeArray8_int(y, bs)
        say(f"host for filter large block width={width} bs={bs}")
        for f in (0, int(y[bs // 2]), int(y.max())):
            want = [3 * bs + i for i in range(bs) if int(y[i]) > f]
            assert c.filter_range(blk, bs, f, nvalue=3).tolist() == want


def test_base_cpp_driver_flow():
    # the reference driver's FOR / Delta flow compiled against include/leco_baselines.hpp
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "base_driver")
    assert os.path.exists(exe), "built by learn-to-compress_amd/build.py"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


@pytest.mark.parametrize("codec", ["for", "delta"])
def test_base_500m_round_trip(codec):
    import torch
    from leco_fix import datagen
    from leco_fix.baselines import BaseColumn
    n = 500_000_000
    say(f"{codec} 500M sorted_pwl_u64 bs=4096")
    vals = datagen.sorted_pwl_u64(n, seed=2, device="cuda")
    col = BaseColumn.encode(codec, vals, 4096)
    out = col.decode()
    assert torch.equal(out, vals)
    del out
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    idx = torch.randint(0, n, (1_000_000,), device="cuda", generator=g)
    assert torch.equal(col.lookup(idx), vals[idx])
