"""GPU: filter_range / filter_range_close on compressed columns (one launch
per column) against the oracle, on every kind of block; the host per-block
surface; capacity handling; and the 500M-key column through size-independent
properties."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def to_dev(torch, a):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


def positions(t):
    return t.cpu().numpy().view(np.uint32)


def mixed_values(rng, n, width, bs):
    """Sorted, constant, raw, unsorted and stepwise stretches in one column."""
    from leco_fix import datagen
    top = (1 << width) - 1
    k = n // 5
    parts = [datagen.sorted_uniform(k, width, int(rng.integers(1 << 20))).astype(np.uint64) >> np.uint64(2),
             np.full(k, int(rng.integers(1, top >> 3)), np.uint64),
             rng.integers(0, top, k, dtype=np.uint64, endpoint=True),
             rng.permutation(datagen.linear_sigma(k, width, int(rng.integers(1 << 20)))).astype(np.uint64),
             datagen.stepwise(n - 4 * k, width, int(rng.integers(1 << 20))).astype(np.uint64)]
    return np.concatenate(parts).astype(np.uint32 if width == 32 else np.uint64)


@pytest.mark.parametrize("bs", [1, 7, 256, 1000, 4096, 10000])
def test_filter_range_matches_oracle(oracle_mod, bs):
    import torch
    from leco_fix import Column
    rng = np.random.default_rng(bs)
    for width in (32, 64):
        n = min(60_000, max(50, 12 * bs))
        vals = mixed_values(rng, n, width, bs)
        say(f"filter_range width={width} bs={bs} n={n}")
        st, off = oracle_mod.encode_column(vals, bs, threads=8)
        col = Column.encode(to_dev(torch, vals), bs)
        assert col.stream_bytes() == st.tobytes()
        f64 = vals.astype(np.float64)
        filters = [0, 1, 2, int(np.quantile(f64, 0.3)), int(np.quantile(f64, 0.77)), int(vals.max()),
                   (1 << width) - 1]
        for f in filters:
            want = oracle_mod.filter_column(st, off, n, bs, vals.dtype, f)
            got = positions(col.filter_range(f))
            assert np.array_equal(got, want), (width, bs, f)


@pytest.mark.parametrize("bs", [7, 256, 1000, 4096])
def test_filter_range_close_matches_oracle(oracle_mod, bs):
    # per-block starting filters: the period at or below each block's minimum
    import torch
    from leco_fix import Column
    from leco_fix._lib import LecoError
    rng = np.random.default_rng(100 + bs)
    for width in (32, 64):
        n = min(60_000, max(50, 12 * bs))
        vals = mixed_values(rng, n, width, bs)
        st, off = oracle_mod.encode_column(vals, bs, threads=8)
        col = Column.encode(to_dev(torch, vals), bs)
        nb = col.num_blocks
        mins = [int(vals[b * bs:(b + 1) * bs].min()) for b in range(nb)]
        mask = (1 << width) - 1
        # (g2 - g1 > base: overlapping periods, which every block walks period by period)
        for base, g1, g2 in ((1 << 28, (1 << 28) // 6, (1 << 28) // 2), (12345, 2057, 6172),
                             ((1 << 31) - 1, 357913941, 1073741823), (40_000, 1000, 70_000)):
            say(f"filter_range_close width={width} bs={bs} n={n} base={base} window=({g1}, {g2})")
            o1 = np.array([(m - m % base + g1) & mask for m in mins], np.uint64)
            o2 = np.array([(m - m % base + g2) & mask for m in mins], np.uint64)
            t1, t2 = to_dev(torch, o1.astype(vals.dtype)), to_dev(torch, o2.astype(vals.dtype))
            try:
                want = oracle_mod.filter_close_column(st, off, n, bs, vals.dtype, o1, o2, base)
            except RuntimeError:
                with pytest.raises(LecoError):
                    col.filter_range_close(t1, t2, base)
                continue
            got = positions(col.filter_range_close(t1, t2, base))
            assert np.array_equal(got, want), (width, bs, base)


def test_filter_capacity_and_count(oracle_mod):
    import torch
    from leco_fix import Column
    from leco_fix._lib import lib
    vals = np.sort(np.random.default_rng(3).integers(0, 1 << 40, 100_000, dtype=np.uint64))
    col = Column.encode(to_dev(torch, vals), 1000)
    f = int(vals[30_000])
    want = np.nonzero(vals > np.uint64(f))[0].astype(np.uint32)
    say("filter capacity 1000 of ~70000")
    out = torch.zeros(1000, dtype=torch.int32, device="cuda")
    out2, result = col._filter(False, f, None, None, 0, out, 1000, None)
    torch.cuda.synchronize()
    assert int(result[0].item()) == len(want)
    assert np.array_equal(positions(out2), want[:1000])
    # zero capacity: count only
    _, result = col._filter(False, f, None, None, 0, torch.zeros(1, dtype=torch.int32, device="cuda"), 0, None)
    assert int(result[0].item()) == len(want)
    # a workspace that is too small is rejected
    res = torch.zeros(2, dtype=torch.int64, device="cuda")
    work = torch.zeros(8, dtype=torch.uint8, device="cuda")
    rc = lib().leco_fix_filter_range_u64(col.stream.data_ptr(), col.offsets.data_ptr(), col.n, 1000, f,
                                         out.data_ptr(), 1000, res.data_ptr(), work.data_ptr(), 8, None)
    assert rc == 1


def test_filter_close_period_limit_flag():
    import torch
    from leco_fix import Column
    from leco_fix._lib import LecoError
    # a steep block and a tiny period: more than 65536 periods in one block
    vals = (np.arange(4096, dtype=np.uint64) * np.uint64(1 << 30))
    col = Column.encode(to_dev(torch, vals), 4096)
    f1 = torch.tensor([1], dtype=torch.int64, device="cuda")
    f2 = torch.tensor([3], dtype=torch.int64, device="cuda")
    say("filter_range_close period limit")
    with pytest.raises(LecoError):
        col.filter_range_close(f1, f2, 5)


def test_filter_host_block_surface(oracle_mod):
    from leco_fix import Leco_int
    from leco_fix._lib import LecoError
    rng = np.random.default_rng(8)
    for width, dt in ((32, np.uint32), (64, np.uint64)):
        codec = Leco_int(dt)
        bs = 700
        codec.init(10, bs)
        for kind in range(4):
            if kind == 0:
                y = np.sort(rng.integers(0, 1 << (width - 2), bs, dtype=np.uint64)).astype(dt)
            elif kind == 1:
                y = rng.integers(0, (1 << width) - 1, bs, dtype=np.uint64).astype(dt)
            elif kind == 2:
                y = np.full(bs, 4242, dt)
            else:
                y = (np.arange(bs, dtype=np.uint64) * np.uint64(977) + np.uint64(5)).astype(dt)
            blk = codec.encodeArray8_int(y, bs)
            say(f"host filter width={width} kind={kind}")
            for f in (0, int(y[bs // 2]), int(y.max())):
                got = codec.filter_range(blk, bs, f, block_id=3)
                want = oracle_mod.filter_block(blk, bs, dt, f, 3 * bs)
                assert np.array_equal(got, want), (width, kind, f)
            span = int(y.max()) - int(y.min())
            wide = min(max(5000, span // 40 + 1), (1 << 31) - 1)
            for f1, f2, base in ((int(y.min()), int(y.min()) + 1000, wide), (100, 300, 4096)):
                try:
                    want = oracle_mod.filter_close_block(blk, bs, dt, f1, f2, base, 2 * bs)
                except RuntimeError:  # more than 65536 periods: both sides refuse
                    with pytest.raises(LecoError):
                        codec.filter_range_close(blk, bs, block_id=2, filter1=f1, filter2=f2, base=base)
                    continue
                got = codec.filter_range_close(blk, bs, block_id=2, filter1=f1, filter2=f2, base=base)
                assert np.array_equal(got, want), (width, kind, f1, f2, base)


@pytest.mark.parametrize("bs", [8193, 20000, 65536])
def test_filter_host_large_block(oracle_mod, bs):
    # one block split into several filter tiles (and look-back slots) on the GPU
    from leco_fix import Leco_int
    rng = np.random.default_rng(bs)
    for width, dt in ((32, np.uint32), (64, np.uint64)):
        codec = Leco_int(dt)
        codec.init(1, bs)
        y = np.sort(rng.integers(0, 1 << (width - 4), bs, dtype=np.uint64)).astype(dt)
        blk = codec.encodeArray8_int(y, bs)
        say(f"host filter large block width={width} bs={bs}")
        for f in (0, int(y[bs // 3]), int(y[-2]), int(y.max())):
            got = codec.filter_range(blk, bs, f, block_id=5)
            want = oracle_mod.filter_block(blk, bs, dt, f, 5 * bs)
            assert np.array_equal(got, want), (width, f)


def test_filter_batch_arena_column(oracle_mod):
    import torch
    from leco_fix.batch import Batch
    from leco_fix import datagen
    cols = [datagen.sorted_uniform(20_000, 64, 1), datagen.stepwise(30_000, 32, 2), datagen.sorted_uniform(9_000, 32, 3)]
    vals = [to_dev(torch, v) for v in cols]
    b = Batch(vals, [1000, 256, 4096])
    b.encode()
    torch.cuda.synchronize()
    say("filter on batch arena columns")
    for c, v in enumerate(cols):
        col = b.column(c)
        f = int(v[len(v) // 2])
        st, off = oracle_mod.encode_column(v, col.block_size)
        want = oracle_mod.filter_column(st, off, len(v), col.block_size, v.dtype, f)
        assert np.array_equal(positions(col.filter_range(f)), want), c


def test_filter_500m_sorted_keys():
    # 500M sorted piecewise-linear uint64 keys, block size 4096: for sorted data
    # the matches of `value > f` are exactly the suffix of positions past the
    # last key <= f; close-filter counts follow the keys' residues.
    import torch
    from leco_fix import Column, datagen
    n = 500_000_000
    say("filter 500M sorted_pwl_u64 bs=4096")
    vals = datagen.sorted_pwl_u64(n, seed=2, device="cuda")
    col = Column.encode(vals, 4096)
    for q in (0.25, 0.5, 0.999):
        f = int(vals[int(q * n)].item())
        expect = int((vals > f).sum().item())
        out = col.filter_range(f)
        assert out.numel() == expect
        if expect:
            first = int(out[0].item())
            assert first == n - expect
            steps = (out[1:].to(torch.int64) - out[:-1].to(torch.int64))
            assert bool((steps == 1).all())
        del out
