"""GPU: batched encode/decode of many mixed columns, hipGraph replay, and the
block-size autotuner, against the plain path and the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def to_dev(torch, a):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


def mixed_columns(seed, count=12, max_n=40_000):
    from leco_fix import datagen
    rng = np.random.default_rng(seed)
    dists = list(datagen.DISTRIBUTIONS.values())
    cols = []
    for c in range(count):
        width = 32 if c % 2 else 64
        n = int(np.exp(rng.uniform(np.log(50), np.log(max_n))))
        bs = int(rng.choice([1, 3, 256, 512, 700, 1000, 1024, 2048, 4096, 5000, 16384]))
        cols.append((dists[c % 4](n, width, int(rng.integers(1 << 30))), bs))
    return cols


def test_batch_matches_plain_and_oracle(oracle_mod):
    import torch
    from leco_fix.batch import Batch
    cols = mixed_columns(1)
    vals = [to_dev(torch, v) for v, _ in cols]
    say(f"batch of {len(cols)} mixed columns")
    b = Batch(vals, [bs for _, bs in cols])
    b.encode()
    outs = b.decode()
    torch.cuda.synchronize()
    for c, (v, bs) in enumerate(cols):
        want_st, want_off = oracle_mod.encode_column(v, bs)
        assert b.column_bytes(c) == want_st.tobytes(), c
        col = b.column(c)
        assert np.array_equal(col.offsets.cpu().numpy().view(np.uint64) - np.uint64(col.base), want_off)
        assert torch.equal(outs[c], vals[c]), c
        idx = torch.randint(0, len(v), (3000,), device="cuda")
        assert torch.equal(col.lookup(idx), vals[c][idx]), c
        a = len(v) // 3
        assert torch.equal(col.decode_range(a, len(v)), vals[c][a:]), c


def test_batch_graph_replay_is_stable():
    import torch
    from leco_fix.batch import Batch
    cols = mixed_columns(2, count=9)
    vals = [to_dev(torch, v) for v, _ in cols]
    b = Batch(vals, [bs for _, bs in cols])
    b.encode(graph=True)
    torch.cuda.synchronize()
    first = [b.column_bytes(c) for c in range(len(cols))]
    for rep in range(4):
        say(f"graph replay {rep}")
        for o in b.outs:
            o.zero_()
        b.encode(graph=True)
        b.decode(graph=True)
        torch.cuda.synchronize()
        for c in range(len(cols)):
            assert torch.equal(b.outs[c], vals[c]), (rep, c)
            assert b.column_bytes(c) == first[c], (rep, c)
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays


def test_batch_decode_grid_sizes():
    import torch
    from leco_fix.batch import Batch
    cols = mixed_columns(3, count=6)
    vals = [to_dev(torch, v) for v, _ in cols]
    b = Batch(vals, [bs for _, bs in cols])
    b.encode()
    for grid in [1, 7, 256, 4096]:
        b.set_decode_grid(grid)
        for o in b.outs:
            o.zero_()
        b.decode(graph=grid > 7)
        torch.cuda.synchronize()
        assert all(torch.equal(o, v) for o, v in zip(b.outs, vals)), grid


def test_autotune_bytes_match_oracle(oracle_mod):
    import torch
    from leco_fix.batch import autotune_block_sizes
    from leco_fix import datagen
    cand = (256, 1024, 4096)
    window, windows = 8192, 3
    cols = [datagen.sorted_uniform(50_000, 64, 5), datagen.stepwise(30_000, 32, 6),
            datagen.linear_sigma(5_000, 32, 7)]   # the last is shorter than a window
    bs, nbytes = autotune_block_sizes([to_dev(torch, v) for v in cols], cand, windows, window, tolerance=0.0)
    for c, v in enumerate(cols):
        slots = len(v) // window
        starts = [0] if slots == 0 else [(slots * k // min(windows, slots)) * window
                                          for k in range(min(windows, slots))]
        lens = [len(v)] if slots == 0 else [window] * len(starts)
        for j, cb in enumerate(cand):
            want = sum(len(oracle_mod.encode_column(v[s:s + ln], cb)[0]) for s, ln in zip(starts, lens))
            assert int(nbytes[c, j]) == want, (c, cb)
        assert bs[c] == cand[int(np.argmin(nbytes[c]))] or nbytes[c][cand.index(bs[c])] == nbytes[c].min()


def test_batch_tile_shapes_decode_exactly(oracle_mod):
    """Per-column decode tiles of small blocks: 4096, 8192 (default) and 16384
    positions per workgroup (at most 128 blocks), against the oracle's bytes
    and the input."""
    import torch
    from leco_fix.batch import Batch
    cols = mixed_columns(4, count=12)
    vals = [to_dev(torch, v) for v, _ in cols]
    tiles = [(4096, 0, 16384)[c % 3] for c in range(len(cols))]
    say(f"batch with tile shapes {tiles}")
    b = Batch(vals, [bs for _, bs in cols], tile_positions=tiles)
    b.encode()
    outs = b.decode()
    torch.cuda.synchronize()
    for c, (v, bs) in enumerate(cols):
        assert b.column_bytes(c) == oracle_mod.encode_column(v, bs)[0].tobytes(), c
        assert torch.equal(outs[c], vals[c]), (c, bs, tiles[c])
    for o in b.outs:
        o.zero_()
    b.decode(graph=True)
    torch.cuda.synchronize()
    assert all(torch.equal(o, v) for o, v in zip(b.outs, vals))
    b.close()


def _windows(n, bs, windows, window):
    slots = n // window
    if slots == 0:
        return [(0, n)]
    ln = max(bs, window // bs * bs)
    nw = min(windows, slots)
    return [((slots * k // nw) * window, min(ln, n - (slots * k // nw) * window)) for k in range(nw)]


def test_autotune_columns_picks_within_tolerance(oracle_mod):
    """autotune_columns: each pick's sample bytes per value equals the
    oracle's on the same windows and is within the tolerance of the best
    coarse candidate; tile shapes are staged exactly for 2048+ values; the
    batch built from the picks decodes exactly."""
    import torch
    from leco_fix import datagen
    from leco_fix.batch import Batch, autotune_columns
    cand = (256, 512, 1024, 2048, 4096, 8192)
    window, windows, tol = 8192, 3, 0.02
    cols = [datagen.sorted_uniform(60_000, 64, 5), datagen.stepwise(40_000, 32, 6),
            datagen.linear_sigma(5_000, 32, 7), datagen.sorted_pwl_u64(70_000, seed=8),
            datagen.sorted_zipf_gaps(50_000, 32, 9)]
    dev = [to_dev(torch, v) for v in cols]
    bss, tiles, info = autotune_columns(dev, cand, windows, window, tolerance=tol)
    say(f"autotune picks {list(zip(bss, tiles))}")

    def bpv(v, bs):
        ws = _windows(len(v), bs, windows, window)
        return sum(len(oracle_mod.encode_column(v[s:s + ln], bs)[0]) for s, ln in ws) / sum(ln for _, ln in ws)
    for c, v in enumerate(cols):
        got = bpv(v, bss[c])
        assert abs(info[c][0] - got) <= 1e-12 * got, (c, info[c][0], got)
        assert got <= (1 + tol) * min(bpv(v, cb) for cb in cand) * (1 + 1e-12), c
        assert (tiles[c] == 0) == (bss[c] >= 2048) and tiles[c] in (0, 8192, 16384)
        assert info[c][1] > 0
    b = Batch(dev, bss, tile_positions=tiles)
    b.encode()
    outs = b.decode()
    torch.cuda.synchronize()
    assert all(torch.equal(o, v) for o, v in zip(outs, dev))
    b.close()


def test_batch_decode_of_long_large_block_columns():
    # long columns of large blocks (several chunks per block, blocks both
    # inside and over the staging buffer) through the batch decode kernels,
    # whose output pointers come from the column table: every value, with and
    # without graph replay, against the column itself and the plain decode
    import torch
    from leco_fix import workloads
    from leco_fix.batch import Batch
    specs = [("linear_sigma", 10_760_805, 64, 3 * 100003 + 324, 16384),
             ("linear_sigma", 4_000_001, 64, 11, 8192),
             ("sorted_uniform", 3_000_000, 64, 12, 2048),
             ("linear_sigma", 5_000_000, 32, 13, 16384),
             ("sorted_zipf", 2_000_000, 64, 14, 1024)]
    vals = [workloads.gen_column(d, n, w, s) for d, n, w, s, _ in specs]
    say(f"batch of {len(specs)} long columns")
    b = Batch(vals, [bs for *_, bs in specs])
    b.encode()
    torch.cuda.synchronize()
    for graph in (False, True, True):
        for o in b.outs:
            o.zero_()
        outs = b.decode(graph=graph)
        torch.cuda.synchronize()
        for c, v in enumerate(vals):
            bad = (outs[c] != v).nonzero()
            assert bad.numel() == 0, (c, graph, bad[:8].flatten().tolist())
    for o in b.outs:
        o.zero_()
    b.decode_per_column()
    torch.cuda.synchronize()
    for c, v in enumerate(vals):
        assert torch.equal(b.outs[c], v), c
    b.close()
