"""GPU parity: the gfx950 library, called through its C ABI, against the
oracle and the golden fixtures.  Streams must be byte-identical, decoded
values and lookups bit-exact."""
import sys

import numpy as np
import pytest

from conftest import golden_cases, load_golden

pytestmark = pytest.mark.gpu

CASES = [c["name"] for c in golden_cases()]


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def to_dev(torch, a: np.ndarray):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


def to_host(t, dtype) -> np.ndarray:
    return t.cpu().numpy().view(dtype)


@pytest.mark.parametrize("name", CASES)
def test_golden_stream_decode_lookup(torch_dev, name):
    from leco_fix import Column
    torch = torch_dev
    g = load_golden(name)
    vals, bs = g["values"], int(g["bs"])
    say(f"golden {name} n={len(vals)} bs={bs} width={8 * vals.dtype.itemsize}")
    col = Column.encode(to_dev(torch, vals), bs)
    assert np.array_equal(to_host(col.offsets, np.uint64), g["offsets"])
    assert col.stream_bytes() == g["stream"].tobytes()
    assert np.array_equal(to_host(col.decode(), vals.dtype), vals)
    idx = torch.from_numpy(g["lookup_idx"].view(np.int64)).to("cuda")
    assert np.array_equal(to_host(col.lookup(idx), vals.dtype), g["lookup_val"])
    assert torch.equal(col.rebuild_params(), col.params)


BS_LIST = [1, 2, 3, 17, 255, 256, 512, 513, 1000, 1024, 1025, 2047, 2048, 2049, 4096, 5000, 10000, 16384]


def test_random_columns_match_oracle(torch_dev, oracle_mod):
    from leco_fix import Column, datagen
    torch = torch_dev
    rng = np.random.default_rng(77)
    dists = list(datagen.DISTRIBUTIONS.items())
    for k, bs in enumerate(BS_LIST):
        for width in (32, 64):
            name, gen = dists[(k + width) % len(dists)]
            n = int(rng.integers(1, 6) * bs + rng.integers(0, bs)) if bs > 1 else int(rng.integers(1, 50))
            n = min(n, 120_000)
            vals = gen(n, width, int(rng.integers(1 << 30)))
            say(f"random {name} width={width} bs={bs} n={n}")
            want_st, want_off = oracle_mod.encode_column(vals, bs, threads=8)
            col = Column.encode(to_dev(torch, vals), bs)
            assert np.array_equal(to_host(col.offsets, np.uint64), want_off), (name, width, bs)
            assert col.stream_bytes() == want_st.tobytes(), (name, width, bs)
            assert np.array_equal(to_host(col.decode(), vals.dtype), vals)
            idx = rng.integers(0, n, 5000, dtype=np.uint64)
            got = col.lookup(torch.from_numpy(idx.view(np.int64)).to("cuda"))
            assert np.array_equal(to_host(got, vals.dtype), vals[idx])
            for _ in range(3):
                a = int(rng.integers(0, n))
                b = int(rng.integers(a + 1, n + 1))
                assert np.array_equal(to_host(col.decode_range(a, b), vals.dtype), vals[a:b])


def test_raw_and_edge_distributions_match_oracle(torch_dev, oracle_mod):
    from leco_fix import Column
    torch = torch_dev
    rng = np.random.default_rng(3)
    for width, bs in [(64, 4096), (32, 4096), (64, 3000), (32, 2500), (64, 512), (32, 1000), (64, 1024)]:
        top = (1 << width) - 1
        parts = [rng.integers(0, top, 3 * bs, dtype=np.uint64, endpoint=True),          # raw blocks
                 np.sort(rng.integers(top >> 1, top, 2 * bs, dtype=np.uint64)),          # >= 2^(w-1)
                 np.full(bs + 7, top, np.uint64)]                                          # saturated
        vals = np.concatenate(parts).astype(np.uint32 if width == 32 else np.uint64)
        say(f"edge width={width} bs={bs} n={len(vals)}")
        want_st, want_off = oracle_mod.encode_column(vals, bs, threads=8)
        col = Column.encode(to_dev(torch, vals), bs)
        assert col.stream_bytes() == want_st.tobytes()
        assert np.array_equal(to_host(col.decode(), vals.dtype), vals)


def test_every_bit_width_small_and_large_blocks(torch_dev, oracle_mod):
    # Block k's residuals span about (k mod (W+1)) bits on top of a line, so
    # consecutive blocks take every width l = 0..W (raw included), for block
    # sizes encoded one per wave (<= 1024, one or two LDS chunks) and one per
    # workgroup.
    from leco_fix import Column
    torch = torch_dev
    rng = np.random.default_rng(11)
    for width in (32, 64):
        for bs in (64, 300, 512, 777, 1024, 1500, 4096):
            nb = width + 3
            n = nb * bs - bs // 3
            i = np.arange(n, dtype=np.uint64)
            bits = (i // np.uint64(bs)) % np.uint64(width + 1)
            noise = rng.integers(0, 1 << 62, n, dtype=np.uint64) | (rng.integers(0, 4, n, dtype=np.uint64) << np.uint64(62))
            noise = np.where(bits == 0, np.uint64(0),
                             noise >> (np.uint64(64) - np.maximum(bits, np.uint64(1))))
            vals = (np.uint64(12345) + np.uint64(977) * i + noise)
            vals = vals.astype(np.uint32) if width == 32 else vals
            say(f"every width: width={width} bs={bs} n={n}")
            want_st, want_off = oracle_mod.encode_column(vals, bs, threads=8)
            col = Column.encode(to_dev(torch, vals), bs)
            assert np.array_equal(to_host(col.offsets, np.uint64), want_off), (width, bs)
            assert col.stream_bytes() == want_st.tobytes(), (width, bs)
            assert np.array_equal(to_host(col.decode(), vals.dtype), vals), (width, bs)
            idx = rng.integers(0, n, 4000, dtype=np.uint64)
            got = col.lookup(torch.from_numpy(idx.view(np.int64)).to("cuda"))
            assert np.array_equal(to_host(got, vals.dtype), vals[idx]), (width, bs)


def test_prediction_range_edges_and_alignment(torch_dev, oracle_mod):
    # Full 4096-value blocks around the +-2^51 prediction range that the
    # encoder's branch-free prediction covers: a fit inside the range whose
    # recentred theta0 leaves it (one large outlier), a line that crosses
    # 2^51 inside the block, lines just inside the range, and values above
    # 2^63; then the same column from an address that is 8 but not 16-byte
    # aligned.  Byte-identical to the oracle either way.
    from leco_fix import Column
    torch = torch_dev
    bs = 4096
    i = np.arange(bs, dtype=np.uint64)
    lim = np.uint64(1 << 51)
    blocks = []
    out_up = np.full(bs, lim - np.uint64(1 << 30), np.uint64) + i
    out_up[2000] += np.uint64(1 << 40)
    blocks.append(out_up)
    out_dn = np.full(bs, lim - np.uint64(1 << 30), np.uint64) + i
    out_dn[100] -= np.uint64(1 << 40)
    blocks.append(out_dn)
    blocks.append(lim - np.uint64(2048) + i)                       # crosses 2^51
    blocks.append(lim - np.uint64(5000) + np.uint64(3) * (i // np.uint64(3)))  # just inside
    blocks.append(np.uint64(1 << 63) + np.uint64(12345) * i)        # above 2^63
    blocks.append((lim - np.uint64(1)) - i)                         # decreasing from just inside
    vals = np.concatenate(blocks + [np.arange(777, dtype=np.uint64)])
    want_st, want_off = oracle_mod.encode_column(vals, bs, threads=8)
    say(f"prediction range edges n={len(vals)}")
    col = Column.encode(to_dev(torch, vals), bs)
    assert np.array_equal(to_host(col.offsets, np.uint64), want_off)
    assert col.stream_bytes() == want_st.tobytes()
    assert np.array_equal(to_host(col.decode(), np.uint64), vals)

    say("prediction range edges: input 8 but not 16-byte aligned")
    padded = to_dev(torch, np.concatenate([np.zeros(1, np.uint64), vals]))
    view = padded[1:]
    assert view.data_ptr() % 16 == 8
    col2 = Column.encode(view, bs)
    assert col2.stream_bytes() == want_st.tobytes()
    for width in (32, 64):
        rng = np.random.default_rng(width)
        y = np.sort(rng.integers(0, 1 << (width - 2), 5 * bs + 100, dtype=np.uint64))
        y = y.astype(np.uint32) if width == 32 else y
        want_st, _ = oracle_mod.encode_column(y, bs, threads=8)
        padded = to_dev(torch, np.concatenate([np.zeros(1, y.dtype), y]))
        say(f"unaligned sorted width={width}")
        assert Column.encode(padded[1:], bs).stream_bytes() == want_st.tobytes()
        assert Column.encode(padded[1:].clone(), bs).stream_bytes() == want_st.tobytes()


def test_reference_cpu_workload_on_gpu(torch_dev, oracle_mod):
    # 10M uint32, y_i = clamp(round(1000 i + N(0, 50^2))), block size 10,000,
    # 1M uniform lookups: byte-identical to the oracle and exact round trip.
    from leco_fix import Column, datagen
    torch = torch_dev
    n, bs = 10_000_000, 10_000
    vals = datagen.linear_noise_u32(n, seed=1)
    say("linear_noise_u32 n=10M bs=10000")
    want_st, want_off = oracle_mod.encode_column(vals, bs, threads=8)
    col = Column.encode(to_dev(torch, vals), bs)
    assert col.nbytes == len(want_st)
    assert col.stream_bytes() == want_st.tobytes()
    assert np.array_equal(to_host(col.decode(), np.uint32), vals)
    idx = np.random.default_rng(1).integers(0, n, 1_000_000, dtype=np.uint64)
    got = col.lookup(torch.from_numpy(idx.view(np.int64)).to("cuda"))
    assert np.array_equal(to_host(got, np.uint32), vals[idx])


def test_sorted_pwl_500m_full_size(torch_dev, oracle_mod):
    # 500M sorted piecewise-linear uint64 keys at block size 4096 generated in
    # HBM: round trip and 64M lookups exact on the whole column; 512 sampled
    # blocks byte-identical to the oracle's encoding of the same values.
    from leco_fix import Column, datagen
    torch = torch_dev
    n, bs = 500_000_000, 4096
    say("sorted_pwl_u64 n=500M bs=4096: generate")
    vals = datagen.sorted_pwl_u64(n, seed=2, device="cuda")
    say("encode")
    col = Column.encode(vals, bs)
    assert 0.05 < col.compression_ratio < 0.5
    say("decode")
    out = col.decode()
    assert torch.equal(out, vals)
    del out
    say("lookup 64M")
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    idx = torch.randint(0, n, (64_000_000,), device="cuda", generator=g, dtype=torch.int64)
    assert torch.equal(col.lookup(idx), vals[idx])
    del idx
    say("sampled blocks vs oracle")
    nb = col.num_blocks
    offs = to_host(col.offsets, np.uint64)
    rng = np.random.default_rng(11)
    blocks = np.unique(np.concatenate([[0, 1, nb - 1], rng.integers(0, nb, 509)]))
    for b in blocks:
        b = int(b)
        lo, hi = b * bs, min((b + 1) * bs, n)
        y = to_host(vals[lo:hi], np.uint64)
        want = oracle_mod.encode_block(y)
        got = col.stream[int(offs[b]):int(offs[b + 1])].cpu().numpy().tobytes()
        assert got == want, b
    assert torch.equal(col.rebuild_params(), col.params)


def test_leco_int_host_surface(torch_dev):
    # the reference's per-block calls on host buffers, against the golden bytes
    from leco_fix import Leco_int
    for name in ["cfg1_kink_u32", "pwl_u64", "bs1_u32", "uniform_u64", "lw1_u64"]:
        g = load_golden(name)
        vals, bs = g["values"], int(g["bs"])
        offs = g["offsets"]
        stream = g["stream"].tobytes()
        codec = Leco_int(vals.dtype)
        codec.init(len(offs) - 1, bs)
        assert codec.get_block_nums() == len(offs) - 1
        for b in range(len(offs) - 1):
            y = vals[b * bs:(b + 1) * bs]
            blk = stream[int(offs[b]):int(offs[b + 1])]
            assert codec.encodeArray8_int(y, len(y), None, b) == blk
            assert np.array_equal(codec.decodeArray8(blk, len(y), None, b), y)
            for j in {0, len(y) // 2, len(y) - 1}:
                assert codec.randomdecodeArray8(blk, j, None, b) == int(y[j])


@pytest.mark.parametrize("width", [32, 64])
def test_cpp_leco_int_driver(torch_dev, width):
    # the reference driver's flow, in C++ against include/leco_int.hpp
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "cpp", "leco_int_driver")
    say(f"cpp driver width={width}")
    r = subprocess.run([exe, "60000", "7", str(width)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")


@pytest.mark.parametrize("bs", [1, 7, 256, 1000, 2048, 4096, 16384])
def test_decode_shapes_agree(torch_dev, bs):
    # staged (one workgroup per block) and value-parallel (one lane per
    # position) decode give the same values on every kind of block
    from leco_fix import Column, datagen
    from leco_fix.column import DECODE_STAGED, DECODE_VALUES
    torch = torch_dev
    rng = np.random.default_rng(bs)
    for width in (32, 64):
        top = (1 << width) - 1
        parts = [datagen.sorted_uniform(5 * bs + 3, width, bs).astype(np.uint64),
                 rng.integers(0, top, 2 * bs + 1, dtype=np.uint64, endpoint=True),       # raw blocks
                 np.full(bs + 5, 12345, np.uint64)]                                        # constant blocks
        vals = np.concatenate(parts).astype(np.uint32 if width == 32 else np.uint64)
        say(f"decode shapes width={width} bs={bs} n={len(vals)}")
        col = Column.encode(to_dev(torch, vals), bs)
        for shape in (DECODE_STAGED, DECODE_VALUES):
            assert np.array_equal(to_host(col.decode(shape=shape), vals.dtype), vals), shape
            a, b = len(vals) // 3, len(vals) - 2
            assert np.array_equal(to_host(col.decode_range(a, b, shape=shape), vals.dtype), vals[a:b]), shape


@pytest.mark.parametrize("bs", [1, 7, 100, 256, 1000, 2047])
def test_small_block_ranges_cross_tiles(torch_dev, bs):
    # small blocks are decoded several per workgroup (tiles of up to 8192
    # positions): ranges that start and end inside tiles and blocks, span
    # several tiles, or sit inside one block, against the values and the
    # value-parallel shape
    from leco_fix import Column, datagen
    from leco_fix.column import DECODE_STAGED, DECODE_VALUES
    torch = torch_dev
    rng = np.random.default_rng(1000 + bs)
    n = 50_000
    for width in (32, 64):
        top = (1 << width) - 1
        vals = datagen.sorted_uniform(n, width, bs).astype(np.uint64)
        raw_at = int(rng.integers(0, n - 3000))
        vals[raw_at:raw_at + 3000] = rng.integers(0, top, 3000, dtype=np.uint64, endpoint=True)
        vals = vals.astype(np.uint32 if width == 32 else np.uint64)
        say(f"small-block ranges width={width} bs={bs} n={n}")
        col = Column.encode(to_dev(torch, vals), bs)
        assert np.array_equal(to_host(col.decode(), vals.dtype), vals)
        ranges = [(0, 1), (n - 1, n), (8191, 8193), (0, n), (bs // 2, bs // 2 + 1)]
        for _ in range(12):
            a = int(rng.integers(0, n))
            ranges.append((a, int(rng.integers(a + 1, min(n, a + 30_000) + 1))))
        for a, b in ranges:
            for shape in (DECODE_STAGED, DECODE_VALUES):
                got = to_host(col.decode_range(a, b, shape=shape), vals.dtype)
                assert np.array_equal(got, vals[a:b]), (width, a, b, shape)


@pytest.mark.parametrize("bs", [256, 1000, 2048, 4096])
def test_decode_into_unaligned_output(torch_dev, bs):
    # whole 64-bit chunks and small-block pairs are written with 16-byte stores
    # when their destination is 16-byte aligned: output buffers that start 8
    # bytes past a 16-byte boundary, ranges of even and odd starts, and both
    # together take the per-value stores instead, with the same values
    from leco_fix import Column, datagen
    torch = torch_dev
    rng = np.random.default_rng(4242 + bs)
    n = 6 * bs + 77
    vals = datagen.sorted_uniform(n, 64, bs)
    say(f"unaligned output bs={bs} n={n}")
    col = Column.encode(to_dev(torch, vals), bs)
    backing = torch.zeros(n + 2, dtype=torch.int64, device="cuda")
    for shift in (0, 1):
        out = backing[shift:shift + n]
        assert (out.data_ptr() % 16 == 0) == (shift == 0)
        col.decode(out=out)
        assert np.array_equal(to_host(out, np.uint64), vals), shift
        for a, b in [(0, n), (1, n), (bs, 3 * bs), (bs + 1, 3 * bs + 1), (2 * bs - 1, n - 1)] + \
                [tuple(sorted(rng.integers(0, n + 1, 2))) for _ in range(6)]:
            if a == b:
                continue
            out = backing[shift:shift + (b - a)]
            col.decode_range(a, b, out=out)
            assert np.array_equal(to_host(out, np.uint64), vals[a:b]), (shift, a, b)
