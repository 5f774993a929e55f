"""GPU: a seeded sweep of block sizes across every kernel boundary (wave
units up to 1024 values, multi-block decode below 2048, whole-block staging,
the one-chunk encoder above 4096, odd and prime sizes) and of lengths, widths
and distributions.  Every case must match the oracle byte for byte (LeCo,
FOR and Delta streams), decode, look up and range-decode to the input, and
give the oracle's filter_range positions."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECIAL = [1, 2, 3, 7, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 1023, 1024, 1025, 2047, 2048, 2049,
           3000, 4095, 4096, 4097, 8191, 8192, 8193, 10007, 65536, 70001]


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def dev(torch, a):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


def host(t, dt):
    return t.cpu().numpy().view(dt)


@pytest.mark.parametrize("seed", [101, 202])
def test_block_size_sweep_matches_oracle(oracle_mod, seed):
    import torch
    from leco_fix import Column, datagen
    from leco_fix.baselines import BaseColumn
    rng = np.random.default_rng(seed)
    dists = list(datagen.DISTRIBUTIONS.items())
    for case in range(90):
        bs = SPECIAL[case] if case < len(SPECIAL) else max(1, int(np.exp(rng.uniform(0, np.log(70000)))))
        width = (32, 64)[(case + seed) % 2]
        name, gen = dists[int(rng.integers(len(dists)))]
        n = int(min(400_000, max(1, rng.integers(1, 12) * bs + rng.integers(0, bs + 1))))
        vals = gen(n, width, int(rng.integers(1 << 30)))
        say(f"case {case} {name} width={width} bs={bs} n={n}")
        st, off = oracle_mod.encode_column(vals, bs, threads=8)
        col = Column.encode(dev(torch, vals), bs)
        assert np.array_equal(host(col.offsets, np.uint64), off), (case, "offsets")
        assert col.stream_bytes() == st.tobytes(), (case, "stream")
        assert np.array_equal(host(col.decode(), vals.dtype), vals), (case, "decode")
        idx = rng.integers(0, n, 3000, dtype=np.uint64)
        got = col.lookup(torch.from_numpy(idx.view(np.int64)).cuda())
        assert np.array_equal(host(got, vals.dtype), vals[idx]), (case, "lookup")
        a = int(rng.integers(0, n))
        b = int(rng.integers(a + 1, n + 1))
        assert np.array_equal(host(col.decode_range(a, b), vals.dtype), vals[a:b]), (case, "range", a, b)
        f = int(vals[int(rng.integers(0, n))])
        want = oracle_mod.filter_column(st, off, n, bs, vals.dtype, f)
        assert np.array_equal(host(col.filter_range(f), np.uint32), want), (case, "filter_range")
        for codec in ("for", "delta"):
            bst, _ = oracle_mod.base_encode_column(codec, vals, bs)
            bc = BaseColumn.encode(codec, dev(torch, vals), bs)
            assert bc.stream_bytes() == bst.tobytes(), (case, codec, "stream")
            assert np.array_equal(host(bc.decode(), vals.dtype), vals), (case, codec, "decode")


def test_close_filter_sweep_on_sorted_columns(oracle_mod):
    """filter_range_close on sorted columns (the scan path: each value tested
    in its own period) with bases from 1 to 2^31 - 1 and windows inside one
    period, the per-block filters derived from block minimums as the
    reference driver does.  Where the oracle stops at the period limit, the
    GPU call must fail too; otherwise the positions must be the oracle's."""
    import torch
    from leco_fix import Column, datagen
    from leco_fix._lib import LecoError
    rng = np.random.default_rng(43)
    gens = [datagen.sorted_uniform, datagen.linear_sigma, datagen.stepwise]
    compared = 0
    for case in range(80):
        width = (32, 64)[case % 2]
        bs = int(rng.choice([64, 256, 1000, 2048, 4096, 8193]))
        n = min(200_000, int(rng.integers(bs, 20 * bs)))
        vals = np.sort(gens[case % 3](n, width, int(rng.integers(1 << 30))).astype(np.uint64))
        if width == 32:  # below 2^31: the 32-bit comparisons are signed
            vals = np.sort(vals & np.uint64(0x7FFFFFFF))
        vals = vals.astype(np.uint32 if width == 32 else np.uint64)
        base = max(1, min(int(np.exp(rng.uniform(0, np.log(2**31 - 1)))), 2**31 - 1))
        g1 = int(rng.integers(0, base))
        g2 = int(rng.integers(g1, base + 1))
        st, off = oracle_mod.encode_column(vals, bs, threads=8)
        col = Column.encode(dev(torch, vals), bs)
        mins = [int(vals[b * bs:(b + 1) * bs].min()) for b in range(col.num_blocks)]
        mask = (1 << width) - 1
        o1 = np.array([(m - m % base + g1) & mask for m in mins], np.uint64)
        o2 = np.array([(m - m % base + g2) & mask for m in mins], np.uint64)
        t1, t2 = dev(torch, o1.astype(vals.dtype)), dev(torch, o2.astype(vals.dtype))
        say(f"close case {case} width={width} bs={bs} n={n} base={base} window=({g1}, {g2})")
        try:
            want = oracle_mod.filter_close_column(st, off, n, bs, vals.dtype, o1, o2, base)
        except RuntimeError:
            with pytest.raises(LecoError):
                col.filter_range_close(t1, t2, base)
            continue
        got = host(col.filter_range_close(t1, t2, base), np.uint32)
        assert np.array_equal(got, want), (case, width, bs, base)
        compared += 1
    assert compared >= 40
