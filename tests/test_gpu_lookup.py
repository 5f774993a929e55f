"""GPU: grouped point lookups (leco_fix_lookup_grouped_*) equal the per-query
kernel and the input on the same queries: random, duplicated, sorted,
reversed and out-of-range indices, power-of-two and other block sizes, both
widths, batch sizes that do not fill the last counting chunk, and one- and
few-query batches."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def _queries(torch, n, nq, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    q = torch.randint(0, n, (nq,), device="cuda", generator=g, dtype=torch.int64)
    k = nq // 8
    q[:k] = q[k:2 * k]                                   # duplicates
    q[2 * k:3 * k] = torch.sort(q[2 * k:3 * k]).values   # a sorted run
    q[3 * k:4 * k] = torch.flip(torch.sort(q[3 * k:4 * k]).values, [0])
    bad = torch.tensor([n, n + 5, 1 << 40, (1 << 62) + 3], device="cuda", dtype=torch.int64)
    q[4 * k:4 * k + 4] = bad                             # out of range: 0
    q[5 * k] = 0
    q[5 * k + 1] = n - 1
    return q


def _check(torch, col, vals, q):
    got = col.lookup(q, grouped=True)
    ref = col.lookup(q, grouped=False)
    assert torch.equal(got, ref)
    ok = q < vals.numel()
    assert torch.equal(got[ok], vals[q[ok]])
    assert int(got[~ok].abs().sum().item()) == 0


@pytest.mark.parametrize("bs,nq", [(4096, 3_000_017), (1000, 1_048_576), (2048, 65_537)])
def test_grouped_lookup_u64(bs, nq):
    import torch
    from leco_fix import Column, datagen
    n = 40_000_000
    vals = datagen.sorted_pwl_u64(n, seed=7, device="cuda")
    col = Column.encode(vals, bs)
    say(f"grouped lookup u64 n={n} bs={bs} nq={nq}")
    _check(torch, col, vals, _queries(torch, n, nq, bs))


@pytest.mark.parametrize("bs", [256, 4096, 10_000])
def test_grouped_lookup_u32(bs):
    import torch
    from leco_fix import Column
    from leco_fix import workloads as W
    n = 25_000_003
    vals = W.gen_column("linear_sigma", n, 32, 31)
    col = Column.encode(vals, bs)
    say(f"grouped lookup u32 n={n} bs={bs}")
    _check(torch, col, vals, _queries(torch, n, 2_000_000, bs + 1))


def test_grouped_lookup_small_batches_and_small_columns():
    import torch
    from leco_fix import Column, datagen
    for n, bs in [(1, 1), (5_000, 7), (3_000_000, 4096)]:
        vals = torch.from_numpy(datagen.sorted_pwl_u64(n, seed=n).view(np.int64)).cuda()
        col = Column.encode(vals, bs)
        for nq in (1, 2, 255, 4097):
            say(f"grouped lookup n={n} bs={bs} nq={nq}")
            g = torch.Generator(device="cuda")
            g.manual_seed(nq)
            q = torch.randint(0, n + 2, (nq,), device="cuda", generator=g, dtype=torch.int64)
            _check(torch, col, vals, q)


def test_grouped_lookup_on_a_side_stream_and_twice_with_one_workspace():
    import ctypes

    import torch
    from leco_fix import Column, datagen
    from leco_fix._lib import lib
    n = 8_000_000
    vals = datagen.sorted_pwl_u64(n, seed=3, device="cuda")
    col = Column.encode(vals, 4096)
    L = lib()
    nq = 1_500_000
    ws = int(L.leco_fix_lookup_grouped_workspace(nq))
    work = torch.empty(ws, dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    outs = []
    for seed in (1, 2):
        q = _queries(torch, n, nq, seed)
        out = torch.empty(nq, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        say(f"grouped lookup side stream, call {seed}")
        assert L.leco_fix_lookup_grouped_u64(col.stream.data_ptr(), col.params.data_ptr(), n, 4096, q.data_ptr(),
                                             nq, out.data_ptr(), work.data_ptr(), ws, ctypes.c_void_p(s.cuda_stream)) == 0
        outs.append((q, out))
    s.synchronize()
    for q, out in outs:
        assert torch.equal(out, col.lookup(q, grouped=False))
