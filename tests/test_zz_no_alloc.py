"""leco_fix_filter_range_noalloc: filter_range with the library's static
scratch instead of a workspace argument, for callers that cannot allocate.

GPU: its output equals leco_fix_filter_range_u32 / _u64's (which takes a
caller workspace) on two calls in a row, the second on a smaller column, so
the second call starts from the first call's scratch contents.  CPU: an input
whose workspace exceeds the scratch is refused before any HIP call."""
import ctypes
import sys

import numpy as np
import pytest


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def test_noalloc_rejects_inputs_beyond_the_scratch():
    from leco_fix import _lib
    L = _lib.lib()
    buf = (ctypes.c_uint8 * 64)()
    # bs 5000 tiles one block each: 2^32 values need about 880 MB of workspace
    n, bs = 1 << 32, 5000
    need = L.leco_fix_filter_workspace_bytes(n, bs)
    assert need > _lib.LECO_FIX_NOALLOC_SCRATCH_BYTES
    assert L.leco_fix_filter_range_noalloc(64, buf, buf, n, bs, 0, buf, 64, buf, None) == _lib.LECO_FIX_ESCRATCH
    assert L.leco_fix_filter_range_noalloc(48, buf, buf, 10, 4, 0, buf, 64, buf, None) == 1  # width
    assert L.leco_fix_filter_range_noalloc(32, buf, buf, 10, 4, 1 << 32, buf, 64, buf, None) == 1  # filter > u32
    assert L.leco_fix_filter_range_noalloc(64, None, buf, 10, 4, 0, buf, 64, buf, None) == 1
    assert b"scratch" in L.leco_fix_strerror(_lib.LECO_FIX_ESCRATCH)


def _filter_both(torch, col, f, cap):
    """(count, flags, positions) from the workspace entry point and from the
    no-alloc one, on the same column and filter."""
    from leco_fix import _lib
    from leco_fix.column import _stream_handle
    L = _lib.lib()
    n, bs, w = col.n, col.block_size, col.width
    wb = int(L.leco_fix_filter_workspace_bytes(n, bs))
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    out_a = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    out_b = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    res_a = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    res_b = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    fn = L.leco_fix_filter_range_u64 if w == 64 else L.leco_fix_filter_range_u32
    s = _stream_handle(None)
    rc = fn(col.stream.data_ptr(), col.offsets.data_ptr(), n, bs, f, out_a.data_ptr(), cap, res_a.data_ptr(),
            work.data_ptr(), wb, s)
    assert rc == 0, rc
    rc = L.leco_fix_filter_range_noalloc(w, col.stream.data_ptr(), col.offsets.data_ptr(), n, bs, f, out_b.data_ptr(),
                                         cap, res_b.data_ptr(), s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return res_a.cpu(), res_b.cpu(), out_a.cpu(), out_b.cpu()


@pytest.mark.gpu
@pytest.mark.parametrize("width,bs", [(64, 4096), (32, 1000), (64, 256)])
def test_noalloc_filter_equals_workspace_filter(width, bs):
    import torch
    from leco_fix import Column, datagen
    first, second = 3_000_000, 700_000  # the second call is smaller: it sees the first call's scratch
    if width == 64:
        vals = datagen.sorted_pwl_u64(first, seed=9)
        dev = torch.from_numpy(vals.view(np.int64)).to("cuda")
    else:
        vals = datagen.linear_noise_u32(first, seed=9, start=0)
        dev = torch.from_numpy(vals.view(np.int32)).to("cuda")
    for n in (first, second):
        col = Column.encode(dev[:n].clone(), bs)
        v = vals[:n]
        for f in (int(v[n // 3]), int(v[(4 * n) // 5])):
            say(f"noalloc width={width} bs={bs} n={n} filter={f}")
            res_a, res_b, out_a, out_b = _filter_both(torch, col, f, n)
            assert torch.equal(res_a, res_b), (res_a, res_b)
            cnt = int(res_a[0])
            assert cnt == int(np.count_nonzero(v > f))
            assert torch.equal(out_a[:cnt], out_b[:cnt])
            assert torch.equal(out_b[cnt:], torch.full((n - cnt,), -1, dtype=torch.int32))
