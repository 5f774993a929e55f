"""GPU: every encoder path on columns of 32M values against the oracle, byte
for byte (streams, offsets, op_max fix-up flags).

Small columns do not load the GPU enough to expose timing-dependent register
hazards: the FOR uint32 encoder once wrote wrong bytes into several percent
of the blocks of a 100M-value column while every small-column test passed.
These columns keep every CU busy for milliseconds per launch."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 32_000_000


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def _check(got_bytes, got_offsets, stream, offsets):
    assert np.array_equal(got_offsets.cpu().numpy().view(np.uint64), offsets)
    g = np.frombuffer(got_bytes, np.uint8)
    assert g.size == stream.size
    bad = np.nonzero(g != stream)[0]
    assert bad.size == 0, f"{bad.size} stream bytes differ, first at {bad[:8].tolist()}"


@pytest.fixture(scope="module")
def columns():
    import torch
    from leco_fix import datagen
    from leco_fix import workloads as W
    pwl = datagen.sorted_pwl_u64(N, seed=5, device="cuda")
    u32 = W.gen_column("sorted_uniform", N, 32, 17)
    cols = {"u64": (pwl, pwl.cpu().numpy().view(np.uint64)), "u32": (u32, u32.cpu().numpy().view(np.uint32))}
    yield cols
    del cols
    torch.cuda.empty_cache()


@pytest.mark.parametrize("width", ["u64", "u32"])
@pytest.mark.parametrize("bs", [256, 2048, 4096, 16384])
def test_leco_encode_at_scale(oracle_mod, columns, width, bs):
    import torch
    from leco_fix import Column
    vals, host = columns[width]
    say(f"LeCo {width} n={N} bs={bs}")
    stream, offsets = oracle_mod.encode_column(host, bs)
    col = Column.encode(vals, bs)
    torch.cuda.synchronize()
    _check(col.stream_bytes(), col.offsets, stream, offsets)


@pytest.mark.parametrize("codec", ["for", "delta"])
@pytest.mark.parametrize("width", ["u64", "u32"])
@pytest.mark.parametrize("bs", [256, 4096, 8192])
def test_base_encode_at_scale(oracle_mod, columns, codec, width, bs):
    import torch
    from leco_fix.baselines import BaseColumn
    vals, host = columns[width]
    say(f"{codec} {width} n={N} bs={bs}")
    stream, offsets = oracle_mod.base_encode_column(codec, host, bs)
    col = BaseColumn.encode(codec, vals, bs)
    torch.cuda.synchronize()
    _check(col.stream_bytes(), col.offsets, stream, offsets)


@pytest.mark.parametrize("bs", [2000, 200])
def test_opmax_encode_at_scale(oracle_mod, bs):
    import torch
    from leco_fix.opmax import OpMaxColumn
    g = torch.Generator(device="cuda")
    g.manual_seed(23)
    i = torch.arange(N, device="cuda", dtype=torch.float64)
    y = torch.round(1e6 + 21.3 * i + torch.randint(-3, 4, (N,), device="cuda", generator=g)).to(torch.int64)
    y = y.to(torch.int32)
    del i
    say(f"op_max n={N} bs={bs}")
    stream, offsets, fix = oracle_mod.opmax_encode_column(y.cpu().numpy().view(np.uint32), bs)
    col = OpMaxColumn.encode(y, bs)
    torch.cuda.synchronize()
    _check(col.stream_bytes(), col.offsets, stream, offsets)
    assert np.array_equal(col.fix.cpu().numpy().view(np.int8), fix.view(np.int8))
