"""GPU: piecewise_fix_op_max (the uint32 IntegerCODEC) against the golden
fixtures written by tests/spec_opmax.py and against the oracle: streams byte
for byte, fix-up flags and lists position for position, the reference
decoder's output, decode with fix-ups == input, lookups == input, the host
one-block surface, and a 200M-value column through round trips."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "opmax_cases.json")) as _f:
    OPMAX = [c["name"] for c in json.load(_f)]


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def dev32(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to("cuda")


def host32(t):
    return t.cpu().numpy().view(np.uint32)


def raw_mask(stream, offs, n, bs):
    """True at positions of raw blocks (tmp_bit >= 31)."""
    m = np.zeros(n, bool)
    for b in range(len(offs) - 1):
        if stream[int(offs[b])] >= 31:
            m[b * bs:min((b + 1) * bs, n)] = True
    return m


@pytest.mark.parametrize("case", OPMAX)
def test_opmax_golden_on_gpu(case):
    import torch
    from leco_fix.opmax import OpMaxColumn
    with np.load(os.path.join(GOLDEN, f"{case}.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    y, bs = g["values"], int(g["bs"])
    say(f"golden {case} n={len(y)} bs={bs}")
    col = OpMaxColumn.encode(dev32(torch, y), bs)
    assert np.array_equal(col.offsets.cpu().numpy().view(np.uint64), g["offsets"])
    assert col.stream_bytes() == g["stream"].tobytes()
    plus, minus = col.fix_lists()
    assert plus.cpu().numpy().tolist() == g["plus"].tolist()
    assert minus.cpu().numpy().tolist() == g["minus"].tolist()
    raw_dec = col.decode(fixed=False)
    assert np.array_equal(host32(raw_dec), g["decoded"])
    assert np.array_equal(host32(col.decode(fixed=True)), y)
    assert np.array_equal(host32(col.decode_lists(plus, minus)), y)
    # the harness's correction on the reference decoder's output
    fixed = host32(col.apply_lists(raw_dec.clone(), plus, minus))
    keep = ~raw_mask(g["stream"], g["offsets"], len(y), bs)
    assert np.array_equal(fixed[keep], y[keep])
    idx = np.arange(len(y), dtype=np.int64)
    assert np.array_equal(host32(col.lookup(torch.from_numpy(idx).cuda())), y)


def _data(kind, n, rng):
    i = np.arange(n, dtype=np.float64)
    if kind == "linear":
        return np.round(1e6 + 21.3 * i + rng.integers(-3, 4, n)).astype(np.uint32)
    if kind == "drift":
        return np.round(3.0e9 + 0.4123457 * (i % 200000) + (i // 200000) * 5e5).astype(np.uint32)
    if kind == "sorted":
        return np.sort(rng.integers(0, 2**32, n, dtype=np.uint64)).astype(np.uint32)
    if kind == "mixed":
        k = n // 4
        return np.concatenate([np.sort(rng.integers(0, 2**32, k, dtype=np.uint64)),
                               np.full(k, 77, np.uint64),
                               rng.integers(0, 2**32, k, dtype=np.uint64),
                               rng.integers(0, 1 << 20, n - 3 * k, dtype=np.uint64)]).astype(np.uint32)
    raise ValueError(kind)


@pytest.mark.parametrize("bs", [1, 3, 7, 64, 200, 1000, 2000, 4096, 65536])
@pytest.mark.parametrize("kind", ["linear", "drift", "mixed"])
def test_opmax_matches_oracle(kind, bs):
    import torch
    from oracle import oracle
    from leco_fix.opmax import OpMaxColumn
    rng = np.random.default_rng(bs * 7 + len(kind))
    n = 300_001 if bs >= 64 else 20_011
    y = _data(kind, n, rng)
    say(f"oracle {kind} n={n} bs={bs}")
    stream, offs, fix = oracle.opmax_encode_column(y, bs)
    col = OpMaxColumn.encode(dev32(torch, y), bs)
    assert np.array_equal(col.offsets.cpu().numpy().view(np.uint64), offs)
    assert col.stream_bytes() == stream.tobytes()
    assert np.array_equal(col.fix.cpu().numpy(), fix)
    assert np.array_equal(host32(col.decode(fixed=False)), oracle.opmax_decode_column(stream, offs, n, bs))
    assert np.array_equal(host32(col.decode(fixed=True)), y)
    plus, minus = col.fix_lists()
    op, om = oracle.opmax_fix_lists(fix)
    assert np.array_equal(plus.cpu().numpy(), op) and np.array_equal(minus.cpu().numpy(), om)
    assert np.array_equal(host32(col.decode_lists(plus, minus)), y)
    idx = rng.integers(0, n, 50_000).astype(np.int64)
    assert np.array_equal(host32(col.lookup(torch.from_numpy(idx).cuda())), y[idx])


def test_opmax_params_rebuilt_from_stream():
    import torch
    from oracle import oracle
    from leco_fix.opmax import OpMaxColumn
    rng = np.random.default_rng(5)
    y = _data("mixed", 100_000, rng)
    stream, offs, fix = oracle.opmax_encode_column(y, 2000)
    col = OpMaxColumn.from_stream(stream.tobytes(), offs, len(y), 2000, fix=fix)
    assert np.array_equal(host32(col.decode(fixed=True)), y)
    idx = rng.integers(0, len(y), 20_000).astype(np.int64)
    assert np.array_equal(host32(col.lookup(torch.from_numpy(idx).cuda())), y[idx])
    enc = OpMaxColumn.encode(dev32(torch, y), 2000)
    assert torch.equal(enc.params, col.params)


def test_opmax_fix_list_capacity():
    import torch
    from leco_fix.opmax import OpMaxColumn
    y = _data("drift", 600_000, np.random.default_rng(9))
    col = OpMaxColumn.encode(dev32(torch, y), 60000)
    plus, minus = col.fix_lists()
    total = len(plus) + len(minus)
    assert total > 0
    cp, cm = col.fix_lists(capacity=3)
    assert cp.tolist() == plus[:3].tolist() and cm.tolist() == minus[:3].tolist()


def test_opmax_host_block_surface():
    from oracle import oracle
    from leco_fix.opmax import piecewise_fix_op_max
    rng = np.random.default_rng(11)
    y = np.concatenate([_data("drift", 60_000, rng)[:59_000], _data("mixed", 4_003, rng)])
    bs = 30_000
    codec = piecewise_fix_op_max()
    nb = (len(y) + bs - 1) // bs
    codec.init(nb, bs, 32)
    blocks = []
    for b in range(nb):
        seg = y[b * bs:(b + 1) * bs]
        blk = codec.encodeArray8(seg, len(seg), None, b)
        oblk, ofix = oracle.opmax_encode_block(seg)
        assert blk == oblk
        blocks.append(blk)
    stream, offs, fix = oracle.opmax_encode_column(y, bs)
    op, om = oracle.opmax_fix_lists(fix)
    assert codec.mul_add_diff_set == op.tolist() and codec.mul_add_diff_set_minus == om.tolist()
    assert len(codec.mul_add_diff_set) + len(codec.mul_add_diff_set_minus) > 0
    recover = np.concatenate([codec.decodeArray8(blk, min(bs, len(y) - b * bs)) for b, blk in enumerate(blocks)])
    assert np.array_equal(recover, oracle.opmax_decode_column(stream, offs, len(y), bs))
    for j in rng.integers(0, len(y), 200):
        b, k = divmod(int(j), bs)
        assert codec.randomdecodeArray8(blocks[b], k) == int(y[j])
    assert codec.get_block_nums() == 1 and codec.name() == "piecewise_fix_op_max"


def test_opmax_200m_round_trip():
    """The published uint32 configuration's shape: 200M values, 100000 blocks (bs 2000)."""
    import torch
    from oracle import oracle
    from leco_fix.opmax import OpMaxColumn
    n, bs = 200_000_000, 2000
    say(f"200M round trip n={n} bs={bs}")
    g = torch.Generator(device="cuda").manual_seed(3)
    i = torch.arange(n, device="cuda", dtype=torch.float64)
    y = torch.round(1e6 + 21.3 * i + torch.randint(-3, 4, (n,), device="cuda", generator=g)).to(torch.int64)
    y = y.to(torch.int32)
    del i
    col = OpMaxColumn.encode(y, bs)
    assert torch.equal(col.decode(fixed=True), y)
    plus, minus = col.fix_lists()
    assert torch.equal(col.decode_lists(plus, minus), y)
    idx = torch.randint(0, n, (1_000_000,), device="cuda", generator=g)
    assert torch.equal(col.lookup(idx), y[idx])
    yh = y.cpu().numpy().view(np.uint32)
    st = col.stream.cpu().numpy()
    offs = col.offsets.cpu().numpy().view(np.uint64)
    fx = col.fix.cpu().numpy()
    for b in np.random.default_rng(1).integers(0, n // bs, 40):
        seg = yh[b * bs:(b + 1) * bs]
        oblk, ofix = oracle.opmax_encode_block(seg)
        assert st[int(offs[b]):int(offs[b + 1])].tobytes() == oblk
        assert np.array_equal(fx[b * bs:(b + 1) * bs], ofix)
    plus, minus = col.fix_lists()
    assert len(plus) == int((col.fix > 0).sum()) and len(minus) == int((col.fix < 0).sum())


def test_opmax_cpp_driver_flow():
    # experiments/fix_int.cpp's flow compiled against include/leco_opmax.hpp
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "opmax_driver")
    assert os.path.exists(exe), "built by learn-to-compress_amd/build.py"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok plus=254 minus=289"), r.stdout
