"""GPU: SOSD files into HBM and column files through save / load, against
numpy-written files, the original column and the oracle."""
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


def to_dev(torch, a):
    st = np.int32 if a.dtype == np.uint32 else np.int64
    return torch.from_numpy(np.ascontiguousarray(a).view(st)).to("cuda")


@pytest.mark.parametrize("width", [32, 64])
def test_sosd_load_save_round_trip(tmp_path, oracle_mod, width):
    import torch
    from leco_fix import Column, datagen, io
    # 20M values: more than one 64 MB staging piece for both widths
    n = 20_000_000
    vals = datagen.sorted_uniform(n, width, 5)
    p = tmp_path / "keys.bin"
    p.write_bytes(struct.pack("<Q", n) + vals.tobytes())
    say(f"sosd load width={width} n={n}")
    dev = io.load_sosd(p, width)
    assert dev.numel() == n
    assert np.array_equal(dev.cpu().numpy().view(vals.dtype), vals)
    q = tmp_path / "again.bin"
    io.save_sosd(q, dev)
    assert q.read_bytes() == p.read_bytes()
    # the reference's flow: load, encode, compare with the oracle
    small = vals[:300_000]
    p2 = tmp_path / "small.bin"
    p2.write_bytes(struct.pack("<Q", len(small)) + small.tobytes())
    col = Column.encode(io.load_sosd(p2, width), 1000)
    st, _ = oracle_mod.encode_column(small, 1000, threads=8)
    assert col.stream_bytes() == st.tobytes()


@pytest.mark.parametrize("width,bs", [(64, 4096), (32, 256), (64, 7)])
def test_column_file_round_trip(tmp_path, width, bs):
    import torch
    from leco_fix import Column, datagen, io
    from leco_fix.column import PARAM_BYTES
    n = 9_000_000 if bs == 4096 else 600_000
    vals = datagen.sorted_uniform(n, width, bs) if bs != 7 else datagen.stepwise(n, width, 3)
    col = Column.encode(to_dev(torch, vals), bs)
    p = tmp_path / "col.leco"
    say(f"column file width={width} bs={bs} n={n}")
    io.save_column(p, col)
    info = io.column_file_info(p)
    assert info == {"n": n, "num_blocks": col.num_blocks, "stream_bytes": col.nbytes, "block_size": bs,
                    "width": width}
    back = io.load_column(p)
    assert back.stream_bytes() == col.stream_bytes()
    assert torch.equal(back.offsets, col.offsets)
    assert torch.equal(back.params, col.params[:col.num_blocks * PARAM_BYTES])
    assert np.array_equal(back.decode().cpu().numpy().view(vals.dtype), vals)
    idx = torch.randint(0, n, (100_000,), device="cuda")
    assert torch.equal(back.lookup(idx), col.lookup(idx))


def test_column_file_from_batch_arena(tmp_path):
    import torch
    from leco_fix import datagen, io
    from leco_fix.batch import Batch
    cols = [datagen.sorted_uniform(50_000, 64, 1), datagen.stepwise(70_000, 32, 2)]
    b = Batch([to_dev(torch, v) for v in cols], [1000, 512])
    b.encode()
    torch.cuda.synchronize()
    for c, v in enumerate(cols):
        col = b.column(c)
        p = tmp_path / f"c{c}.leco"
        say(f"column file from batch arena column {c} (base {col.base})")
        io.save_column(p, col)
        back = io.load_column(p)
        assert back.stream_bytes() == b.column_bytes(c)
        assert np.array_equal(back.decode().cpu().numpy().view(v.dtype), v)


def test_column_load_rejects_small_buffers(tmp_path):
    import torch
    from leco_fix import Column, datagen, io
    from leco_fix._lib import LecoError, lib
    vals = datagen.sorted_uniform(100_000, 64, 9)
    col = Column.encode(to_dev(torch, vals), 1000)
    p = tmp_path / "col.leco"
    io.save_column(p, col)
    st = torch.empty(col.nbytes, dtype=torch.uint8, device="cuda")      # no room for the tail pad
    offs = torch.empty(col.num_blocks + 1, dtype=torch.int64, device="cuda")
    rc = lib().leco_fix_column_load(bytes(p), st.data_ptr(), st.numel(), offs.data_ptr(), None, None)
    assert rc == 1
    with pytest.raises(LecoError):
        io.load_sosd(tmp_path / "missing.bin", 64)


@pytest.mark.parametrize("width,bs", [(64, 7), (64, 4096), (32, 1000)])
def test_column_load_checks_header_bytes_against_spans(tmp_path, oracle_mod, width, bs):
    """A block whose header byte does not give exactly its span (a flipped
    width byte, or offsets shifted by a multiple of 8 so that every span is
    still a possible size) is refused while the stream is copied, before the
    param table is built; the buffers are not touched by any kernel.  (A
    width that gives the same span, e.g. 6 and 7 bits for 7 values, cannot be
    told apart and is skipped: every read stays inside the block.)  The
    unaltered file loads and decodes."""
    import torch
    from test_io import column_file

    from leco_fix import datagen, io
    from leco_fix._lib import lib
    n = 3000 * bs // 7 + 5 if bs == 7 else 30 * bs + 5
    vals = datagen.sorted_pwl_u64(n, seed=6) if width == 64 else datagen.linear_noise_u32(n, seed=6)
    st, off = oracle_mod.encode_column(vals, bs)
    off = np.asarray(off, np.uint64)
    p = tmp_path / "col.leco"
    dst = torch.zeros(len(st) + 64, dtype=torch.uint8, device="cuda")
    offs = torch.zeros(len(off), dtype=torch.int64, device="cuda")
    prm = torch.zeros(32 * (len(off) - 1), dtype=torch.uint8, device="cuda")

    def load(stream_bytes, offsets):
        p.write_bytes(column_file(n, bs, width, stream_bytes, offsets))
        return lib().leco_fix_column_load(bytes(p), dst.data_ptr(), dst.numel(), offs.data_ptr(), prm.data_ptr(),
                                          None)
    nb = len(off) - 1
    for b in (0, nb // 2, nb - 1):
        for flip in (1, 0x40, 0x7F, 0xFF):
            bad = bytearray(st.tobytes())
            bad[int(off[b])] ^= flip
            l = bad[int(off[b])]
            m = min(bs, n - b * bs)
            if l <= width and oracle_mod.block_bytes(l, width, m) == int(off[b + 1] - off[b]):
                continue  # this width gives the same span: a stream the check cannot tell apart
            say(f"column load: width={width} bs={bs} block {b} header {st[int(off[b])]} -> {l}")
            assert load(bytes(bad), off) == 1, (b, flip)
    if bs == 7:
        # an offset moved by 8: both neighbouring spans are still possible sizes
        moved = off.copy()
        widths = st[off[:-1].astype(np.int64)]
        k = next(i for i in range(1, nb) if 0 < widths[i - 1] < width - 1 and 0 < widths[i] < width - 1)
        moved[k] = np.uint64(int(moved[k]) + 8)
        say(f"column load: offset {k} moved by 8")
        assert load(st.tobytes(), moved) == 1
    assert load(st.tobytes(), off) == 0
    col = io.load_column(p)
    assert np.array_equal(col.decode().cpu().numpy().view(vals.dtype), vals)


@pytest.mark.parametrize("width", [32, 64])
def test_text_load_values(tmp_path, istream_values, width):
    """leco_fix_text_load's values against libstdc++'s extraction loop
    (tests/cpp/istream_values.cpp) on the edge cases, on fuzz and on a file
    of more than one 64 MB staging piece of values."""
    from leco_fix import io
    from text_corpus import edge_cases, fuzz_bytes
    host = np.uint32 if width == 32 else np.uint64
    rng = np.random.default_rng(width + 1)
    say(f"text load edge cases width={width}")
    files = dict(edge_cases(width))
    for k in range(40):
        files[f"fuzz{k}"] = fuzz_bytes(rng, int(rng.integers(0, 400)))
    for name, body in files.items():
        p = tmp_path / f"{name}.txt"
        p.write_bytes(body)
        got = io.load_text(p, width).cpu().numpy().view(host)
        assert got.astype(np.uint64).tolist() == istream_values(p, width).tolist(), name
    n = 17_000_000 if width == 32 else 9_000_000     # 68 / 72 MB of values
    say(f"text load large width={width} n={n}")
    v = rng.integers(0, (1 << width) - 1, n, dtype=np.uint64, endpoint=True)
    p = tmp_path / "large.txt"
    p.write_text("\n".join(map(str, v.tolist())) + "\n")
    got = io.load_text(p, width)
    assert got.numel() == n
    assert np.array_equal(got.cpu().numpy().view(host), v.astype(host))
