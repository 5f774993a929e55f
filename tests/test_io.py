"""CPU: the file formats' host-side parsers (no GPU calls) against files
written here from the documented layouts -- the reference's SOSD binary
(experiments/leco_int_fix_template.cpp:21-40: uint64 count, then the raw
values) and the column file (64-byte header, block offsets, stream)."""
import struct

import numpy as np
import pytest

from leco_fix._lib import LecoError


def fnv1a(b: bytes) -> int:
    h = 0xcbf29ce484222325
    for x in b:
        h = ((h ^ x) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def column_file(n, bs, width, stream: bytes, offsets):
    nb = (n + bs - 1) // bs
    head = b"LECOCOL1" + struct.pack("<IIQIIQQ", 1, width, n, bs, 0, nb, len(stream))
    head += struct.pack("<QQ", fnv1a(head), 0)
    body = np.asarray(offsets, np.uint64).tobytes()
    start = (64 + 8 * (nb + 1) + 63) // 64 * 64
    return head + body + bytes(start - 64 - len(body)) + stream


@pytest.mark.parametrize("width", [32, 64])
def test_sosd_info_reads_the_reference_layout(tmp_path, width):
    from leco_fix import io
    vals = np.arange(1000, dtype=np.uint32 if width == 32 else np.uint64) * 7
    p = tmp_path / "keys"
    p.write_bytes(struct.pack("<Q", len(vals)) + vals.tobytes())
    assert io.sosd_info(p, width) == 1000
    p.write_bytes(struct.pack("<Q", 1001) + vals.tobytes())   # count past the end of the file
    with pytest.raises(LecoError):
        io.sosd_info(p, width)


def test_column_file_info_and_rejections(tmp_path):
    from leco_fix import io
    n, bs = 10_000, 1000
    offsets = [17 * b for b in range(11)]
    good = column_file(n, bs, 64, bytes(170), offsets)
    p = tmp_path / "col"
    p.write_bytes(good)
    assert io.column_file_info(p) == {"n": n, "num_blocks": 10, "stream_bytes": 170, "block_size": bs, "width": 64}
    bad = bytearray(good)
    bad[20] ^= 1                                    # n changed: the header check fails
    p.write_bytes(bytes(bad))
    with pytest.raises(LecoError):
        io.column_file_info(p)
    p.write_bytes(b"LECOCOL2" + good[8:])           # unknown magic
    with pytest.raises(LecoError):
        io.column_file_info(p)
    p.write_bytes(good[:-1])                        # truncated stream
    with pytest.raises(LecoError):
        io.column_file_info(p)
    p.write_bytes(column_file(n, bs, 64, bytes(170), offsets)[:64])
    with pytest.raises(LecoError):
        io.column_file_info(p)


def _load_rc(path, cap):
    """leco_fix_column_load with host stand-ins for the device buffers: a
    file that fails the host-side checks is refused before any HIP call, so
    the stand-ins are never touched."""
    import ctypes
    from leco_fix import lib
    buf = (ctypes.c_uint8 * 64)()
    return lib().leco_fix_column_load(str(path).encode(), buf, cap, buf, None, None)


@pytest.mark.parametrize("width,bs", [(64, 4096), (32, 1000), (64, 7)])
def test_column_load_refuses_spans_no_header_byte_gives(tmp_path, oracle_mod, width, bs):
    """A shifted or altered block offset leaves a block whose span no header
    byte can give (no width 0..W, raw, or the 126/127 forms): refused with
    EINVAL before anything is copied or launched.  (A shift by a multiple of
    8 can leave possible spans; the header-byte check while the stream is
    copied catches those, tests/test_gpu_io.py.)"""
    from leco_fix import datagen
    n = 20 * bs + 3
    vals = datagen.sorted_pwl_u64(n, seed=4) if width == 64 else datagen.linear_noise_u32(n, seed=4)
    st, off = oracle_mod.encode_column(vals, bs)
    off = np.asarray(off, np.uint64)
    p = tmp_path / "col"
    cap = len(st) + 64
    for b, delta in ((3, 5), (7, -1), (1, 1), (19, 3)):
        bad = off.copy()
        bad[b] = np.uint64(int(bad[b]) + delta)
        p.write_bytes(column_file(n, bs, width, st.tobytes(), bad))
        assert _load_rc(p, cap) == 1, (b, delta)
    bad = off.copy()
    bad[5], bad[6] = bad[6], bad[5]               # not monotone
    p.write_bytes(column_file(n, bs, width, st.tobytes(), bad))
    assert _load_rc(p, cap) == 1
