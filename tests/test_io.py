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
