"""CPU checks of the C ABI boundary: the library loads, exports every function
include/leco_fix.h declares, its host-only size functions agree with the
format, and argument errors come back as codes without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "leco_fix.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(leco_fix_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from leco_fix import _lib
    L = _lib.lib()
    names = declared_functions()
    assert names == sorted(_lib.SIGNATURES)
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_param_struct_layout():
    from leco_fix import BlockParam
    assert ctypes.sizeof(BlockParam) == 32
    assert [f for f, _ in BlockParam._fields_] == ["theta0", "theta1", "payload_offset", "bits", "length"]


@pytest.mark.parametrize("width", [32, 64])
def test_size_functions_match_format(oracle_mod, width):
    from leco_fix import lib
    L = lib()
    for n in [1, 2, 7, 100, 4096, 10000]:
        for l in range(0, width + 1):
            assert L.leco_fix_block_bytes(l, width, n) == oracle_mod.block_bytes(l, width, n)
    for n, bs in [(1, 1), (10, 3), (500_000_000, 4096), (10_000_000, 10_000), (12345, 16384)]:
        assert L.leco_fix_block_count(n, bs) == -(-n // bs)
        assert L.leco_fix_stream_bound(n, bs, width) == oracle_mod.stream_bound(n, bs, width) + 64


def test_argument_errors_are_codes():
    from leco_fix import lib
    L = lib()
    EINVAL = 1
    buf = (ctypes.c_uint8 * 64)()
    assert L.leco_fix_encode_column_u64(None, 0, 4096, None, 0, None, None, None) == EINVAL
    assert L.leco_fix_encode_column_u32(buf, 10, 0, buf, 64, buf, buf, None) == EINVAL
    assert L.leco_fix_encode_column_u64(buf, 10, 1 << 20, buf, 64, buf, buf, None) == EINVAL
    # capacity below the bound
    assert L.leco_fix_encode_column_u64(buf, 10, 4, buf, 64, buf, buf, None) == EINVAL
    assert L.leco_fix_decode_range_u64(buf, buf, 10, 4, 5, 5, buf, None) == EINVAL
    assert L.leco_fix_decode_range_u64(buf, buf, 10, 4, 0, 11, buf, None) == EINVAL
    assert L.leco_fix_lookup_u32(buf, buf, 10, 4, buf, 0, buf, None) == 0  # nothing to do
    assert L.leco_fix_build_params(buf, buf, 10, 4, 48, buf, None) == EINVAL
    assert L.leco_fix_strerror(EINVAL) == b"invalid argument"


def test_cpp_driver_is_built():
    # the C++ Leco_int<T> drop-in header compiles and links against the library
    assert os.access(os.path.join(ROOT, "tests", "cpp", "leco_int_driver"), os.X_OK)


def test_host_mirror_surface():
    from leco_fix import Leco_int
    for name in ["init", "encodeArray8_int", "encodeArray8", "decodeArray8", "randomdecodeArray8",
                 "get_block_nums"]:
        assert callable(getattr(Leco_int, name))
    c = Leco_int(np.uint32)
    c.init(7, 1000)
    assert c.get_block_nums() == 7 and c.block_size == 1000
    with pytest.raises(TypeError):
        Leco_int(np.int64)


def test_host_allocation_failure_is_an_error_code(tmp_path):
    """A host allocation that fails inside an entry point comes back as
    LECO_FIX_ENOMEM instead of an exception unwinding into the C caller (which
    would abort it).  leco_fix_text_info is host only and allocates a 16 MB
    read buffer; a child process caps its address space just above what it
    already uses, then calls it."""
    import subprocess
    import sys
    p = tmp_path / "v.txt"
    p.write_text("1 2 3\n")
    code = f"""
import ctypes, resource, sys
sys.path.insert(0, {os.path.join(ROOT, "learn-to-compress_amd")!r})
from leco_fix import _lib
L = _lib.lib()
cnt = ctypes.c_uint64(0)
assert L.leco_fix_text_info({str(p)!r}.encode(), 64, ctypes.byref(cnt)) == 0 and cnt.value == 3
used = 0
for line in open("/proc/self/status"):
    if line.startswith("VmSize:"):
        used = int(line.split()[1]) * 1024
resource.setrlimit(resource.RLIMIT_AS, (used + (4 << 20), resource.RLIM_INFINITY))
rc = L.leco_fix_text_info({str(p)!r}.encode(), 64, ctypes.byref(cnt))
print("rc", rc)
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "rc 3" in r.stdout, r.stdout


def test_grouped_lookup_workspace_and_argument_checks():
    # no device memory is touched: every call below is refused before any launch
    from leco_fix._lib import LECO_FIX_EINVAL, lib
    L = lib()
    head = 2 * 4096 * 4
    for nq in (0, 1, 1000, 1 << 31, (1 << 31) + 5, 1 << 40):
        assert L.leco_fix_lookup_grouped_workspace(nq) == head + 8 * min(nq, 1 << 31)
    fake = 1 << 20  # an address the checks reject the call before using
    ws = L.leco_fix_lookup_grouped_workspace(100)
    args = (fake, fake, 1000, 16, fake, 100, fake)
    assert L.leco_fix_lookup_grouped_u64(*args, fake, ws - 1, None) == LECO_FIX_EINVAL   # short workspace
    assert L.leco_fix_lookup_grouped_u64(*args, fake + 4, ws, None) == LECO_FIX_EINVAL   # misaligned workspace
    assert L.leco_fix_lookup_grouped_u64(fake, fake, (1 << 33) + 1, 16, fake, 100, fake, fake, ws, None) == \
        LECO_FIX_EINVAL                                                                   # indices past 33 bits
    assert L.leco_fix_lookup_grouped_u32(fake, fake, 1000, 0, fake, 100, fake, fake, ws, None) == LECO_FIX_EINVAL
    assert L.leco_fix_lookup_grouped_u32(fake, fake, 1000, 16, fake, 0, fake, fake, 0, None) == 0  # empty batch
