"""CPU: the drivers' host-side arithmetic in leco_fix.cli -- argument parsing
against libc's own atoi / atoll, output formatting against libc's printf %g
(what C++'s default double output prints), the partition and ratio formulas of
experiments/leco_int_fix_template.cpp:99-147 and experiments/fix_int.cpp, the
zone map and the close drivers' v % base check on CPU tensors against Python
integers, and a clean refusal without a GPU."""
import ctypes

import numpy as np
import pytest
import torch

from leco_fix import cli

LIBC = ctypes.CDLL(None)
LIBC.atoll.restype = ctypes.c_longlong
LIBC.atoll.argtypes = [ctypes.c_char_p]
LIBC.atoi.restype = ctypes.c_int
LIBC.atoi.argtypes = [ctypes.c_char_p]
LIBC.snprintf.restype = ctypes.c_int

ARGS = ["0", "12", "-7", "+9", "  42", "\t-3x", "17abc", "abc", "", "-", "+-1", "007", "2147483647",
        "2147483648", "-2147483648", "4294967297", "9223372036854775807", "-9223372036854775808",
        "18446744073709551615"]


@pytest.mark.parametrize("s", ARGS)
def test_atoll_atoi_match_libc(s):
    # atoll / atoi have undefined results past the range of their type; compare inside it
    v = cli.c_atoll(s)
    if -(1 << 63) <= v < (1 << 63):
        assert v == LIBC.atoll(s.encode())
    if -(1 << 31) <= v < (1 << 31):
        assert cli.c_atoi(s) == LIBC.atoi(s.encode())


def test_cfmt_matches_printf_g():
    rng = np.random.default_rng(4)
    xs = [0.0, 1.0, 100.0, 123456.0, 1234567.0, 0.0001, 0.00001, 12.5, 2.0 / 3.0, 1e300, 37.4999995, 1.0 / 7e9,
          float("inf")] + (10.0 ** rng.uniform(-12, 12, 300)).tolist()
    buf = ctypes.create_string_buffer(64)
    for x in xs:
        LIBC.snprintf(buf, 64, b"%g", ctypes.c_double(x))
        assert cli.cfmt(x) == buf.value.decode(), x
    assert cli.cfmt(7) == "7" and cli.cfmt(0) == "0"


def test_partition_and_ratios():
    assert cli._partition(200_000_000, 100_000) == (2000, 100_000)
    assert cli._partition(1003, 10) == (100, 11)          # ragged tail: one more block
    assert cli._partition(7, 7) == (1, 7)
    for bad in [(5, 6), (10, 0), (10, -1)]:
        with pytest.raises(SystemExit):
            cli._partition(*bad)
    cr, crm, crw = cli._ratios(1000, 12, 11, 8.0 * 1003)
    assert cr == 1000 * 100.0 / 8024.0 and crm == 132 * 100.0 / 8024.0 and crw == (1000 - 132.0) * 100.0 / 8024.0


@pytest.mark.parametrize("width", [32, 64])
def test_block_min_and_close_truth(width):
    rng = np.random.default_rng(width)
    n, bs = 10_007, 100
    v = rng.integers(0, 1 << width, n, dtype=np.uint64) if width == 64 else rng.integers(0, 1 << 32, n).astype(
        np.uint64)
    v[:50] = (1 << width) - 1
    st = torch.from_numpy(v.view(np.int64).copy()) if width == 64 else torch.from_numpy(
        v.astype(np.uint32).view(np.int32).copy())
    mins = cli._block_min(st, bs, width).numpy()
    exp = [int(v[b:b + bs].min()) for b in range(0, n, bs)]
    got = [int(x) & ((1 << 64) - 1) for x in mins.tolist()]
    assert got == exp
    for base, f1, f2 in [(1000, 100, 900), (7, 0, 6), ((1 << 31) - 1, 5, 1 << 30), (3, 1 << 40, 1 << 41)]:
        truth = cli._close_truth(st, width, f1, f2, base).numpy()
        exp = np.array([f1 < (int(x) % base) < f2 for x in v.tolist()])
        assert np.array_equal(truth, exp), (base, f1, f2)


def test_main_refuses_without_gpu(monkeypatch):
    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    with pytest.raises(SystemExit, match="no GPU"):
        cli.main(["leco_int_fix", "x.bin", "10", "12", "1"])
