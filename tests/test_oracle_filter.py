"""CPU checks that pin the oracle's filter_range / filter_range_close
(Leco_int<T>, headers/piecewise_fix_integer_template.h:206-303) before it
checks the GPU filters.

- filter_range on non-decreasing data equals the brute-force set {i : y_i >
  filter} (the model threshold only skips values that cannot pass), for
  every width, block size and distribution below;
- filter_range_close with the reference driver's zonemap-derived per-block
  filters gives the driver's own expected count (values with y % base in
  (filter1, filter2), experiments/leco_int_fix_template.cpp:299-310) and,
  period after period, exactly those positions;
- the documented behaviours: uint64 threshold wrap when filter + 1 < 2^(l-1),
  raw blocks filtered by value, every reported position passes the predicate
  on unsorted data;
- per-block calls agree with the column calls.
"""
import struct

import numpy as np
import pytest
import torch

import spec_leco as spec

from leco_fix import datagen
from leco_fix.column import close_filters_from_zonemap

WIDTHS = (32, 64)
SORTED = {
    "sorted_uniform": lambda n, w, s: datagen.sorted_uniform(n, w, s),
    "zipf_gaps": lambda n, w, s: datagen.sorted_zipf_gaps(n, w, s),
    "stepwise": lambda n, w, s: datagen.stepwise(n, w, s),
}


def _col(oracle_mod, vals, bs):
    st, off = oracle_mod.encode_column(vals, bs)
    return st, off


def _bits_of(st, off):
    return np.array([st[int(o)] for o in off[:-1]], np.int64)


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("bs", [64, 1000, 4096])
@pytest.mark.parametrize("dist", list(SORTED))
def test_filter_range_sorted_equals_brute_force(oracle_mod, width, bs, dist):
    vals = SORTED[dist](30_000, width, bs + width)
    st, off = _col(oracle_mod, vals, bs)
    bits = _bits_of(st, off)
    qs = [0.001, 0.1, 0.37, 0.5, 0.9, 0.999]
    filters = [int(np.quantile(vals.astype(np.float64), q)) for q in qs] + [int(vals.max())]
    for f in filters:
        if width == 64 and f + 1 < (1 << max(int(bits.max()) - 1, 0)):
            continue  # the uint64 threshold wrap, pinned separately below
        got = oracle_mod.filter_column(st, off, len(vals), bs, vals.dtype, f)
        want = np.nonzero(vals.astype(np.uint64) > np.uint64(f))[0].astype(np.uint32)
        assert np.array_equal(got, want), (dist, width, bs, f)


def test_filter_range_u32_small_filters_scan_everything(oracle_mod):
    # 32-bit columns compute the threshold in int64: filter + 1 - 2^(l-1) < 0
    # just starts at position 0
    vals = datagen.linear_sigma(20_000, 32, 3)
    st, off = _col(oracle_mod, vals, 500)
    for f in (0, 1, 7):
        got = oracle_mod.filter_column(st, off, len(vals), 500, vals.dtype, f)
        want = np.nonzero(vals > f)[0].astype(np.uint32)
        assert np.array_equal(got, want)


def test_filter_range_u64_threshold_wrap(oracle_mod):
    # T = uint64: filter + 1 - 2^(l-1) wraps, the threshold lands past the
    # block (INT_MIN after the int conversion) and the block reports nothing;
    # blocks with 2^(l-1) <= filter + 1 behave normally
    vals = datagen.linear_sigma(20_000, 64, 4)
    bs = 1000
    st, off = _col(oracle_mod, vals, bs)
    bits = _bits_of(st, off)
    assert (bits[bits < 64] >= 3).any()
    f = 2
    got = oracle_mod.filter_column(st, off, len(vals), bs, vals.dtype, f)
    want = []
    for b in range(len(bits)):
        if bits[b] < 64 and f + 1 < (1 << max(int(bits[b]) - 1, 0)):
            continue  # raw blocks (bits 64) are filtered by value
        seg = vals[b * bs:(b + 1) * bs]
        want.extend((b * bs + np.nonzero(seg > np.uint64(f))[0]).tolist())
    assert got.tolist() == want


def _brute_close(vals, bs, f1, f2, base, st=None, off=None):
    """Per block, period after period from (f1[b], f2[b]): positions whose value lies in the period.
    Sorted data: a value is in at most one period and periods follow the data upward.  With the stream
    given, a block of bit width 0 also has the virtual position m (value p(m) = cvtt64(fma(t1, m, t0))):
    the reference's l == 0 loop runs to i <= thre2 (:285)."""
    out = []
    nb = (len(vals) + bs - 1) // bs
    for b in range(nb):
        seg = [int(v) for v in vals[b * bs:(b + 1) * bs]]
        if st is not None and st[int(off[b])] == 0:
            t0, t1 = struct.unpack("<dd", st[int(off[b]) + 1:int(off[b]) + 17].tobytes())
            seg.append(spec.predict(t0, t1, len(seg)))
        lo, hi, top = int(f1[b]), int(f2[b]), max(seg)
        while lo < top:
            out.extend(b * bs + i for i, v in enumerate(seg) if lo < v < hi)
            lo += base
            hi += base
    return out


def _window(vals, width):
    """Non-decreasing data moved into [lo, lo + span): block minimums stay above every filter1, f + k*base
    never wraps and the driver's float64 filter arithmetic is exact (below 2^53) -- the conditions under
    which the driver's own count check holds (its zonemap formula wraps for minimums below filter1,
    T += base wraps at the top, and doubles round above 2^53)."""
    lo, span = (1 << 30, 1 << 29) if width == 32 else (1 << 44, 1 << 36)
    v = vals.astype(np.float64)
    u = (v - v.min()) / max(v.max() - v.min(), 1.0)
    return (np.uint64(lo) + np.floor(u * (span - 1)).astype(np.uint64)).astype(vals.dtype)


CLOSE_CASES = [(d, bs) for d in SORTED for bs in (100, 1000, 4096) if not (d == "stepwise" and bs == 100)]


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("dist,bs", CLOSE_CASES)
def test_filter_close_driver_check(oracle_mod, width, dist, bs):
    # (stepwise data at block size 100 is mostly constant blocks, whose fitted
    # slope can round slightly negative; the reference's threshold then turns
    # into INT_MIN and the block reports nothing, so its own check fails
    # there -- the exact-constant case is pinned below)
    vals = _window(SORTED[dist](30_000, width, 7 * bs + width), width)
    st, off = _col(oracle_mod, vals, bs)
    nb = len(off) - 1
    mins = np.array([vals[b * bs:(b + 1) * bs].min() for b in range(nb)], np.uint64)
    span = float(vals.max()) - float(vals.min())
    for base in (int(span / 3) + 1, int(span / 50) + 1, int(span / 2000) + 7):
        base = min(base, (1 << 31) - 1)
        f1, f2 = base // 5, base // 5 + base // 2
        t1, t2 = close_filters_from_zonemap(torch.from_numpy(mins.view(np.int64)), f1, f2, base, width)
        t1 = t1.numpy().astype(np.int64).astype(np.uint64) & np.uint64((1 << width) - 1)
        t2 = t2.numpy().astype(np.int64).astype(np.uint64) & np.uint64((1 << width) - 1)
        got = oracle_mod.filter_close_column(st, off, len(vals), bs, vals.dtype, t1, t2, base)
        assert got.tolist() == _brute_close(vals, bs, t1, t2, base, st, off), (dist, width, bs, base)
        v = vals.astype(np.uint64)
        truecount = int(np.count_nonzero(((v % np.uint64(base)) > f1) & ((v % np.uint64(base)) < f2)))
        if not (_bits_of(st, off) == 0).any():
            assert len(got) == truecount, (dist, width, bs, base)  # the driver's own check


@pytest.mark.parametrize("width", WIDTHS)
def test_filter_close_exact_constant_blocks(oracle_mod, width):
    # Constant blocks of small values fit exactly: theta1 = +0, theta0 = c, l = 0.
    # The close loop's window is then [0, m] inclusive (:285): a passing
    # constant reports positions 0..m of its block, m being the next block's
    # first position; a failing one reports nothing.
    bs, nb, base = 250, 12, 1000
    consts = [5000 + 137 * k for k in range(nb)]
    vals = np.repeat(np.array(consts, np.uint64), bs).astype(np.uint32 if width == 32 else np.uint64)
    st, off = _col(oracle_mod, vals, bs)
    assert (_bits_of(st, off) == 0).all()
    for b in range(nb):
        t0, t1 = struct.unpack("<dd", st[int(off[b]) + 1:int(off[b]) + 17].tobytes())
        assert (t0, t1) == (float(consts[b]), 0.0) and not np.signbit(t1)
    f1, f2 = 200, 700
    mins = torch.tensor(consts, dtype=torch.int64)
    t1s, t2s = close_filters_from_zonemap(mins, f1, f2, base, width)
    t1s = t1s.numpy().astype(np.int64).astype(np.uint64)
    t2s = t2s.numpy().astype(np.int64).astype(np.uint64)
    got = oracle_mod.filter_close_column(st, off, len(vals), bs, vals.dtype, t1s, t2s, base).tolist()
    want = []
    for b, c in enumerate(consts):
        if f1 < c % base < f2:
            want.extend(range(b * bs, b * bs + bs + 1))
    assert got == want


@pytest.mark.parametrize("width", WIDTHS)
def test_filter_raw_blocks_by_value(oracle_mod, width):
    rng = np.random.default_rng(width)
    top = (1 << width) - 1
    vals = rng.integers(0, top, 5000, dtype=np.uint64, endpoint=True).astype(np.uint32 if width == 32 else np.uint64)
    bs = 500
    st, off = _col(oracle_mod, vals, bs)
    assert (_bits_of(st, off) == width).all()
    for f in (0, top // 3, top - 5):
        got = oracle_mod.filter_column(st, off, len(vals), bs, vals.dtype, f)
        assert np.array_equal(got, np.nonzero(vals.astype(np.uint64) > np.uint64(f))[0].astype(np.uint32))
    base = min(top // 7, (1 << 31) - 1)  # the reference's `int base`
    f1, f2 = base // 4, base // 2
    nb = len(off) - 1
    got = oracle_mod.filter_close_column(st, off, len(vals), bs, vals.dtype, np.full(nb, f1, np.uint64),
                                         np.full(nb, f2, np.uint64), base)
    want = [i for i, v in enumerate(vals.tolist()) if v > f1 and 0 < (v - f1 - 1) % base + 1 < f2 - f1]
    assert got.tolist() == want


def test_filter_unsorted_reports_only_passing_positions(oracle_mod):
    rng = np.random.default_rng(9)
    for width in WIDTHS:
        vals = datagen.linear_sigma(20_000, width, 5)
        vals = vals.copy()
        rng.shuffle(vals[:7000])
        st, off = _col(oracle_mod, vals, 700)
        f = int(np.median(vals.astype(np.float64)))
        got = oracle_mod.filter_column(st, off, len(vals), 700, vals.dtype, f)
        assert (vals[got] > f).all()
        assert np.all(np.diff(got.astype(np.int64)) > 0)
        # past the shuffled prefix the data is sorted again: nothing is missed there
        tail = got[got >= 7000]
        assert np.array_equal(tail, 7000 + np.nonzero(vals[7000:] > f)[0].astype(np.uint32))


def test_filter_block_calls_match_column(oracle_mod):
    vals = datagen.sorted_uniform(9_000, 64, 12)
    bs = 1000
    st, off = _col(oracle_mod, vals, bs)
    f = int(vals[4321])
    col = oracle_mod.filter_column(st, off, len(vals), bs, vals.dtype, f)
    per = []
    for b in range(len(off) - 1):
        blk = st[int(off[b]):int(off[b + 1])].tobytes()
        m = min(bs, len(vals) - b * bs)
        per.extend(oracle_mod.filter_block(blk, m, vals.dtype, f, b * bs).tolist())
    assert col.tolist() == per
