"""The reference's benchmark drivers on the gfx950 library: the same
arguments, input files and output line, so its benchmark scripts and the
parsers of their logs run unchanged against the GPU.

    python -m leco_fix.cli leco_int_fix  <file> <blocks> <model_size> <binary> [filter1 [filter2 base]]
    python -m leco_fix.cli FOR_int_fix   <file> <blocks> <model_size> <binary> [filter1 [filter2 base]]
    python -m leco_fix.cli Delta_int_fix <file> <blocks> <model_size> <binary>
    python -m leco_fix.cli fix_int piecewise_fix_op_max <file> <blocks> <model_size>

leco_int_fix, FOR_int_fix and Delta_int_fix follow
experiments/leco_int_fix_template.cpp, FOR_int_fix_template.cpp and
Delta_int_fix_template.cpp (uint64 values, their leco_type); fix_int follows
experiments/fix_int.cpp for the uint32 codec piecewise_fix_op_max (its other
method names are codecs that this library does not carry).  As there:

- inputs are read from ../data/<file> (--data-dir moves it), SOSD binary when
  <binary> is nonzero, else text (Delta_int_fix reads a text input from <file>
  itself, as its driver does);
- block_size = N / blocks, then blocks = ceil(N / block_size);
- cr = compressed bytes * 100 / raw bytes, cr_model = model_size * blocks * 100
  / raw bytes, cr_wo_model the difference;
- da_ns = time of 10 full decodes / (10 N) in ns (op_max: decode plus the
  fix-up lists, the reference harness's correction), ra_ns = time of N
  lookups at uniform random positions / N, compress_throughput (fix_int) =
  raw bytes / encode seconds / 1e9;
- filter1 alone runs filter_range(filter1) over every block (the drivers skip
  blocks whose zone-map maximum is <= filter1, which hold no match);
  filter1 filter2 base run filter_range_close, for leco_int_fix with each
  block's starting filters from its zone-map minimum; both print the match
  count and the filter's seconds; a time that was not measured repeats the
  last interval measured before it, as the drivers' shared start/end do;
- the sum of the looked-up values (modulo 2^width) is appended to ./fix_log.

Times are wall-clock seconds around device-synchronized work on resident
HBM data, after one untimed warm-up of each operation.  Every decode, lookup
and filter result is checked against the input outside the timed region (the
reference's fix_int and close-filter loops check inside theirs), and a
mismatch prints the drivers' message.  Outputs use C++'s default stream
formatting of doubles (printf %g)."""
from __future__ import annotations

import argparse
import os
import sys
import time

import torch

U32_MAX = (1 << 32) - 1
FIX_METHODS = {"leco_int_fix": "piecewise_fix_op_max", "FOR_int_fix": "FOR_my", "Delta_int_fix": "delta_my"}
REPEAT = 10


def c_atoll(s: str) -> int:
    """atoll: leading whitespace, an optional sign, then digits; 0 without digits."""
    s = s.lstrip(" \t\n\v\f\r")
    sign = 1
    if s[:1] in ("+", "-"):
        sign = -1 if s[0] == "-" else 1
        s = s[1:]
    digits = ""
    for ch in s:
        if not ch.isdigit() or not ch.isascii():
            break
        digits += ch
    return sign * int(digits) if digits else 0


def c_atoi(s: str) -> int:
    """atoi into a 32-bit int (the drivers' int arguments)."""
    v = c_atoll(s) & U32_MAX
    return v - (1 << 32) if v >= (1 << 31) else v


def cfmt(x) -> str:
    """std::cout << x: integers as they are, doubles as %g."""
    if isinstance(x, float):
        return "%g" % x
    return str(x)


def _timed(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def _host(t: torch.Tensor, width: int):
    import numpy as np
    a = t.cpu().numpy()
    return a.view(np.uint32 if width == 32 else np.uint64)


def _report_decode_mismatch(data, rec, width, what, out):
    j = int(torch.nonzero(data != rec)[0, 0].item())
    print(f" num: {j} true is: {int(_host(data[j:j + 1], width)[0])} predict is: "
          f"{int(_host(rec[j:j + 1], width)[0])}", file=out)
    print(f"something wrong! {what}", file=out)


def _report_lookup_mismatch(idx, data, got, width, what, out):
    exp = data[idx]
    k = int(torch.nonzero(exp != got)[0, 0].item())
    print(f"num: {int(idx[k].item())}true is: {int(_host(exp[k:k + 1], width)[0])} predict is: "
          f"{int(_host(got[k:k + 1], width)[0])}", file=out)
    print(f"something wrong! {what}", file=out)


def _partition(n: int, blocks: int) -> tuple[int, int]:
    """experiments/leco_int_fix_template.cpp:99-104: block_size = N / blocks, blocks = ceil(N / block_size)."""
    if blocks <= 0 or n // blocks == 0:
        raise SystemExit(f"cannot split {n} values into {blocks} blocks")
    bs = n // blocks
    return bs, (n + bs - 1) // bs


def _ratios(totalsize: int, model_size: int, blocks: int, raw: float):
    total_model = float(model_size * blocks)
    return totalsize * 100.0 / raw, total_model * 100.0 / raw, (totalsize - total_model) * 100.0 / raw


def _random_positions(n: int, seed: int) -> torch.Tensor:
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(0, n, (n,), device="cuda", generator=g, dtype=torch.int64)


def _append_fix_log(path: str | None, got: torch.Tensor, width: int) -> None:
    if not path:
        return
    mark = int(got.to(torch.int64).sum().item()) & ((1 << width) - 1)
    with open(path, "a") as f:
        f.write(f"{mark}\n")


MIN64 = -(1 << 63)


def _ukey(t: torch.Tensor, width: int) -> torch.Tensor:
    """int64 keys whose signed order is the unsigned order of the values."""
    v = t.to(torch.int64)
    return v & U32_MAX if width == 32 else v ^ MIN64


def _block_min(data: torch.Tensor, bs: int, width: int) -> torch.Tensor:
    """Each block's minimum (the drivers' zone map), unsigned bits in int64."""
    n = data.numel()
    nb = (n + bs - 1) // bs
    key = _ukey(data, width)
    key = torch.cat([key, key[-1:].expand(nb * bs - n)]).view(nb, bs).amin(1)
    return key if width == 32 else key ^ MIN64


def _close_truth(data: torch.Tensor, width: int, filter1: int, filter2: int, base: int) -> torch.Tensor:
    """v % base in (filter1, filter2) in unsigned arithmetic: the close drivers' own check."""
    v = data.to(torch.int64)
    if width == 32:
        r = (v & U32_MAX) % base
    else:
        # v < 0 holds 2^64 + v: add 2^63 mod base to the remainder of the low 63 bits
        r = torch.remainder(v & ((1 << 63) - 1), base)
        c = (1 << 63) % base
        r = torch.where(v < 0, torch.where(r >= base - c, r - (base - c), r + c), r)
    top = (1 << 63) - 1                                        # r < base < 2^31: larger filters act alike
    return (r > min(filter1, top)) & (r < min(filter2, top))


def run_fix_driver(prog: str, argv: list[str], data_dir: str, width: int, seed: int, fix_log: str | None,
                   out=sys.stdout) -> str:
    from . import io
    from .baselines import BaseColumn
    from .column import Column, close_filters_from_zonemap

    if len(argv) < 4:
        raise SystemExit(f"usage: {prog} <file> <blocks> <model_size> <binary> [filter1 [filter2 base]]")
    method = FIX_METHODS[prog]
    source_file = argv[0]
    blocks = c_atoi(argv[1])
    model_size = c_atoi(argv[2])
    binary = c_atoi(argv[3]) != 0
    mask = (1 << width) - 1
    filter_exp = close_exp = False
    filter1 = filter2 = base = 0
    if prog != "Delta_int_fix":
        if len(argv) > 4:
            filter1 = c_atoll(argv[4]) & mask
            filter_exp = True
        if len(argv) > 5:
            if len(argv) < 7:
                raise SystemExit(f"{prog}: filter2 needs a base")
            filter2 = c_atoll(argv[5]) & mask
            base = c_atoll(argv[6]) & mask
            filter_exp, close_exp = False, True
            if not 0 < base < (1 << 31):
                raise SystemExit(f"{prog}: base must be in [1, 2^31) (the codecs take it as int)")

    if binary:
        path = os.path.join(data_dir, source_file)
        data = io.load_sosd(path, width)
    else:
        path = source_file if prog == "Delta_int_fix" else os.path.join(data_dir, source_file)
        data = io.load_text(path, width)
    n = data.numel()
    bs, blocks = _partition(n, blocks)

    if prog == "leco_int_fix":
        col = Column.encode(data, bs)
    else:
        col = BaseColumn.encode("for" if prog == "FOR_int_fix" else "delta", data, bs)
    totalsize = col.nbytes
    compressrate, cr_model, cr_wo_model = _ratios(totalsize, model_size, blocks, float(width // 8 * n))

    recover = torch.empty_like(data)
    col.decode(out=recover)                                    # warm-up
    _, t = _timed(lambda: [col.decode(out=recover) for _ in range(REPEAT)])
    da_ns = t / (n * REPEAT) * 1e9
    if not torch.equal(recover, data):
        _report_decode_mismatch(data, recover, width, "decompress all failed", out)

    idx = _random_positions(n, seed)
    got = torch.empty_like(data)
    col.lookup(idx[:1024], out=got[:1024])                     # warm-up
    _, t_ra = _timed(lambda: col.lookup(idx, out=got))
    ra_ns = t_ra / n * 1e9
    if not torch.equal(got, data[idx]):
        _report_lookup_mismatch(idx, data, got, width, "random access failed", out)
    _append_fix_log(fix_log, got, width)
    del idx, got, recover

    last = t_ra                                                # the drivers' shared start/end
    total_counter = 0
    if filter_exp:
        col.filter_range(filter1)                              # warm-up
        pos, last = _timed(lambda: col.filter_range(filter1))
        total_counter = int(pos.numel())
    filter_totaltime = last
    if close_exp:
        if prog == "leco_int_fix":
            f1, f2 = close_filters_from_zonemap(_block_min(data, bs, width), filter1, filter2, base, width)
            run = lambda: col.filter_range_close(f1, f2, base)  # noqa: E731
        else:
            run = lambda: col.filter_range_close(filter1, filter2, base)  # noqa: E731
        run()                                                  # warm-up
        pos, last = _timed(run)
        total_counter = int(pos.numel())
        # the drivers' per-block check of the count against v % base
        nb = (n + bs - 1) // bs
        want = torch.bincount(torch.nonzero(_close_truth(data, width, filter1, filter2, base))[:, 0] // bs,
                              minlength=nb)
        have = torch.bincount((pos.to(torch.int64) & U32_MAX) // bs, minlength=nb)
        for b in torch.nonzero(want != have)[:, 0].tolist():
            print(f"wrong{b} {int(want[b])} {int(have[b])}", file=out)
    filter_close_totaltime = last

    fields = [method, source_file, blocks, compressrate, cr_model, cr_wo_model, da_ns, ra_ns, 0]
    if prog != "Delta_int_fix":
        fields += [total_counter, filter_totaltime, filter_close_totaltime]
    line = " ".join(cfmt(x) for x in fields)
    print(line, file=out, flush=True)
    return line


def run_fix_int(argv: list[str], data_dir: str, seed: int, fix_log: str | None, out=sys.stdout) -> str:
    """experiments/fix_int.cpp for piecewise_fix_op_max (uint32 text input)."""
    from . import io
    from .opmax import OpMaxColumn

    if len(argv) < 4:
        raise SystemExit("usage: fix_int <method> <file> <blocks> <model_size>")
    method, source_file = argv[0], argv[1]
    if method != "piecewise_fix_op_max":
        raise SystemExit(f"name {method} does not refer to a CODEC of this library (piecewise_fix_op_max)")
    blocks = c_atoi(argv[2])
    model_size = c_atoi(argv[3])
    data = io.load_text(os.path.join(data_dir, source_file), 32)
    n = data.numel()
    if n == 0:
        print("Empty vector", file=out)
        return "Empty vector"
    bs, blocks = _partition(n, blocks)

    def encode():
        c = OpMaxColumn.encode(data, bs)
        return c, c.fix_lists()

    encode()                                                   # warm-up
    (col, (plus, minus)), compress_time = _timed(encode)
    compress_throughput = n * 4 / (compress_time * 1e9)
    compressrate, cr_model, cr_wo_model = _ratios(col.nbytes, model_size, blocks, float(4 * n))

    recover = torch.empty_like(data)
    col.decode_lists(plus, minus, out=recover)                 # warm-up
    _, t = _timed(lambda: [col.decode_lists(plus, minus, out=recover) for _ in range(REPEAT)])
    da_ns = t / (REPEAT * n) * 1e9
    if not torch.equal(recover, data):
        _report_decode_mismatch(data, recover, 32, "decompress failed", out)

    idx = _random_positions(n, seed)
    got = torch.empty_like(data)
    col.lookup(idx[:1024], out=got[:1024])                     # warm-up
    _, t_ra = _timed(lambda: col.lookup(idx, out=got))
    ra_ns = t_ra / n * 1e9
    if not torch.equal(got, data[idx]):
        _report_lookup_mismatch(idx, data, got, 32, "decompress failed", out)
    _append_fix_log(fix_log, got, 32)

    fields = [method, source_file, blocks, compressrate, cr_model, cr_wo_model, da_ns, ra_ns, compress_throughput]
    line = " ".join(cfmt(x) for x in fields)
    print(line, file=out, flush=True)
    return line


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m leco_fix.cli", description=__doc__.split("\n\n")[0],
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--data-dir", default=os.path.join("..", "data"), help="directory of the input files (../data)")
    ap.add_argument("--width", type=int, choices=(32, 64), default=64,
                    help="value width of the _fix drivers (their leco_type: 64)")
    ap.add_argument("--seed", type=int, default=1, help="seed of the random access positions")
    ap.add_argument("--fix-log", default="fix_log", help="file the lookup checksum is appended to ('' for none)")
    ap.add_argument("driver", choices=sorted(FIX_METHODS) + ["fix_int"])
    ap.add_argument("args", nargs=argparse.REMAINDER)
    a = ap.parse_args(argv)
    if not torch.cuda.is_available():
        raise SystemExit("the drivers run on the GPU library: no GPU is visible")
    if a.driver == "fix_int":
        run_fix_int(a.args, a.data_dir, a.seed, a.fix_log or None)
    else:
        run_fix_driver(a.driver, a.args, a.data_dir, a.width, a.seed, a.fix_log or None)
    return 0


if __name__ == "__main__":
    sys.exit(main())
