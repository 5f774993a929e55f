"""GPU: the reference's drivers (leco_fix.cli) on small files in the
reference's layout (../data/<file> from the working directory): every field of
the output line against the oracle's compressed sizes and the data itself,
the match counts of both filters on sorted data (where the model threshold
skips nothing), the fix_log checksum, and `python -m leco_fix.cli` as a
process."""
import io as _io
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


@pytest.fixture
def layout(tmp_path, monkeypatch):
    (tmp_path / "data").mkdir()
    run = tmp_path / "run"
    run.mkdir()
    monkeypatch.chdir(run)
    return tmp_path


def sorted_u64(n, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.integers(0, 1 << 44, n, dtype=np.uint64))


def dense_u64(n, seed):
    """Sorted keys with gaps 1..20: a block of 2000 spans a few dozen periods of 1000."""
    rng = np.random.default_rng(seed)
    return (10 ** 12 + np.cumsum(rng.integers(1, 21, n))).astype(np.uint64)


def write_sosd(path, v):
    path.write_bytes(struct.pack("<Q", len(v)) + v.tobytes())


def run(args):
    from leco_fix import cli
    out = _io.StringIO()
    if args[0] == "fix_int":
        cli.run_fix_int(args[1:], os.path.join("..", "data"), 1, "fix_log", out=out)
    else:
        cli.run_fix_driver(args[0], args[1:], os.path.join("..", "data"), 64, 1, "fix_log", out=out)
    return out.getvalue().splitlines()


def lookup_mark(v, width, seed=1):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    idx = torch.randint(0, len(v), (len(v),), device="cuda", generator=g, dtype=torch.int64).cpu().numpy()
    return int(sum(int(x) for x in v[idx].tolist())) & ((1 << width) - 1)


def check_common(fields, method, fname, n, bs, nbytes, model_size, width):
    blocks = (n + bs - 1) // bs
    raw = float(width // 8 * n)
    assert fields[0] == method and fields[1] == fname and fields[2] == str(blocks)
    assert fields[3] == "%g" % (nbytes * 100.0 / raw)
    assert fields[4] == "%g" % (model_size * blocks * 100.0 / raw)
    assert fields[5] == "%g" % ((nbytes - float(model_size * blocks)) * 100.0 / raw)
    assert float(fields[6]) > 0 and float(fields[7]) > 0


@pytest.mark.parametrize("prog,codec", [("leco_int_fix", "leco"), ("FOR_int_fix", "for"), ("Delta_int_fix", "delta")])
def test_fix_drivers_plain(layout, oracle_mod, prog, codec):
    n, blocks = 200_003, 100                           # block_size 2000, 101 blocks
    v = sorted_u64(n, 3)
    write_sosd(layout / "data" / "keys_u64", v)
    say(f"cli {prog} plain n={n}")
    lines = run([prog, "keys_u64", str(blocks), "12", "1"])
    assert len(lines) == 1, lines
    f = lines[0].split()
    if codec == "leco":
        stream, offs = oracle_mod.encode_column(v, 2000)
    else:
        stream, offs = oracle_mod.base_encode_column(codec, v, 2000)
    method = {"leco": "piecewise_fix_op_max", "for": "FOR_my", "delta": "delta_my"}[codec]
    check_common(f, method, "keys_u64", n, 2000, int(offs[-1]), 12, 64)
    if codec == "delta":
        assert len(f) == 9 and f[8] == "0"
    else:
        assert len(f) == 12 and f[8] == "0" and f[9] == "0"
        # no filter ran: both filter times repeat the random-access interval
        assert f[10] == f[11] and float(f[10]) > 0
    assert int(open("fix_log").read().split()[-1]) == lookup_mark(v, 64)


@pytest.mark.parametrize("prog", ["leco_int_fix", "FOR_int_fix"])
def test_fix_drivers_filters(layout, prog):
    n, blocks = 300_000, 150
    v = dense_u64(n, 5)
    write_sosd(layout / "data" / "keys_u64", v)
    filt = int(v[n // 3])
    say(f"cli {prog} filter n={n}")
    f = run([prog, "keys_u64", str(blocks), "12", "1", str(filt)])[0].split()
    assert int(f[9]) == int((v > filt).sum()) and float(f[10]) > 0
    base, f1, f2 = 1000, 100, 600
    say(f"cli {prog} close n={n}")
    lines = run([prog, "keys_u64", str(blocks), "12", "1", str(f1), str(f2), str(base)])
    assert len(lines) == 1, lines[:5]                   # no "wrong" lines
    f = lines[0].split()
    r = v % base
    assert int(f[9]) == int(((r > f1) & (r < f2)).sum())
    assert float(f[11]) > 0


def test_fix_driver_text_inputs(layout):
    n = 50_000
    v = sorted_u64(n, 7)
    text = "\n".join(map(str, v.tolist())) + "\n"
    (layout / "data" / "keys.txt").write_text(text)
    (layout / "run" / "keys.txt").write_text(text)     # Delta_int_fix reads text from <file> itself
    say("cli text inputs")
    a = run(["leco_int_fix", "keys.txt", "50", "12", "0"])[0].split()
    b = run(["Delta_int_fix", "keys.txt", "50", "12", "0"])[0].split()
    assert a[2] == b[2] == "50"
    (layout / "run" / "keys.txt").unlink()
    with pytest.raises(Exception):
        run(["Delta_int_fix", "keys.txt", "50", "12", "0"])


def test_fix_int_opmax(layout, oracle_mod):
    n, blocks = 120_007, 60                            # block_size 2000, 61 blocks
    i = np.arange(n, dtype=np.float64)
    v = np.round(3.0e9 + 0.4123457 * i).astype(np.uint32)
    (layout / "data" / "keys_u32.txt").write_text("\n".join(map(str, v.tolist())) + "\n")
    say("cli fix_int piecewise_fix_op_max")
    lines = run(["fix_int", "piecewise_fix_op_max", "keys_u32.txt", str(blocks), "12"])
    assert len(lines) == 1, lines
    f = lines[0].split()
    stream, offs, fix = oracle_mod.opmax_encode_column(v, 2000)
    assert len(f) == 9
    check_common(f, "piecewise_fix_op_max", "keys_u32.txt", n, 2000, int(offs[-1]), 12, 32)
    assert float(f[8]) > 0
    assert int(open("fix_log").read().split()[-1]) == lookup_mark(v, 32)
    with pytest.raises(SystemExit):
        run(["fix_int", "FOR_my", "keys_u32.txt", str(blocks), "4"])


def test_cli_module_process(layout):
    v = sorted_u64(20_000, 9)
    write_sosd(layout / "data" / "k", v)
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "learn-to-compress_amd"), env.get("PYTHONPATH", "")])
    say("cli as a process")
    r = subprocess.run([sys.executable, "-m", "leco_fix.cli", "--fix-log", "", "leco_int_fix", "k", "10", "12", "1"],
                       capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    f = r.stdout.strip().split()
    assert f[0] == "piecewise_fix_op_max" and f[2] == "10" and len(f) == 12
    assert not os.path.exists("fix_log")
