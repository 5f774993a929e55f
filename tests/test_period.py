"""The close filter's period index (learn-to-compress_amd/csrc/leco_period.h):
one FP64 multiply by 1/base and a remainder correction, run on the host
against integer division n / base at n = q*base - 1, q*base, q*base + 1 for
every q < 2^16 at bases 1, 2, 3, 2^31 - 2, 2^31 - 1, around multiples of
random bases, and at seeded random n."""
import os
import subprocess

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "period_check")


def test_period_quotient_equals_integer_division():
    assert os.path.exists(EXE), "built by learn-to-compress_amd/build.py"
    r = subprocess.run([EXE, "2000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout
