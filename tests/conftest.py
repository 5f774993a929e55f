import glob
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "learn-to-compress_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_cases():
    with open(os.path.join(GOLDEN, "cases.json")) as f:
        return json.load(f)


def load_golden(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def istream_values(tmp_path_factory):
    """tests/cpp/istream_values.cpp built with g++: (path, width) -> the values
    libstdc++'s extraction loop reads from a text file."""
    import subprocess
    exe = tmp_path_factory.mktemp("istream") / "istream_values"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "cpp", "istream_values.cpp")],
                   check=True, capture_output=True)

    def read(path, width):
        out = subprocess.run([str(exe), str(path), str(width)], check=True, capture_output=True, text=True).stdout
        lines = out.split()
        n = int(lines[0])
        return np.array([int(x) for x in lines[1:1 + n]], dtype=np.uint64)
    return read
