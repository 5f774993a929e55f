"""CPU: the text input parser (leco_fix_text_info, host only) against
libstdc++'s own stream extraction (tests/cpp/istream_values.cpp, the loop of
experiments/leco_int_fix_template.cpp:43-58 and experiments/fix_int.cpp:30-44):
the value count must agree on every edge case and on random byte soup, which
pins where the input ends (end of file without a newline, overflow, signs,
stray characters).  The values themselves are compared on the GPU, where
leco_fix_text_load can run."""
import numpy as np
import pytest

from text_corpus import edge_cases, fuzz_bytes, random_text


@pytest.mark.parametrize("width", [32, 64])
def test_text_count_edge_cases(tmp_path, istream_values, width):
    from leco_fix import io
    for name, body in edge_cases(width).items():
        p = tmp_path / f"{name}.txt"
        p.write_bytes(body)
        assert io.text_info(p, width) == len(istream_values(p, width)), name


@pytest.mark.parametrize("width", [32, 64])
def test_text_count_random(tmp_path, istream_values, width):
    from leco_fix import io
    rng = np.random.default_rng(width)
    for trailing in (True, False):
        body, v = random_text(rng, 5000, width, trailing)
        p = tmp_path / f"r{trailing}.txt"
        p.write_bytes(body)
        assert io.text_info(p, width) == len(v) - (0 if trailing else 1)
        assert io.text_info(p, width) == len(istream_values(p, width))
    for k in range(200):
        p = tmp_path / f"f{k}.txt"
        p.write_bytes(fuzz_bytes(rng, int(rng.integers(0, 300))))
        assert io.text_info(p, width) == len(istream_values(p, width)), p.read_bytes()


def test_text_info_rejects(tmp_path):
    from leco_fix import io
    from leco_fix._lib import LecoError
    with pytest.raises(LecoError):
        io.text_info(tmp_path / "missing.txt", 64)
    p = tmp_path / "a.txt"
    p.write_bytes(b"1\n")
    with pytest.raises(LecoError):
        io.text_info(p, 16)
