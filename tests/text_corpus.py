"""Text inputs for the text loader's tests: well-formed value files and
byte strings that exercise every way the reference's extraction loop ends
(no trailing newline, overflow, signs, stray characters)."""
import numpy as np


def edge_cases(width):
    top = (1 << width) - 1
    return {
        "newline_end": b"1\n2\n3\n",
        "eof_end": b"1\n2\n3",                       # the last value is not kept
        "spaces": b"  \t7 8\r\n9\v10\f 11 \n",
        "empty": b"",
        "only_space": b" \n\n ",
        "max": f"{top}\n{top}\n".encode(),
        "overflow": f"5\n{top + 1}\n6\n".encode(),  # ends at the value past the maximum
        "leading_zeros": b"000012 0 00\n",
        "junk_after": b"12abc 13\n",                 # 12 kept, then extraction fails
        "junk_first": b"x1 2\n",
        "minus": b"-5 7\n",                          # negated modulo 2^width
        "plus": b"+5 +0 7\n",
        "minus_zero": b"-0 1\n",
        "sign_glued": b"5-3 4\n",                    # 5, then -3
        "lone_sign": b"4 - 5\n",
        "double_sign": b"4 --5\n",
        "sign_eof": b"4 -",
        "hex": b"0x10 5\n",
        "huge_digits": b"1 " + b"9" * 40 + b"\n",
        "nul": b"3\x004\n",
        "minus_max": f"-{top} 2\n".encode(),
        "minus_over": f"-{top + 1} 2\n".encode(),
    }


def random_text(rng, n, width, trailing=True):
    hi = (1 << width) - 1
    v = rng.integers(0, hi, n, dtype=np.uint64, endpoint=True)
    seps = np.array([b" ", b"\n", b"\t", b"  ", b"\r\n"], dtype=object)
    pick = rng.integers(0, len(seps), n)
    parts = []
    for x, s in zip(v.tolist(), pick.tolist()):
        parts.append(str(x).encode())
        parts.append(seps[s])
    if not trailing:
        parts.pop()
    return b"".join(parts), v


def fuzz_bytes(rng, length):
    alphabet = np.frombuffer(b"0123456789012345678901234567890123456789  \n\n\t+-x", dtype=np.uint8)
    return alphabet[rng.integers(0, len(alphabet), length)].tobytes()
