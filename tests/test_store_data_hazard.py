"""Static check of the built library's gfx950 code: no vector-memory store of
more than 8 bytes (dwordx3 / dwordx4) may have the upper part of its data
registers rewritten by a later load while the store can still be reading them.

On MI355X a 16-byte store followed by an LDS load into the same registers
wrote the load's data as its upper 8 bytes, now and then, under load (batched
decodes returned wrong values).  The kernels store 8 bytes per lane wherever
stored registers are refilled by loads; this test disassembles the library
that the tests and the benchmark load and fails on any store whose data
registers a load (LDS, global, flat, buffer or scratch) rewrites within the
next NEAR instructions, or rewrites in its upper 8 bytes within FAR
instructions, along every path (fall-through and branch targets) up to a wait
for all vector memory operations.  CPU only: it needs the ROCm LLVM tools and
the built library, and skips without them."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "learn-to-compress_amd", "lib", "libleco_fix.so")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"
NEAR, FAR = 2, 64

STORE = re.compile(r"^\s+(global|flat|buffer|scratch)_store_(dwordx3|dwordx4|b96|b128)\s+(.*?)\s*//\s*([0-9A-F]+):")
LOAD = re.compile(r"^\s+(ds_read\w*|ds_load\w*|ds_bpermute\w*|ds_permute\w*|ds_swizzle\w*|\w+_load\w*)\s+"
                  r"(v\[\d+:\d+\]|v\d+)")
BRANCH = re.compile(r"^\s+s_(cbranch_\w+|branch)\s+.*<(\w+)\+0x([0-9a-f]+)>")
INSTR = re.compile(r"^\s+([a-z_0-9]+)\b.*//\s*([0-9A-F]+):")
FUNC = re.compile(r"^([0-9a-f]+) <(\w+)>:$")


def vregs(tok):
    m = re.match(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return int(m.group(1)), int(m.group(2))
    m = re.match(r"v(\d+)$", tok)
    return (int(m.group(1)),) * 2 if m else None


def store_data(kind, operands):
    """The data register range of a store: buffer stores name it first,
    global, flat and scratch stores second (after the address or `off`)."""
    toks = [t.strip() for t in operands.split(",")]
    return vregs(toks[0] if kind == "buffer" else toks[1])


def code_objects(tmp):
    subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={tmp}/fatbin.bin", LIB, f"{tmp}/lib.copy"],
                   check=True, capture_output=True)
    data = open(f"{tmp}/fatbin.bin", "rb").read()
    starts = [m.start() for m in re.finditer(rb"__CLANG_OFFLOAD_BUNDLE__", data)] + [len(data)]
    out = []
    for i in range(len(starts) - 1):
        part = f"{tmp}/bundle{i}.bin"
        with open(part, "wb") as f:
            f.write(data[starts[i]:starts[i + 1]])
        co = f"{tmp}/bundle{i}.co"
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", f"--input={part}", f"--targets={TARGET}",
                        f"--output={co}", "--unbundle"], check=True, capture_output=True)
        out.append(co)
    return out


def scan(lines):
    """(function, store, distance, load) for every store whose data registers a
    load rewrites too early."""
    addr_line, func_addr, fn_of = {}, {}, {}
    fn = None
    for i, line in enumerate(lines):
        m = FUNC.match(line)
        if m:
            fn = m.group(2)
            func_addr[fn] = int(m.group(1), 16)
            continue
        m = INSTR.match(line)
        if m:
            addr_line[int(m.group(2), 16)] = i
            fn_of[i] = fn
    hits = []
    for i, line in enumerate(lines):
        m = STORE.match(line)
        if not m:
            continue
        data = store_data(m.group(1), m.group(3))
        if data is None:
            continue
        upper = (data[0] + 2, data[1])
        stack, seen = [(i + 1, 0)], set()
        while stack:
            j, d = stack.pop()
            while j < len(lines) and d < FAR:
                if (j, d) in seen:
                    break
                seen.add((j, d))
                t = lines[j]
                if FUNC.match(t):
                    break
                im = INSTR.match(t)
                if not im:
                    j += 1
                    continue
                op = im.group(1)
                d += 1
                if op == "s_endpgm" or op.startswith("s_setpc") or op.startswith("s_swappc"):
                    break
                if op == "s_waitcnt" and re.search(r"vmcnt\(0\)", t):
                    break
                lm = LOAD.match(t)
                if lm and not op.startswith("ds_write") and not op.startswith("ds_store"):
                    r = vregs(lm.group(2))
                    if r:
                        def overlaps(a):
                            return not (r[1] < a[0] or r[0] > a[1])
                        if (d <= NEAR and overlaps(data)) or overlaps(upper):
                            hits.append((fn_of.get(i), line.split("//")[0].strip(), d, t.split("//")[0].strip()))
                bm = BRANCH.match(t)
                if bm and bm.group(2) in func_addr:
                    target = func_addr[bm.group(2)] + int(bm.group(3), 16)
                    if target in addr_line:
                        stack.append((addr_line[target], d))
                    if bm.group(1) == "branch":
                        break
                j += 1
    return hits


def test_no_wide_store_data_rewritten_by_a_load(tmp_path):
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump"):
        if not os.path.exists(f"{LLVM}/{tool}"):
            pytest.skip(f"{LLVM}/{tool} is not installed")
    if not os.path.exists(LIB):
        pytest.skip("the library is not built")
    hits, stores = [], 0
    for co in code_objects(str(tmp_path)):
        text = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", co], check=True,
                              capture_output=True, text=True).stdout
        lines = text.split("\n")
        stores += sum(1 for line in lines if STORE.match(line))
        hits += scan(lines)
    assert stores > 0, "no wide stores found: the disassembly format changed"
    assert not hits, "\n".join(f"{f}: {s} / {d} instructions later: {ld}" for f, s, d, ld in hits[:20])


def test_scan_finds_the_known_bad_pattern():
    """The scan itself: a 16-byte store followed at once by an LDS load into
    its registers (the pattern that wrote wrong values) is reported; the same
    load into other registers, or after a wait for all stores, is not."""
    head = ["0000000000001000 <_Zk>:"]

    def asm(body):
        out = list(head)
        for k, ins in enumerate(body):
            out.append(f"\t{ins:<60}// {0x1000 + 8 * k:012X}: 00000000")
        return out

    bad = asm(["global_store_dwordx4 v[24:25], v[2:5], off nt", "ds_read2_b64 v[2:5], v36 offset1:1", "s_endpgm"])
    assert len(scan(bad)) == 1
    upper_late = asm(["global_store_dwordx4 v[24:25], v[2:5], off"] + ["v_add_u32_e32 v9, v9, v9"] * 30 +
                     ["ds_read_b64 v[4:5], v36", "s_endpgm"])
    assert len(scan(upper_late)) == 1
    lower_late = asm(["global_store_dwordx4 v[24:25], v[2:5], off"] + ["v_add_u32_e32 v9, v9, v9"] * 5 +
                     ["ds_read_b64 v[2:3], v36", "s_endpgm"])
    assert scan(lower_late) == []
    other = asm(["global_store_dwordx4 v[24:25], v[2:5], off", "ds_read2_b64 v[6:9], v36 offset1:1", "s_endpgm"])
    assert scan(other) == []
    waited = asm(["global_store_dwordx4 v[24:25], v[2:5], off", "s_waitcnt vmcnt(0)",
                  "ds_read2_b64 v[2:5], v36 offset1:1", "s_endpgm"])
    assert scan(waited) == []
    narrow = asm(["global_store_dwordx2 v[24:25], v[2:3], off", "ds_read_b64 v[2:3], v36", "s_endpgm"])
    assert scan(narrow) == []
