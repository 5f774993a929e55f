"""Static checks of the built library's gfx950 code for two register hazards
seen on MI355X: a memory instruction still queued in its pipe reads its
VGPR operands after a load of the OTHER pipe has already refilled them.

1. A vector-memory store of more than 8 bytes (dwordx3 / dwordx4) followed
   at once by an LDS load into the same registers wrote the load's data as
   its upper 8 bytes, now and then, under load (batched decodes returned
   wrong values).  Stores whose registers loads refill are 8 bytes wide.
2. An LDS atomic (ds_or of the encoders' image packer) whose address and
   data registers a later global load refilled ORed the wrong bits at the
   wrong place, in several percent of a 500M-value FOR column's blocks.
   Fields are computed before any LDS atomic of a chunk, and waits for the
   LDS queue separate the atomics from later loads.

The tests disassemble the library that the tests and the benchmark load
(llvm-objdump of our own code object) and fail on: a wide store whose data
registers a load (LDS, global, flat, buffer or scratch) rewrites within the
next NEAR instructions, or in its upper 8 bytes within FAR instructions; an
LDS atomic, permute or swizzle whose operand registers a global, flat,
buffer or scratch load rewrites within FAR instructions with no wait for
the LDS queue (lgkmcnt(0)) in between.  Paths follow fall-through and branch
targets.  CPU only: it needs the ROCm LLVM tools and the built library, and
skips without them."""
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


VMEM_LOAD = re.compile(r"^\s+((?:global|flat|buffer|scratch)_load\w*)\s+(v\[\d+:\d+\]|v\d+)")
DS_OP = re.compile(r"^\s+(ds_\w+)\s+([^/]*?)\s*//")
PLAIN_DS = ("ds_read", "ds_write", "ds_load", "ds_store", "ds_nop")


def index(lines):
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
    return addr_line, func_addr, fn_of


def walk(lines, start, addr_line, func_addr, stop, visit):
    """visit(line, distance) on every instruction reachable from line `start`
    within FAR instructions; stop(op, line) ends a path (after counting it)."""
    stack, seen = [(start, 0)], set()
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
            if op == "s_endpgm" or op.startswith("s_setpc") or op.startswith("s_swappc") or stop(op, t):
                break
            visit(t, d)
            bm = BRANCH.match(t)
            if bm and bm.group(2) in func_addr:
                target = func_addr[bm.group(2)] + int(bm.group(3), 16)
                if target in addr_line:
                    stack.append((addr_line[target], d))
                if bm.group(1) == "branch":
                    break
            j += 1


def scan_lds_atomics(lines):
    """(function, LDS op, distance, load) for every LDS atomic / permute whose
    operand registers a vector-memory load rewrites before an lgkmcnt(0)."""
    addr_line, func_addr, fn_of = index(lines)
    hits = []
    for i, line in enumerate(lines):
        m = DS_OP.match(line)
        if not m or m.group(1).startswith(PLAIN_DS):
            continue
        op = m.group(1)
        toks = [t.split()[0] for t in m.group(2).split(",") if t.strip()]
        returns = op.startswith(("ds_bpermute", "ds_permute", "ds_swizzle")) or "_rtn" in op
        srcs = [r for r in (vregs(t) for t in (toks[1:] if returns else toks)) if r]
        if not srcs:
            continue

        def visit(t, d, i=i, srcs=srcs, line=line):
            lm = VMEM_LOAD.match(t)
            if lm:
                r = vregs(lm.group(2))
                if r and any(not (r[1] < s[0] or r[0] > s[1]) for s in srcs):
                    hits.append((fn_of.get(i), line.split("//")[0].strip(), d, t.split("//")[0].strip()))

        walk(lines, i + 1, addr_line, func_addr,
             lambda op_, t: op_ == "s_waitcnt" and re.search(r"lgkmcnt\(0\)", t) is not None, visit)
    return hits


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


def test_no_memory_operand_rewritten_by_another_pipes_load(tmp_path):
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump"):
        if not os.path.exists(f"{LLVM}/{tool}"):
            pytest.skip(f"{LLVM}/{tool} is not installed")
    if not os.path.exists(LIB):
        pytest.skip("the library is not built")
    hits, lds_hits, stores, atomics = [], [], 0, 0
    for co in code_objects(str(tmp_path)):
        text = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", co], check=True,
                              capture_output=True, text=True).stdout
        lines = text.split("\n")
        stores += sum(1 for line in lines if STORE.match(line))
        atomics += sum(1 for line in lines if DS_OP.match(line) and not DS_OP.match(line).group(1).startswith(PLAIN_DS))
        hits += scan(lines)
        lds_hits += scan_lds_atomics(lines)
    assert stores > 0 and atomics > 0, "no wide stores or LDS atomics found: the disassembly format changed"
    assert not hits, "\n".join(f"{f}: {s} / {d} instructions later: {ld}" for f, s, d, ld in hits[:20])
    assert not lds_hits, "\n".join(f"{f}: {s} / {d} instructions later: {ld}" for f, s, d, ld in lds_hits[:20])


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


def test_lds_scan_finds_the_known_bad_pattern():
    """The LDS-atomic scan: an image OR whose address register the next
    round's global load refills (the FOR encoder's pattern) is reported; the
    same load after an lgkmcnt(0) wait, or into other registers, is not; a
    plain LDS write is not checked."""
    head = ["0000000000001000 <_Zk>:"]

    def asm(body):
        out = list(head)
        for k, ins in enumerate(body):
            out.append(f"\t{ins:<60}// {0x1000 + 8 * k:012X}: 00000000")
        return out

    bad = asm(["ds_or_b32 v40, v42 offset:4", "s_or_b64 exec, exec, s[18:19]", "global_load_dword v40, v[36:37], off",
               "s_endpgm"])
    assert len(scan_lds_atomics(bad)) == 1
    data = asm(["ds_or_b32 v40, v42 offset:4"] + ["v_add_u32_e32 v9, v9, v9"] * 20 +
               ["scratch_load_dwordx2 v[42:43], off, off offset:8", "s_endpgm"])
    assert len(scan_lds_atomics(data)) == 1
    waited = asm(["ds_or_b32 v40, v42 offset:4", "s_waitcnt lgkmcnt(0)", "global_load_dword v40, v[36:37], off",
                  "s_endpgm"])
    assert scan_lds_atomics(waited) == []
    other = asm(["ds_or_b32 v40, v42 offset:4", "global_load_dword v41, v[36:37], off", "s_endpgm"])
    assert scan_lds_atomics(other) == []
    plain = asm(["ds_write_b128 v1, v[4:7]", "global_load_dwordx4 v[4:7], v[4:5], off", "s_endpgm"])
    assert scan_lds_atomics(plain) == []
    perm = asm(["ds_bpermute_b32 v33, v1, v31", "global_load_dword v31, v2, s[6:7]", "s_endpgm"])
    assert len(scan_lds_atomics(perm)) == 1
