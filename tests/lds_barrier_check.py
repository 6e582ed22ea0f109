"""Static check of the built gfx950 code: no s_barrier is reachable with an LDS write whose
completion has not been waited for (s_waitcnt lgkmcnt(0)).

A workgroup barrier only publishes LDS writes that have completed. The compiler normally waits
for them before each barrier, but it missed that wait on the back edge of the FP32 site kernel's
work-queue loop (thread 0's write of the next site was still in flight at the loop-top barrier,
and the other wavefronts read the old value). This check walks the control-flow graph of every
kernel in the objects the build wrote, so a missing wait anywhere fails a CPU test.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LLVM = os.path.join(ROCM, "llvm", "bin")
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

# DS instructions that write LDS memory (cross-lane ds_bpermute / ds_permute / ds_swizzle and
# reads do not)
_WRITES = re.compile(r"^ds_(write|store|add_|sub_|rsub_|inc_|dec_|min_|max_|and_|or_|xor_|mskor|cmpst|cmpswap|"
                     r"wrxchg|wrap|condxchg|append|consume)")
_WAIT0 = re.compile(r"^s_waitcnt\b.*\blgkmcnt\(0\)")
_INSN = re.compile(r"^\s+(\S+)(.*?)//\s*([0-9A-Fa-f]+):(.*)$")
_FUNC = re.compile(r"^([0-9a-f]+) <(\S+)>:$")
_TARGET = re.compile(r"<(\S+)\+0x([0-9a-f]+)>")


def disassemble(obj: str) -> str:
    """gfx950 disassembly of the device code in a HIP object file ("" if it holds none)."""
    sections = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "-S", obj], check=True, capture_output=True,
                              text=True).stdout
    if ".hip_fatbin" not in sections:
        return ""
    with tempfile.TemporaryDirectory() as d:
        fat, co = os.path.join(d, "fatbin"), os.path.join(d, "co")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fat}", obj,
                        os.path.join(d, "copy.o")], check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--targets={TARGET}",
                        f"--input={fat}", f"--output={co}"], check=True, capture_output=True)
        r = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--no-show-raw-insn", co], check=True,
                           capture_output=True, text=True)
        return r.stdout


def functions(text: str):
    """(name, base address, [(address, opcode, operands and branch target)]) per function."""
    out, cur = [], None
    for line in text.split("\n"):
        m = _FUNC.match(line)
        if m:
            cur = (m.group(2), int(m.group(1), 16), [])
            out.append(cur)
            continue
        m = _INSN.match(line)
        if m and cur is not None:
            cur[2].append((int(m.group(3), 16), m.group(1), m.group(2) + " " + m.group(4)))
    return out


def pending_barriers(insns, base: int):
    """Addresses of s_barrier instructions that some path reaches with an LDS write not yet waited for."""
    index = {a: i for i, (a, _, _) in enumerate(insns)}
    succ = []
    for i, (a, op, rest) in enumerate(insns):
        s = []
        t = _TARGET.search(rest)
        if op.startswith("s_branch") or op.startswith("s_cbranch"):
            if t:
                s.append(index.get(base + int(t.group(2), 16)))
        if not (op == "s_branch" or op == "s_endpgm" or op.startswith("s_setpc")) and i + 1 < len(insns):
            s.append(i + 1)
        succ.append([j for j in s if j is not None])
    # forward dataflow over instructions: state = an LDS write may be in flight
    state = [None] * len(insns)
    state[0] = False
    work, bad = [0], set()
    while work:
        i = work.pop()
        p = state[i]
        a, op, rest = insns[i]
        if op == "s_barrier" and p:
            bad.add(a)
        if _WRITES.match(op):
            p = True
        elif _WAIT0.match(op + rest):
            p = False
        for j in succ[i]:
            if state[j] is None or (p and not state[j]):
                state[j] = bool(state[j]) or p
                work.append(j)
    return sorted(bad)


def check_objects(objs):
    """{kernel: [barrier addresses]} for every kernel with a barrier reachable with a pending LDS write."""
    found = {}
    for obj in objs:
        for name, base, insns in functions(disassemble(obj)):
            bad = pending_barriers(insns, base)
            if bad:
                found[name] = bad
    return found
