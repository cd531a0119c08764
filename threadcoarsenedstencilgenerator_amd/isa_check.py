"""Checks a built gfx950 code object for the wide-store data hazard.

A vector-memory store of more than 64 bits (``*_store_dwordx3``/``x4``) reads its
data VGPRs after issue.  On gfx950 a VALU instruction that overwrites those VGPRs
too soon after the store can put its own values into the stored data for some
lanes.  The ISA documents one wait state for this case and exempts buffer stores
that take an SGPR ``soffset``; the compiler follows the document, so an exempt
store can be followed at once by such a write.  Measured on MI355X: kernels with
that adjacency (no instruction between the store and the write) stored wrong
values in a varying number of points from run to run; kernels with at least one
instruction between never did.  The emitter therefore gives every STREAM
store an ``soffset`` of 0 (the row/plane offset goes into the per-lane offset), so
the compiler inserts the wait states itself, and the build runs
:func:`check_code_object` on every kernel it compiles.

The check walks the disassembly (``llvm-objdump -d``) in program order.  After
each wide store it counts wait states (one per instruction, N+1 for ``s_nop N``)
until an instruction writes one of the store's data registers; fewer than
``MIN_WAIT_STATES`` is a violation.  Branches inside that window are followed
to their targets (and conditional ones also fall through), so a loop's back
edge is checked like straight-line code.
"""
from __future__ import annotations

import os
import re
import subprocess

OBJDUMP = os.environ.get("TCSG_OBJDUMP", "/opt/rocm/llvm/bin/llvm-objdump")
# wait states the compiler's own hazard rule gives this case on gfx950 (s_nop 1 after the
# store), one more than the smallest gap seen to be safe
MIN_WAIT_STATES = 2

_WIDE_STORE = re.compile(r"^(buffer|global|flat|scratch)_store_(dwordx[34]|b96|b128|format_xyzw?)\b")
_REG = re.compile(r"^([va])(?:\[(\d+):(\d+)\]|(\d+))$")


def _regs(tok: str):
    m = _REG.match(tok.strip().rstrip(","))
    if not m:
        return set()
    kind = m.group(1)
    if m.group(4) is not None:
        return {(kind, int(m.group(4)))}
    return {(kind, i) for i in range(int(m.group(2)), int(m.group(3)) + 1)}


def _written(ins: str):
    """Vector registers a VALU instruction writes (its first operand).  Memory
    instructions are processed in order behind the store, so a later load's
    write is not this hazard."""
    op, _, rest = ins.partition(" ")
    if op.startswith("v_"):
        if not rest:
            return set()
        first = rest.split(",")[0].strip()
        return _regs(first)
    return set()


def parse(text: str):
    """[(address or None, instruction)] in program order, from llvm-objdump's
    listing (each line ends with `// ADDRESS: ENCODING`) or an assembler file."""
    out = []
    for line in text.splitlines():
        code, _, comment = line.partition("//")
        s = code.strip()
        if not s or s.endswith(":") or s.startswith(("Disassembly", ".")) or "file format" in s:
            continue
        m = re.match(r"\s*([0-9A-Fa-f]+):", comment)
        out.append((int(m.group(1), 16) if m else None, s))
    return out


def _branch_target(addr, ins):
    """Address a SOPP branch jumps to (simm16 words after the next instruction)."""
    m = re.match(r"s_(?:c)?branch\w*\s+(-?\d+)", ins)
    if addr is None or not m:
        return None
    k = int(m.group(1))
    if k >= 1 << 15:
        k -= 1 << 16
    return addr + 4 + 4 * k


def _waits(ins):
    if ins.startswith("s_nop"):
        m = re.match(r"s_nop\s+(0x[0-9a-fA-F]+|\d+)", ins)
        return (int(m.group(1), 0) if m else 0) + 1
    return 1


def violations(text: str):
    """(store, offending instruction, wait states) for every wide store whose data
    registers a VALU instruction writes within MIN_WAIT_STATES wait states on any
    path (fall-through and branch targets).  A branch whose target cannot be
    resolved counts as a violation."""
    prog = parse(text)
    where = {a: i for i, (a, _) in enumerate(prog) if a is not None}
    bad = []

    def walk(i, data, waits, store, seen):
        while i < len(prog) and waits < MIN_WAIT_STATES:
            addr, ins = prog[i]
            op = ins.partition(" ")[0]
            if _written(ins) & data:
                bad.append((store, ins, waits))
                return
            if op == "s_endpgm" or op.startswith("s_setpc"):
                return
            if op.startswith(("s_branch", "s_cbranch")):
                t = _branch_target(addr, ins)
                if t is None or t not in where:
                    bad.append((store, ins, waits))
                    return
                if (t, waits) not in seen:
                    seen.add((t, waits))
                    walk(where[t], data, waits + 1, store, seen)
                if op.startswith("s_branch"):
                    return
            waits += _waits(ins)
            i += 1

    for i, (addr, ins) in enumerate(prog):
        op, _, rest = ins.partition(" ")
        if not _WIDE_STORE.match(op):
            continue
        # data operand: first for buffer stores, second (after the address) for global/flat/scratch
        ops = [t.strip() for t in rest.split(",")] if rest else []
        k = 0 if op.startswith("buffer") else 1
        data = _regs(ops[k]) if len(ops) > k else set()
        if data:
            walk(i + 1, data, 0, ins, set())
    return bad


def disassemble(co_path: str) -> str:
    r = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", co_path], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"llvm-objdump failed on {co_path}: {r.stderr}")
    return r.stdout


def check_code_object(co_path: str) -> None:
    """Raises RuntimeError naming the first offending instructions."""
    bad = violations(disassemble(co_path))
    if bad:
        lines = "\n".join(f"  {s}  ->  {n}  ({w} wait states)" for s, n, w in bad[:8])
        raise RuntimeError(f"{os.path.basename(co_path)}: {len(bad)} wide store(s) whose data registers are "
                           f"overwritten within {MIN_WAIT_STATES} wait states:\n{lines}")
