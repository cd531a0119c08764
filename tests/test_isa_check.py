"""The wide-store data hazard check (isa_check) on hand-written listings and on
built STREAM kernels.  CPU only: hipcc cross-compiles for gfx950 and
llvm-objdump disassembles the code objects built here."""
import os
import subprocess
import tempfile

import pytest

from threadcoarsenedstencilgenerator_amd import build as B
from threadcoarsenedstencilgenerator_amd import isa_check, manifest
from threadcoarsenedstencilgenerator_amd.generator.emit import emit_source
from threadcoarsenedstencilgenerator_amd.spec import F32C

STORE_SGPR = "\tbuffer_store_dwordx4 v[0:3], v7, s[0:3], s13 offen nt"
STORE_ZERO = "\tbuffer_store_dwordx4 v[0:3], v7, s[0:3], 0 offen nt"


def listing(*lines):
    return "k.co:\tfile format elf64-amdgpu\n\nDisassembly of section .text:\n\n0000000000001900 <k>:\n" + \
        "\n".join(lines) + "\n\ts_endpgm\n"


def test_adjacent_overwrite_is_flagged():
    bad = isa_check.violations(listing(STORE_SGPR, "\tv_add_f64 v[0:1], v[4:5], v[2:3]"))
    assert len(bad) == 1 and bad[0][2] == 0


def test_overwrite_after_one_instruction_is_below_the_compiler_rule():
    # one wait state: seen safe on the hardware, but below the two the compiler gives the case
    bad = isa_check.violations(listing(STORE_SGPR, "\tv_mul_f64 v[8:9], v[4:5], v[2:3]",
                                       "\tv_mov_b32_dpp v1, v5 wave_shr:1 row_mask:0xf bank_mask:0xf"))
    assert len(bad) == 1 and bad[0][2] == 1


def test_s_nop_counts_its_wait_states():
    assert not isa_check.violations(listing(STORE_ZERO, "\ts_nop 1", "\tv_add_f64 v[0:1], v[4:5], v[2:3]"))
    assert isa_check.violations(listing(STORE_ZERO, "\ts_nop 0", "\tv_add_f64 v[0:1], v[4:5], v[2:3]"))


def test_writes_of_other_registers_loads_and_narrow_stores_pass():
    assert not isa_check.violations(listing(STORE_SGPR, "\tv_add_f64 v[8:9], v[4:5], v[2:3]",
                                            "\tv_add_f64 v[10:11], v[4:5], v[2:3]", "\tv_add_f64 v[0:1], v[4:5], v[2:3]"))
    # a vector load behind the store is processed in order: not this hazard
    assert not isa_check.violations(listing(STORE_SGPR, "\tbuffer_load_dwordx4 v[0:3], v9, s[4:7], 0 offen"))
    # 64-bit stores are outside the rule
    assert not isa_check.violations(listing("\tbuffer_store_dwordx2 v[0:1], v7, s[0:3], s13 offen",
                                            "\tv_add_f64 v[0:1], v[4:5], v[2:3]"))


def test_branch_inside_the_window_is_flagged():
    bad = isa_check.violations(listing(STORE_SGPR, "\ts_cbranch_execz 12"))
    assert len(bad) == 1


def test_agpr_data_registers_count():
    bad = isa_check.violations(listing("\tglobal_store_dwordx4 v2, a[4:7], s[0:1]", "\tv_accvgpr_write_b32 a5, v1"))
    assert len(bad) == 1


@pytest.mark.parametrize("name", ["box2d_r1_2x4_dppw_f64", "star3d_r4_4x2_dppw_f32c"])
def test_built_stream_kernels_have_no_violation(name):
    """The two code shapes whose earlier builds had a store followed at once by a
    write of its data (and returned wrong points on the GPU) compile clean now."""
    if name == "box2d_r1_2x4_dppw_f64":
        d = manifest.cfg1_desc("dppw").replace(xc=2, yc=4, max_threads=256)
    else:
        d = manifest.cfg2_desc(4, 2, 8, shape="dppw", max_threads=256, dtype=F32C).replace(
            prefetch=0, comp_group=21, zfirst=False, cl_size=32)
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "k.hip")
        co = os.path.join(td, "k.co")
        with open(src, "w") as f:
            f.write(emit_source(d))
        subprocess.run([B.HIPCC, *B.HIP_FLAGS, "--genco", "--no-gpu-bundle-output", "-x", "hip", src, "-o", co],
                       check=True, capture_output=True)
        text = isa_check.disassemble(co)
        assert "buffer_store_dwordx4" in text
        assert isa_check.violations(text) == []


def _addressed(*lines, base=0x1900):
    out = []
    for k, l in enumerate(lines):
        out.append(f"\t{l:<60}// {base + 4 * k:012X}: 00000000")
    return "k.co:\tfile format elf64-amdgpu\n\n0000000000001900 <k>:\n" + "\n".join(out) + "\n"


def test_branch_targets_are_followed():
    # loop: the back edge (simm16 -4 words -> 0x1904) lands on a write of the store's data
    loop = _addressed("v_mov_b32_e32 v9, 0", "v_add_f64 v[0:1], v[4:5], v[2:3]", "s_nop 4",
                      "buffer_store_dwordx4 v[0:3], v7, s[0:3], 0 offen nt", "s_cbranch_execnz 65532", "s_endpgm")
    bad = isa_check.violations(loop)
    assert len(bad) == 1 and bad[0][1].startswith("v_add_f64") and bad[0][2] == 1
    # back edge to 0x1900: the write is two instructions into the body, clean on both paths
    ok = _addressed("v_mov_b32_e32 v9, 0", "v_mov_b32_e32 v10, 0", "v_add_f64 v[0:1], v[4:5], v[2:3]",
                    "buffer_store_dwordx4 v[0:3], v7, s[0:3], 0 offen nt", "s_cbranch_execnz 65531", "s_endpgm")
    assert isa_check.violations(ok) == []
