"""Production kernels over the full 512^3 benchmark grid, bit for bit
against the CPU oracle (one and three time steps), plus the FP32-storage
kernel's error against the FP64 result after 20 steps.

The small-grid parity suite covers the option matrix; this file covers the
exact configurations bench.py runs, whose z-chunks (64-128 planes) the small
grids do not reach."""
import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import spec, tuned
from threadcoarsenedstencilgenerator_amd.spec import F32C, F64

pytestmark = pytest.mark.gpu
GRID = (520, 520, 520)


def _run(desc, block, steps, grid=GRID):
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    k = rt.Kernel(desc)
    g = k.geometry(grid)
    mask = rt.mask_seeded(2, desc.n)
    k.set_mask(mask)
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), rt.INIT_NORMAL, 1)
    torch.cuda.synchronize()
    a0 = grids.host(0)
    fin = k.run(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1), steps)
    torch.cuda.synchronize()
    got = grids.host(fin)
    del grids
    torch.cuda.empty_cache()
    og = oracle.Geometry(3, desc.range, desc.alignment, desc.cl_size, grid,
                         (1, 1, 1) if desc.remainder else desc.coarsening)
    run = oracle.run_f32c if desc.dtype == spec.F32C else oracle.run_f64
    ref = run(desc, og, mask, a0.copy(), a0.copy(), steps)
    return got, ref


@pytest.mark.parametrize("which", ["f64", "f32c"])
@pytest.mark.parametrize("steps", [1, 3])
def test_production_kernel_bitexact_512(which, steps):
    cfg = tuned.CFG2_F64 if which == "f64" else tuned.CFG2_F32C
    print(f"full-size {which} {cfg['desc'].key()} block {cfg['block']} steps {steps}", flush=True)
    got, ref = _run(cfg["desc"], tuple(cfg["block"]), steps)
    ui = np.uint32 if got.dtype == np.float32 else np.uint64
    bad = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert bad.size == 0, f"{bad.size} points differ from the oracle"


@pytest.mark.parametrize("which", ["f64", "f32c", "f64_unstaged", "f32c_unstaged"])
def test_remainder_kernel_bitexact_511(which):
    """The workload's kernels generated with remainder handling (the staged production
    kernels, and those loading their own rows) on a 511^3 interior (no axis a multiple of
    the coarsening or of the staged block's tile): every point updated, bit for bit against
    the untruncated oracle over 3 steps."""
    cfg = tuned.remainder_cfg(F32C if which.startswith("f32c") else F64, staged=not which.endswith("unstaged"))
    d = cfg["desc"]
    grid = (519, 519, 519)
    print(f"full-size remainder {which} {d.key()} grid {grid}", flush=True)
    got, ref = _run(d, tuple(cfg["block"]), 3, grid)
    ui = np.uint32 if got.dtype == np.float32 else np.uint64
    bad = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert bad.size == 0, f"{bad.size} points differ from the oracle"


@pytest.mark.parametrize("which", ["f64", "f32c"])
def test_production_kernel_bitexact_above_4gib(which):
    """The production kernels on a 1024x1024x512 interior: FP64 buffers of 4.53 GB
    (FP32: 2.3 GB), beyond what 32-bit offsets from the buffer start reach.  Each
    block addresses its own window of planes; one step is bit for bit the oracle's."""
    cfg = tuned.CFG2_F64 if which == "f64" else tuned.CFG2_F32C
    grid = (1032, 1032, 520)
    print(f"above-4GiB {which} {cfg['desc'].key()} grid {grid} block {cfg['block']}", flush=True)
    got, ref = _run(cfg["desc"], tuple(cfg["block"]), 1, grid)
    if which == "f64":
        assert got.nbytes > 4 << 30
    ui = np.uint32 if got.dtype == np.float32 else np.uint64
    bad = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert bad.size == 0, f"{bad.size} points differ from the oracle"


def test_f32c_error_vs_f64_after_20_steps():
    """FP32 storage, compensated accumulation: error of the 20-step state
    against the FP64 result of the same workload (which the test above ties
    bit-for-bit to the oracle), computed in FP64 on the GPU.  Tolerance
    (the accuracy bound for this variant): normwise max |u32 - u64| / max |u64| <= 1e-6."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d64, b64 = tuned.CFG2_F64["desc"], tuple(tuned.CFG2_F64["block"])
    d32, b32 = tuned.CFG2_F32C["desc"], tuple(tuned.CFG2_F32C["block"])
    k64, k32 = rt.Kernel(d64), rt.Kernel(d32)
    g = k64.geometry(GRID)
    mask = rt.mask_seeded(2, d64.n)
    k64.set_mask(mask)
    k32.set_mask(mask)
    g64 = rt.Grids(g, np.float64)
    for i in range(2):
        k64.init(g, g64.ptr(i), rt.INIT_NORMAL, 1)
    gf = k32.geometry(GRID)  # FP32 rows padded to their own 128-byte lines
    g32 = rt.Grids(gf, np.float32)
    for i in range(2):
        # the seeded fill computes each value in FP64 and stores it in the
        # grid's type: the FP32 grid starts from the FP32 rounding of the FP64 one
        k32.init(gf, g32.ptr(i), rt.INIT_NORMAL, 1)
    f64 = k64.run(g, b64, rt.grid_blocks(g, b64), g64.ptr(0), g64.ptr(1), 20)
    f32 = k32.run(gf, b32, rt.grid_blocks(gf, b32), g32.ptr(0), g32.ptr(1), 20)
    torch.cuda.synchronize()
    mx, mb, s2, b2, _ = rt.error_f32_vs_f64(gf, g32.ptr(f32), g, g64.ptr(f64), 1e-3)
    print(f"f32c vs f64 after 20 steps: normwise max {mx / mb:.3e}, rms {np.sqrt(s2 / b2):.3e}", flush=True)
    assert mx / mb <= 1e-6
