"""Temporal kernels on the GPU: each launch advances tsteps time steps with the
tile and its tsteps*r halo staged in LDS (no reference counterpart; the
reference runs one step per launch, hip/generateLaunchUpdateGrid.py:129-139).
k launches equal k*tsteps steps of the oracle bit for bit, on the small-grid
matrix and on the 16384^2 2D Box r=1 workload, and equal k*tsteps launches of
the one-step kernel of the same stencil."""
import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import manifest, spec, tuned

pytestmark = pytest.mark.gpu


def _run(d, size, launches, mode=1, seed=4, halo=0.375):
    """(final grid from the GPU, oracle's after launches*tsteps steps)."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    k = rt.Kernel(d)
    assert k.steps_per_launch == d.tsteps
    mask = oracle.mask_seeded(6, d.n)
    if d.coeff == spec.VARIABLE:
        k.set_mask(mask)
    g = k.geometry(size)
    og = oracle.Geometry(d.dimension, d.range, d.alignment, d.cl_size, size,
                         (1, 1, 1) if d.remainder else d.coarsening)
    assert (og.xdim, og.ydim, og.zdim) == (g.size.x, g.size.y, g.size.z)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, mode, seed, a, halo=halo)
    dt = np.float32 if d.dtype == spec.F32C else np.float64
    a = a.astype(dt)
    grids = rt.Grids(g, k.dtype)
    grids.upload(0, a)
    grids.upload(1, a)
    block = (64, d.max_threads // 64, 1)
    fin = k.run(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1), launches)
    torch.cuda.synchronize()
    got = grids.host(fin)
    run = oracle.run_f32c if d.dtype == spec.F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), launches * d.tsteps)
    del grids
    k.close()
    return got, ref


@pytest.mark.parametrize("case", manifest.temporal_cases(), ids=[c[0].key() for c in manifest.temporal_cases()])
def test_temporal_kernel_matches_oracle(case):
    d, size = case
    print(f"temporal {d.key()} size {size}", flush=True)
    got, ref = _run(d, size, 3)
    ui = np.uint32 if d.dtype == spec.F32C else np.uint64
    bad = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert bad.size == 0, f"{bad.size} points differ, first {bad[:5]}"


@pytest.mark.parametrize("case", manifest.temporal3d_cases(), ids=[c[0].key() for c in manifest.temporal3d_cases()])
def test_temporal3d_kernel_matches_oracle(case):
    """3D kernels marching through z, tsteps steps per launch: 3 launches equal
    3*tsteps oracle steps bit for bit on grids of several tiles per axis."""
    d, size = case
    print(f"temporal 3D {d.key()} size {size}", flush=True)
    got, ref = _run(d, size, 3)
    ui = np.uint32 if d.dtype == spec.F32C else np.uint64
    bad = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert bad.size == 0, f"{bad.size} points differ, first {bad[:5]}"


def test_temporal_refuses_other_block_shapes():
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, _ = manifest.temporal_cases()[1]
    k = rt.Kernel(d)
    assert k.block_ok((64, d.max_threads // 64, 1))
    assert not k.block_ok((128, d.max_threads // 128, 1)) and not k.block_ok((64, 1, 1))
    k.close()


def test_temporal_workload_matches_oracle_and_one_step_kernel():
    """The tuned temporal kernel on the 16384^2 2D Box Jacobi r=1 grid: two
    launches equal the oracle and the reference-shaped one-step kernel (PLAIN,
    no lane exchange) bit for bit."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    cfg = tuned.BOX2D_R1_TEMPORAL
    d = cfg["desc"]
    size = tuned.BOX2D_R1_GRID
    print(f"temporal workload {d.key()} size {size}", flush=True)
    got, ref = _run(d, size, 2, mode=0, seed=9, halo=0.0)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    one = dict(desc=manifest.cfg1_desc("plain"), block=(64, 4, 1))
    k = rt.Kernel(one["desc"])
    g = k.geometry(size)
    grids = rt.Grids(g, k.dtype)
    og = oracle.Geometry(2, 1, 0, 0, size, (1, 1, 1))
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 0, 9, a)
    grids.upload(0, a)
    grids.upload(1, a)
    b = tuple(one["block"])
    fin = k.run(g, b, rt.grid_blocks(g, b), grids.ptr(0), grids.ptr(1), 2 * d.tsteps)
    torch.cuda.synchronize()
    one_got = grids.host(fin)
    bad = np.flatnonzero(one_got.view(np.uint64) != got.view(np.uint64))
    xs, ys = bad % g.size.x, bad // g.size.x
    assert bad.size == 0, (f"PLAIN one-step kernel differs from the temporal one (= the oracle) at {bad.size} points; "
                           f"x in [{xs.min()}, {xs.max()}], y in [{ys.min()}, {ys.max()}], first (x, y) "
                           f"{list(zip(xs[:6].tolist(), ys[:6].tolist()))}")
    del grids
    k.close()


def test_reference_timing_protocol_counts_steps():
    """tcsg_benchmark_update_grid (the reference's doubling rounds) on a
    temporal kernel: each launch contributes steps_per_launch per-step
    durations, equal shares of the launch, so the reference's Lup/s per
    duration stays a per-step rate."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, size = manifest.temporal_cases()[2]      # 4 steps per launch
    k = rt.Kernel(d)
    k.set_mask(oracle.mask_seeded(6, d.n))
    g = k.geometry(size)
    grids = rt.Grids(g, k.dtype)
    block = (64, d.max_threads // 64, 1)
    durs = k.benchmark(g, block, rt.grid_blocks(g, block), 0.0, grids.ptr(0), grids.ptr(1))
    torch.cuda.synchronize()
    assert len(durs) == d.tsteps and d.tsteps == 4   # one round of one launch
    assert all(x == durs[0] and x > 0 for x in durs)
    del grids
    k.close()


def test_temporal_f32c_workload_matches_oracle():
    """The FP32-storage temporal kernel of the 16384^2 workload: one launch
    equals the FP32 oracle (same compensated summation) bit for bit."""
    d = tuned.BOX2D_R1_TEMPORAL_F32C["desc"]
    print(f"temporal workload {d.key()} size {tuned.BOX2D_R1_GRID}", flush=True)
    got, ref = _run(d, tuned.BOX2D_R1_GRID, 1, mode=0, seed=9, halo=0.0)
    bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, f"{bad.size} points differ, first {bad[:5]}"


def test_temporal3d_workload_matches_oracle_and_one_step_kernel():
    """The tuned 3D temporal kernel on the 512^3 3D Star r=1 Variable grid: two
    launches (4 steps) equal the oracle and 4 launches of the one-step STREAM kernel
    of the same stencil bit for bit."""
    import torch

    import oracle
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    cfg, one = tuned.STAR3D_R1_TEMPORAL, tuned.STAR3D_R1_ONE_STEP
    size = tuned.STAR3D_R1_GRID
    mask = rt.mask_seeded(2, cfg["desc"].n)
    finals = []
    a0 = None
    for c, launches in ((cfg, 2), (one, 4)):
        d, b = c["desc"], tuple(c["block"])
        print(f"temporal 3D workload {d.key()} size {size} block {b} launches {launches}", flush=True)
        k = rt.Kernel(d)
        k.set_mask(mask)
        g = k.geometry(size)
        grids = rt.Grids(g, k.dtype)
        for i in range(2):
            k.init(g, grids.ptr(i), rt.INIT_NORMAL, 3)
        torch.cuda.synchronize()
        if a0 is None:
            a0 = grids.host(0)
        fin = k.run(g, b, rt.grid_blocks(g, b), grids.ptr(0), grids.ptr(1), launches)
        torch.cuda.synchronize()
        finals.append(grids.host(fin))
        del grids
        k.close()
        torch.cuda.empty_cache()
    og = oracle.Geometry(3, 1, 0, 0, size, (1, 1, 1))
    ref = oracle.run_f64(cfg["desc"], og, mask, a0.copy(), a0.copy(), 4)
    assert np.array_equal(finals[0].view(np.uint64), ref.view(np.uint64))
    assert np.array_equal(finals[1].view(np.uint64), finals[0].view(np.uint64))
