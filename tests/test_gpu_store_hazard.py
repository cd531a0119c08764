"""STREAM kernels whose earlier builds stored wrong values in a varying number of
points (a 16-byte buffer store followed at once by a VALU write of its data
registers): the 2D Box Jacobi r=1 2x4 kernels on the 16384^2 grid and the FP32
kernels with four points per lane in x on the 512^3 grid.  Built with the
soffset-0 store form (the compiler inserts the wait states, and the build checks
every code object), each must equal the oracle bit for bit over the whole grid."""
import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import manifest, spec

pytestmark = pytest.mark.gpu


def _run(d, size, block, steps, mode, seed):
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    k = rt.Kernel(d)
    g = k.geometry(size)
    dim = d.dimension
    og = oracle.Geometry(dim, d.range, d.alignment, d.cl_size, size, (1, 1, 1) if d.remainder else d.coarsening)
    mask = oracle.mask_seeded(3, d.n)
    if d.coeff == spec.VARIABLE:
        k.set_mask(mask)
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), mode, seed)
    torch.cuda.synchronize()
    a0 = grids.host(0)
    fin = k.run(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1), steps)
    torch.cuda.synchronize()
    got = grids.host(fin)
    del grids
    torch.cuda.empty_cache()
    k.close()
    run = oracle.run_f32c if d.dtype == spec.F32C else oracle.run_f64
    ref = run(d, og, mask, a0.copy(), a0.copy(), steps)
    return got, ref, (og.xdim, og.ydim)


@pytest.mark.parametrize("case", manifest.store_hazard_cases(), ids=lambda c: c[0])
def test_wide_store_kernels_bitexact_full_grid(case):
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    name, d, size, block, steps = case
    print(f"store-hazard case {name}: {d.key()} grid {size} block {block} steps {steps}", flush=True)
    mode = rt.INIT_U01 if d.dimension == 2 else rt.INIT_NORMAL
    got, ref, (xdim, ydim) = _run(d, size, block, steps, mode, 5)
    ui = np.uint32 if got.dtype == np.float32 else np.uint64
    bad = np.flatnonzero(got.view(ui) != ref.view(ui))
    if bad.size:
        x = bad % xdim
        y = (bad // xdim) % ydim
        pytest.fail(f"{bad.size} points differ; x mod 64 in {sorted(set((x % 64).tolist()))[:20]}, "
                    f"first (x, y) {list(zip(x[:6].tolist(), y[:6].tolist()))}")
