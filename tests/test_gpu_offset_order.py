"""2D STREAM kernels with offset-major statements (each weight used by every output of
the thread at once; rows loaded and x-neighbours lane-shifted where first needed) on the
GPU: bit-exact against the oracle, which sums each output's terms in the reference's
order (hip/generate2DStencil_hip.py:245-298)."""
import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import manifest, spec
from threadcoarsenedstencilgenerator_amd.spec import F32C

pytestmark = pytest.mark.gpu

CASES = manifest.stream_offset_cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0].key() for c in CASES])
def test_offset_order_stream_kernel_bitexact(case):
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, size, blocks = case
    k = rt.Kernel(d)
    mask = rt.mask_seeded(7, d.n)
    if d.coeff == spec.VARIABLE:
        k.set_mask(mask)
    g = k.geometry(size)
    gr = rt.Grids(g, k.dtype)
    k.init(g, gr.ptr(0), rt.INIT_NORMAL, 5)
    k.init(g, gr.ptr(1), rt.INIT_NORMAL, 5)
    torch.cuda.synchronize()
    a = gr.host(0)
    og = oracle.Geometry(2, d.range, d.alignment, d.cl_size, size, (1, 1, 1) if d.remainder else d.coarsening)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 3)
    ui = np.uint32 if d.dtype == F32C else np.uint64
    for block in blocks:
        print(f"offset order {d.key()} block {block}", flush=True)
        gr.upload(0, a)
        gr.upload(1, a)
        fin = k.run(g, block, rt.grid_blocks(g, block), gr.ptr(0), gr.ptr(1), 3)
        torch.cuda.synchronize()
        got = gr.host(fin)
        diff = np.flatnonzero(got.view(ui) != ref.view(ui))
        assert diff.size == 0, f"block {block}: {diff.size} points differ, first {diff[:5]}"
