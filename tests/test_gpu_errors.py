"""Error behaviour of the C ABI on the GPU: every refused call returns its
status code before anything is launched (no kernel runs on a bad argument),
and a refused call leaves the library usable: the next valid run is still
bit-exact against the oracle.  The reference's generated interface checks
its compiled-in parameters (hip/generateLaunchUpdateGrid.py:106-119) and
nothing else; these are the checks that replace its unchecked launches."""
import os

import numpy as np
import pytest

from threadcoarsenedstencilgenerator_amd import build, manifest, spec, tuned

from .parity import assert_bitwise, run_case

pytestmark = pytest.mark.gpu

OK, ARG, HIP, MISMATCH, BLOCK_SHAPE, TOO_LARGE, NO_MASK, FILE = range(8)


def _status(fn):
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    with pytest.raises(rt.TcsgError) as e:
        fn()
    return e.value.status


def test_load_errors():
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d = tuned.CFG2_F64["desc"]
    assert _status(lambda: rt.Kernel(d, os.path.join(build.KDIR, "no_such_kernel.co"))) == FILE
    # a code object whose baked-in parameters differ from the descriptor asked for
    other = manifest.cfg1_desc("plain")
    assert _status(lambda: rt.Kernel(d, build.kernel_path(other))) == MISMATCH


def test_launch_argument_errors():
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, block = tuned.CFG2_F64["desc"], tuple(tuned.CFG2_F64["block"])
    k = rt.Kernel(d)
    size = (72, 36, 136)
    g = k.geometry(size)
    grids = rt.Grids(g, k.dtype)
    blocks = rt.grid_blocks(g, block)
    p0, p1 = grids.ptr(0), grids.ptr(1)
    print(f"errors on {d.key()} size {size}", flush=True)
    # Variable kernel before its weights are set
    assert _status(lambda: k.update(g, block, blocks, p0, p1)) == NO_MASK
    assert _status(lambda: k.set_mask(np.ones(d.n - 1) / d.n)) == ARG
    k.set_mask(rt.mask_seeded(3, d.n))
    # block rows that are not whole waves, and more threads than the launch bound
    assert _status(lambda: k.update(g, (96, 2, 1), rt.grid_blocks(g, (96, 2, 1)), p0, p1)) == BLOCK_SHAPE
    over = (128, 2 * d.max_threads // 128, 1)
    assert _status(lambda: k.update(g, over, rt.grid_blocks(g, over), p0, p1)) == BLOCK_SHAPE
    # a grid that does not cover the interior, the same buffer as input and output
    short = (blocks[0], blocks[1], blocks[2] - 1)
    assert _status(lambda: k.update(g, block, short, p0, p1)) == ARG
    assert _status(lambda: k.update(g, block, blocks, p0, p0)) == ARG
    # a geometry this kernel's coarsening does not produce (another kernel's truncation)
    g_other = rt.Kernel(manifest.cfg2_desc(1, 1, 1)).geometry((72, 37, 139))
    assert _status(lambda: k.update(g_other, block, rt.grid_blocks(g_other, block), p0, p1)) == ARG
    torch.cuda.synchronize()
    del grids


def test_grid_limits_refused_before_launch():
    """More than 65535 blocks in y is refused as a block-shape error (the
    launch API's limit) instead of failing inside the runtime."""
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d = manifest.cfg1_desc("plain")
    k = rt.Kernel(d)
    g = k.geometry((20, 140010))
    grids = rt.Grids(g, k.dtype)
    block = (16, 1, 1)
    blocks = rt.grid_blocks(g, block)
    assert blocks[1] > 65535
    assert _status(lambda: k.update(g, block, blocks, grids.ptr(0), grids.ptr(1))) == BLOCK_SHAPE
    del grids


def test_valid_run_after_refused_calls():
    """After the refusals above in this process, a valid run is bit-exact."""
    d, size = manifest.parity_cases()[-1]
    block = manifest.parity_blocks(d)[0]
    got, ref, _, _ = run_case(d, size, steps=2, block=block)
    assert_bitwise(got, ref, d.key())
