"""STREAM kernels that stage each plane's rows in LDS, on the GPU: bit for bit
against the oracle on small grids (several blocks per axis, partial last blocks,
one and two z slices per block, FP32C, Box and Column cases), and the launch
check that runs such a kernel with the block shape it was generated for only."""
import pytest

from threadcoarsenedstencilgenerator_amd import manifest

from .parity import assert_bitwise, run_case

pytestmark = pytest.mark.gpu

CASES = manifest.staged_cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0].key() for c in CASES])
def test_staged_kernel_bitexact(case):
    d, size, block = case
    print(f"staged {d.key()} size {size} block {block}", flush=True)
    got, ref, _, _ = run_case(d, size, steps=3, block=block, mode=2)
    assert_bitwise(got, ref, f"{d.key()} block {block}")


REMAINDER_CASES = manifest.staged_remainder_cases()


@pytest.mark.parametrize("case", REMAINDER_CASES, ids=[c[0].key() for c in REMAINDER_CASES])
def test_staged_remainder_kernel_bitexact(case):
    """Remainder handling with LDS staging (the last block of rows and of planes starts
    early, the last lanes in x drop their columns past the interior) on grids that are
    multiples of neither the coarsening nor the tile: every point against the untruncated
    oracle over 3 steps."""
    d, size, block = case
    print(f"staged remainder {d.key()} size {size} block {block}", flush=True)
    got, ref, _, _ = run_case(d, size, steps=3, block=block, mode=2)
    assert_bitwise(got, ref, f"{d.key()} block {block}")


def test_staged_remainder_refuses_an_interior_below_one_tile():
    """The runtime refuses a grid whose interior is smaller than one block's tile in y (the
    last block of rows could not start early enough), as TCSG_ERR_ARG, before any launch."""
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, size, block = REMAINDER_CASES[1]
    small = (size[0], 2 * d.range + block[1] * d.yc - 1, size[2])
    print(f"staged remainder {d.key()} refuses size {small}", flush=True)
    k = rt.Kernel(d)
    g = k.geometry(small)
    k.set_mask(rt.mask_seeded(3, d.n))
    grids = rt.Grids(g, k.dtype)
    with pytest.raises(rt.TcsgError) as e:
        k.update(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1))
    assert e.value.status == 1   # TCSG_ERR_ARG
    del grids


def test_staged_kernel_runs_its_block_only():
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, size, block = CASES[1]
    k = rt.Kernel(d)
    g = k.geometry(size)
    k.set_mask(rt.mask_seeded(3, d.n))
    grids = rt.Grids(g, k.dtype)
    for other in ((128, 2, 1), (64, 4, 1), (128, 4, 2), (256, 2, 1)):
        print(f"staged {d.key()} refuses block {other}", flush=True)
        with pytest.raises(rt.TcsgError) as e:
            k.update(g, other, rt.grid_blocks(g, other), grids.ptr(0), grids.ptr(1))
        assert e.value.status == 4   # TCSG_ERR_BLOCK_SHAPE
    k.update(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1))
    del grids


def test_staged_one_step_r1_workload_bitexact():
    """The one-step baseline of the 3D temporal workload (3D Star r=1 on 512^3, LDS
    staging for blocks of 256x2 threads): 2 steps bit for bit against the oracle."""
    from threadcoarsenedstencilgenerator_amd import tuned
    cfg = tuned.STAR3D_R1_ONE_STEP
    print(f"staged one-step {cfg['desc'].key()} size {tuned.STAR3D_R1_GRID}", flush=True)
    got, ref, _, _ = run_case(cfg["desc"], tuned.STAR3D_R1_GRID, steps=2, block=tuple(cfg["block"]), mode=2)
    assert_bitwise(got, ref, cfg["desc"].key())
