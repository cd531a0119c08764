"""Generated kernels vs the CPU oracle, bit for bit, over the option matrix
(2D/3D x Box/Star x Row/Column x Jacobi/Variable x range/coarsening/
alignment) in the three code shapes (plain, stream, stream+DPP), on small
grids with truncated and padded geometry and a few time steps."""
import pytest

from threadcoarsenedstencilgenerator_amd import manifest

from .parity import assert_bitwise, assert_golden, run_case

pytestmark = pytest.mark.gpu

CASES = manifest.parity_cases()


def _id(case):
    d, size = case
    return d.key()


def _blocks(d):
    return manifest.parity_blocks(d)


@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_parity_f64(case):
    d, size = case
    for block in _blocks(d):
        print(f"case {d.key()} size {size} block {block}", flush=True)
        got, ref, _, _ = run_case(d, size, steps=3, block=block)
        assert_bitwise(got, ref, f"{d.key()} block {block}")
        assert_golden(got, d.key())


F32 = manifest.f32c_cases()


@pytest.mark.parametrize("case", F32, ids=[_id(c) for c in F32])
def test_parity_f32c(case):
    d, size = case
    block = manifest.parity_blocks(d)[0]
    print(f"case {d.key()} size {size} block {block}", flush=True)
    got, ref, _, _ = run_case(d, size, steps=3, block=block, mode=2)
    assert_bitwise(got, ref, d.key())
    assert_golden(got, d.key())
