"""Remainder handling on the GPU: STREAM kernels generated for grids whose
interiors are not multiples of the coarsening update every interior point,
bit for bit against the oracle on the untruncated grid, while the kernels
without it keep the reference's truncation (hip/benchmark.cpp:139-161)."""
import numpy as np
import pytest

from threadcoarsenedstencilgenerator_amd import manifest, spec

from .parity import assert_bitwise, assert_golden, run_case

pytestmark = pytest.mark.gpu

CASES = manifest.remainder_cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0].key() for c in CASES])
def test_remainder_parity(case):
    d, size = case
    mode = 2 if d.dtype == spec.F32C else None
    for block in manifest.parity_blocks(d):
        print(f"case {d.key()} size {size} block {block}", flush=True)
        got, ref, _, og = run_case(d, size, steps=3, block=block, mode=mode)
        assert og.interior_points == int(np.prod([s - 2 * d.range for s in size]))
        assert_bitwise(got, ref, f"{d.key()} block {block}")
        assert_golden(got, d.key())


def test_geometry_of_a_remainder_kernel():
    """The loaded kernel reports the option and its geometry is the
    generator's untruncated one; the same shape without the option truncates
    and refuses the untruncated geometry."""
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, size = CASES[0]
    k = rt.Kernel(d)
    assert k.remainder
    g = k.geometry(size)
    sg = spec.geometry(d, size)
    assert (g.size.x, g.size.y, g.size.z) == (sg.xdim, sg.ydim, sg.zdim)
    assert g.threads.tup() == sg.threads and g.elements.tup() == sg.elements
    plain = rt.Kernel(d.replace(remainder=False))
    assert not plain.remainder
    gt = plain.geometry(size)
    assert gt.elements.tup() != g.elements.tup()
    grids = rt.Grids(g, k.dtype)
    if d.coeff == spec.VARIABLE:
        plain.set_mask(rt.mask_seeded(3, d.n))
    block = manifest.parity_blocks(d)[0]
    with pytest.raises(rt.TcsgError):
        plain.update(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1))
