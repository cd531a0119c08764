"""The emitted STREAM kernels (the exact text compiled for gfx950) run on
the host through the wave-level CPU emulation, bit for bit against the
oracle, with no buffer access that would pass the hardware's per-lane range
check while lying past the buffer.  Covers the x-neighbour exchange modes
(loads, segment DPP, wave-aligned DPP), both dtypes and several block
shapes, including blocks whose rows span or split waves."""
import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import manifest, spec
from threadcoarsenedstencilgenerator_amd.generator import emulate
from threadcoarsenedstencilgenerator_amd.spec import (ALIGN_NONE, ALIGN_READ, ALIGN_WRITE, BOX, COLUMN, F32C,
                                                      JACOBI, ROW, STAR, VARIABLE, StencilDesc)


def _cases():
    out = []
    base = dict(variant=spec.STREAM, xcd_remap=True, nt_store=True)
    for xext in ("load", "dpp", "dppw"):
        blocks = [(64, 1, 1), (128, 2, 1)] if xext == "dppw" else [(16, 4, 1), (96, 2, 1)]
        out += [
            (StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=4, xc=1, yc=2,
                         zc=3, alignment=ALIGN_WRITE, cl_size=16, xext=xext, **base), (30, 14, 17), blocks),
            (StencilDesc(dimension=3, neighbourhood=BOX, orientation=COLUMN, coeff=JACOBI, range=2, xc=2, yc=2,
                         zc=2, alignment=ALIGN_NONE, cl_size=0, xext=xext, zfirst=True, **base), (21, 12, 11),
             blocks),
            (StencilDesc(dimension=2, neighbourhood=BOX, orientation=ROW, coeff=VARIABLE, range=3, xc=2, yc=3,
                         alignment=ALIGN_READ, cl_size=8, xext=xext, prefetch=2, **base), (45, 23), blocks),
            (StencilDesc(dimension=2, neighbourhood=STAR, orientation=COLUMN, coeff=JACOBI, range=3, xc=1, yc=2,
                         alignment=ALIGN_WRITE, cl_size=4, xext=xext, **base), (40, 19), blocks),
            (StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=2, xc=2, yc=2,
                         zc=2, alignment=ALIGN_WRITE, cl_size=16, dtype=F32C, comp_group=3, xext=xext, **base),
             (24, 12, 12), blocks),
        ]
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0].key() for c in CASES])
def test_emulated_stream_kernel_matches_oracle(case):
    d, size, blocks = case
    geo = spec.geometry(d, size)
    og = oracle.Geometry(d.dimension, d.range, d.alignment, d.cl_size, size, d.coarsening)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a)
    mask = oracle.mask_seeded(6, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2)
    for block in blocks:
        nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
        got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk)
        assert bad == 0, f"{bad} accesses whose full offset does not fit in 32 bits"
        ui = np.uint32 if dt == np.float32 else np.uint64
        diff = np.flatnonzero(got.view(ui) != ref.view(ui))
        assert diff.size == 0, f"block {block}: {diff.size} points differ, first {diff[:5]}"


@pytest.mark.parametrize("dtype", ["f64", "f32c"])
@pytest.mark.parametrize("nwg", [5, 11])
def test_emulated_batched_kernel_matches_oracle(dtype, nwg):
    """The persistent kernel over a tile table of uneven grids (fewer
    workgroups than tiles, so each workgroup loops) equals the oracle per grid:
    with fewer than 8 workgroups, which stride the whole table, and with 11,
    which split it into eight per-XCD ranges of unequal workgroup counts."""
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d = manifest.cfg4_desc(shape="dppw", batched=True, dtype=F32C if dtype == "f32c" else spec.F64)
    sizes = rt.batch_sizes(4, 7, 8, 24, d.coarsening, d.range)
    dt = np.float32 if d.dtype == F32C else np.float64
    geoms, a0, a1, refs = [], [], [], []
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    for i, size in enumerate(sizes):
        geo = spec.geometry(d, size)
        og = oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, d.coarsening)
        a = np.zeros(og.n_points)
        oracle.init_seeded(og, 0, 10 + i, a)
        a = a.astype(dt)
        geoms.append(geo)
        a0.append(a.copy())
        a1.append(a.copy())
        refs.append(run(d, og, None, a.copy(), a.copy(), 3))
    got, bad = emulate.run_batched(d, geoms, a0, a1, None, 3, (64, 2, 1), nwg)
    assert bad == 0
    for g, r in zip(got, refs):
        ui = np.uint32 if dt == np.float32 else np.uint64
        assert np.array_equal(g.view(ui), r.view(ui))
