"""The emitted STREAM kernels (the exact text compiled for gfx950) run on
the host through the wave-level CPU emulation, bit for bit against the
oracle, with no buffer access that would pass the hardware's per-lane range
check while lying past the buffer.  Covers the x-neighbour exchange modes
(loads, segment DPP, wave-aligned DPP), both dtypes and several block
shapes, including blocks whose rows span or split waves.  The temporal
kernels' phases (load, sub-steps, store; the barriers between them) run
for every thread of a workgroup in turn."""
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
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
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


OFFSET_CASES = manifest.stream_offset_cases()


@pytest.mark.parametrize("case", OFFSET_CASES, ids=[c[0].key() for c in OFFSET_CASES])
def test_emulated_offset_order_stream_kernel_matches_oracle(case):
    """Offset-major statements in a 2D STREAM kernel (each weight used by every output at once,
    rows loaded and x-neighbours lane-shifted where first needed): bit-exact against the oracle
    after 2 steps, which sums each output's terms in the reference's order."""
    d, size, blocks = case
    geo = spec.geometry(d, size)
    og = oracle.Geometry(2, d.range, d.alignment, d.cl_size, size, (1, 1, 1) if d.remainder else d.coarsening)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 5, a, halo=0.375)
    mask = oracle.mask_seeded(7, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2)
    ui = np.uint32 if dt == np.float32 else np.uint64
    for block in blocks:
        nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
        got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk)
        assert bad == 0
        diff = np.flatnonzero(got.view(ui) != ref.view(ui))
        assert diff.size == 0, f"block {block}: {diff.size} points differ, first {diff[:5]}"


@pytest.mark.parametrize("dtype", ["f64", "f32c"])
def test_emulated_packed_batched_kernel_matches_oracle(dtype):
    """Packed tiles (runs of block x * block y consecutive threads of a grid's
    xy thread space, continuing on the next thread row) over uneven grids whose
    thread rows the block's x extent does not divide, with remainder handling
    on every axis: every grid equals the untruncated oracle after 3 steps
    (STREAM, x-neighbours by loads; the CPU emulation covers STREAM kernels)."""
    dt_ = F32C if dtype == "f32c" else spec.F64
    d = manifest.cfg4_desc(shape="stream", batched=True, dtype=dt_).replace(remainder=True, pack=True, zc=8)
    r = d.range
    sizes = [(2 * r + 13, 2 * r + 7, 2 * r + 9), (2 * r + 30, 2 * r + 11, 2 * r + 17), (2 * r + 5, 2 * r + 21, 2 * r + 8),
             (2 * r + 46, 2 * r + 3, 2 * r + 12)]
    dt = np.float32 if dt_ == F32C else np.float64
    run = oracle.run_f32c if dt_ == F32C else oracle.run_f64
    geoms, a0, a1, refs = [], [], [], []
    for i, size in enumerate(sizes):
        geo = spec.geometry(d, size)
        og = oracle.Geometry(3, r, d.alignment, d.cl_size, size, (1, 1, 1))
        a = np.zeros(og.n_points)
        oracle.init_seeded(og, 0, 30 + i, a)
        a = a.astype(dt)
        geoms.append(geo)
        a0.append(a.copy())
        a1.append(a.copy())
        refs.append(run(d, og, None, a.copy(), a.copy(), 3))
    got, bad = emulate.run_batched(d, geoms, a0, a1, None, 3, (32, 4, 1), 7)
    assert bad == 0
    ui = np.uint32 if dt == np.float32 else np.uint64
    for g, ref in zip(got, refs):
        assert np.array_equal(g.view(ui), ref.view(ui))


@pytest.mark.parametrize("dtype", ["f64", "f32c"])
def test_emulated_packed_star_variable_batch_matches_oracle(dtype):
    """Packed tiles with a Variable mask, range 3, Write alignment and coarsening 2x3x5 with
    remainder handling, over grids whose thread rows no block row divides: every grid equals
    the oracle after 2 steps."""
    dt_ = F32C if dtype == "f32c" else spec.F64
    d = manifest.packed_star_desc(dt_)
    r = d.range
    sizes = [(2 * r + 11, 2 * r + 4, 2 * r + 6), (2 * r + 25, 2 * r + 7, 2 * r + 11), (2 * r + 6, 2 * r + 13, 2 * r + 5)]
    dt = np.float32 if dt_ == F32C else np.float64
    run = oracle.run_f32c if dt_ == F32C else oracle.run_f64
    mask = oracle.mask_seeded(8, d.n)
    geoms, a0, a1, refs = [], [], [], []
    for i, size in enumerate(sizes):
        geo = spec.geometry(d, size)
        og = oracle.Geometry(3, r, d.alignment, d.cl_size, size, (1, 1, 1))
        a = np.zeros(og.n_points)
        oracle.init_seeded(og, 1, 40 + i, a, halo=0.25)
        a = a.astype(dt)
        geoms.append(geo)
        a0.append(a.copy())
        a1.append(a.copy())
        refs.append(run(d, og, mask, a.copy(), a.copy(), 2))
    got, bad = emulate.run_batched(d, geoms, a0, a1, mask.astype(dt), 2, (32, 2, 1), 5)
    assert bad == 0
    ui = np.uint32 if dt == np.float32 else np.uint64
    for g, ref in zip(got, refs):
        assert np.array_equal(g.view(ui), ref.view(ui))


def test_pack_refuses_lane_exchange():
    """Packed tiles put lanes of one wave on different thread rows, so code
    shapes that pass x-neighbours between lanes (DPP), LDS staging, 2D and
    unbatched kernels refuse the option."""
    base = manifest.cfg4_desc(shape="stream", batched=True).replace(pack=True)
    assert base.pack
    manifest.cfg4_desc(shape="plain", batched=True).replace(pack=True)
    for bad in (manifest.cfg4_desc(shape="dpp", batched=True), manifest.cfg4_desc(shape="dppw", batched=True),
                manifest.cfg4_desc(shape="stream", batched=False)):
        with pytest.raises(ValueError):
            bad.replace(pack=True)


def _remainder_cases():
    """Grids whose interiors are not multiples of the coarsening on at least
    one axis (every axis for some), generated with remainder handling."""
    sizes = {0: (30, 15, 18), 1: (22, 13, 12), 2: (45, 23), 3: (40, 21), 4: (25, 13, 13)}
    out = []
    for d, size, blocks in CASES:
        i = len(out) % 5
        out.append((d.replace(remainder=True), sizes[i], blocks))
    # four FP64 columns per thread (two 16-byte stores per output row) and three
    # FP64 columns (a 16-byte store and a single one)
    base = dict(variant=spec.STREAM, xcd_remap=True, nt_store=True, remainder=True)
    out.append((StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=2, xc=4, yc=1,
                            zc=2, alignment=ALIGN_WRITE, cl_size=16, xext="dppw", **base),
                (27, 9, 10), [(64, 1, 1), (128, 1, 1)]))
    out.append((StencilDesc(dimension=2, neighbourhood=BOX, orientation=COLUMN, coeff=JACOBI, range=1, xc=3, yc=2,
                            alignment=ALIGN_NONE, cl_size=0, xext="dpp", **base), (33, 16), [(16, 4, 1), (96, 2, 1)]))
    # wave-aligned kernels with paired edge loads over rows of more than two waves
    out += [(d.replace(edge_pairs=True), (size[0] + 150 * d.xc + 1, *size[1:]), blocks)
            for d, size, blocks in out[:len(CASES)] if d.xext == "dppw"]
    return out


REMAINDER_CASES = _remainder_cases()


@pytest.mark.parametrize("case", REMAINDER_CASES, ids=[c[0].key() for c in REMAINDER_CASES])
def test_emulated_remainder_kernel_matches_oracle(case):
    """Every interior point is updated, also on the last thread of an axis
    whose interior is not a multiple of the coarsening: the oracle without
    truncation (coarsening 1) equals the kernel bit for bit, halo included, and
    the untruncated geometry keeps the user's extents."""
    d, size, blocks = case
    geo = spec.geometry(d, size)
    og = oracle.Geometry(d.dimension, d.range, d.alignment, d.cl_size, size, (1, 1, 1))
    assert (og.xdim, og.ydim, og.zdim, og.lo, og.ro) == (geo.xdim, geo.ydim, geo.zdim, geo.lo, geo.ro)
    assert (geo.ydim, geo.zdim) == (tuple(size) + (1,))[1:3]
    assert any(e % c for e, c in zip(geo.elements, d.coarsening))
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
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


def test_emulated_batched_remainder_kernel_matches_oracle():
    """The persistent kernel with remainder handling over grids of arbitrary
    extents (none a multiple of the 2x2x4 coarsening; one interior a single
    column wide) equals the untruncated oracle."""
    d = manifest.cfg4_desc(shape="dppw", batched=True).replace(remainder=True)
    r = d.range
    sizes = [(2 * r + 5, 2 * r + 7, 2 * r + 7), (2 * r + 11, 2 * r + 3, 2 * r + 9), (2 * r + 1, 2 * r + 5, 2 * r + 6)]
    geoms, a0, a1, refs = [], [], [], []
    for i, size in enumerate(sizes):
        geo = spec.geometry(d, size)
        og = oracle.Geometry(3, r, d.alignment, d.cl_size, size, (1, 1, 1))
        a = np.zeros(og.n_points)
        oracle.init_seeded(og, 0, 20 + i, a)
        geoms.append(geo)
        a0.append(a.copy())
        a1.append(a.copy())
        refs.append(oracle.run_f64(d, og, None, a.copy(), a.copy(), 3))
    got, bad = emulate.run_batched(d, geoms, a0, a1, None, 3, (64, 2, 1), 9)
    assert bad == 0
    for g, ref in zip(got, refs):
        assert np.array_equal(g.view(np.uint64), ref.view(np.uint64))


def test_remainder_geometry_limits():
    """Remainder handling keeps the user's extents; interiors narrower than
    the coarsening are accepted in x (the columns past the end are dropped)
    and refused in y and z (the last thread's shifted start would precede
    the interior)."""
    d = manifest.cfg4_desc(shape="dppw").replace(remainder=True)
    r = d.range
    g = spec.geometry(d, (2 * r + 1, 2 * r + 2, 2 * r + 4))
    assert g.elements == (1, 2, 4) and g.threads == (1, 1, 1)
    for size in ((2 * r + 5, 2 * r + 1, 2 * r + 9), (2 * r + 5, 2 * r + 3, 2 * r + 3)):
        with pytest.raises(ValueError):
            spec.geometry(d, size)
    t = spec.geometry(d.replace(remainder=False), (2 * r + 5, 2 * r + 3, 2 * r + 9))
    assert t.elements == (4, 2, 8)
    # PLAIN kernels shift the last thread in x as well
    p = manifest.cfg4_desc(shape="plain").replace(remainder=True)
    assert spec.geometry(p, (2 * r + 3, 2 * r + 3, 2 * r + 5)).threads == (2, 2, 2)
    with pytest.raises(ValueError):
        spec.geometry(p, (2 * r + 1, 2 * r + 2, 2 * r + 4))


# the wave-aligned cases with x widened past two waves of threads, so that waves also meet inside
# the interior and their edge lanes carry the neighbouring wave's columns (the halo is nonzero, so
# the grid's own border counts too); each also with edge_pairs
WIDE_CASES = [(d.replace(edge_pairs=pairs), (size[0] + 150 * d.xc, *size[1:]), blocks)
              for d, size, blocks in CASES if d.xext == "dppw" for pairs in (False, True)]


@pytest.mark.parametrize("case", WIDE_CASES, ids=[c[0].key() for c in WIDE_CASES])
def test_emulated_wide_rows_match_oracle(case):
    """Wave-aligned kernels over rows of several waves equal the oracle, also
    when the edge lanes fetch two rows' extra columns per load (the second
    row's moved into place by a DPP row shift)."""
    d, size, blocks = case
    geo = spec.geometry(d, size)
    og = oracle.Geometry(d.dimension, d.range, d.alignment, d.cl_size, size, d.coarsening)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
    mask = oracle.mask_seeded(6, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2)
    for block in blocks:
        nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
        got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk)
        assert bad == 0
        ui = np.uint32 if dt == np.float32 else np.uint64
        diff = np.flatnonzero(got.view(ui) != ref.view(ui))
        assert diff.size == 0, f"block {block}: {diff.size} points differ, first {diff[:5]}"


TEMPORAL_CASES = manifest.temporal_cases()


@pytest.mark.parametrize("case", TEMPORAL_CASES, ids=[c[0].key() for c in TEMPORAL_CASES])
def test_emulated_temporal_kernel_matches_oracle(case):
    """Two launches of a kernel that advances tsteps steps per launch equal
    2*tsteps steps of the oracle bit for bit, halo and padding included."""
    d, size = case
    geo = spec.geometry(d, size)
    og = oracle.Geometry(2, d.range, d.alignment, d.cl_size, size, (1, 1, 1) if d.remainder else d.coarsening)
    assert (og.xdim, og.ydim) == (geo.xdim, geo.ydim)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
    mask = oracle.mask_seeded(6, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2 * d.tsteps)
    block = (64, d.max_threads // 64, 1)
    nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
    assert nblk[0] > 1 and nblk[1] > 1
    got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk)
    assert bad == 0, f"{bad} LDS accesses outside the tile buffers"
    ui = np.uint32 if dt == np.float32 else np.uint64
    diff = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert diff.size == 0, f"{diff.size} points differ, first {diff[:5]}"


def test_temporal_descriptor_limits():
    T = spec.TEMPORAL
    # 3D: a ring of 2r+1 planes per level s < tsteps, level s covering the tile plus (tsteps-s) r per side
    d3 = StencilDesc(dimension=3, variant=T, range=1, yc=2, zc=128, tsteps=2, max_threads=512)
    assert d3.tile == (64, 16, 68, 20) and d3.lds_bytes == 3 * (68 * 20 + 66 * 18) * 8
    with pytest.raises(ValueError, match="LDS"):
        StencilDesc(dimension=3, variant=T, range=2, yc=4, zc=16, tsteps=3, max_threads=512)
    with pytest.raises(ValueError, match="LDS"):
        StencilDesc(dimension=2, variant=T, range=4, xc=4, yc=8, tsteps=4, max_threads=256)
    with pytest.raises(ValueError, match="temporal variant"):
        StencilDesc(dimension=2, variant=spec.PLAIN, tsteps=2)
    with pytest.raises(ValueError, match="code-shape"):
        StencilDesc(dimension=2, variant=T, tsteps=2, xcd_remap=True)
    d = StencilDesc(dimension=2, neighbourhood=BOX, range=1, yc=8, variant=T, tsteps=4, max_threads=256)
    assert d.tile == (64, 32, 72, 40) and d.lds_bytes == 2 * 72 * 40 * 8
    assert d.params_blob()[13] == 2 and d.params_blob()[18] == 4 << 8


TEMPORAL3D_CASES = manifest.temporal3d_cases()


@pytest.mark.parametrize("case", TEMPORAL3D_CASES, ids=[c[0].key() for c in TEMPORAL3D_CASES])
def test_emulated_temporal3d_kernel_matches_oracle(case):
    """Two launches of a 3D kernel that marches through z advancing tsteps steps
    per launch equal 2*tsteps steps of the oracle bit for bit, halo and padding
    included, with every LDS access inside the workgroup's rings."""
    d, size = case
    geo = spec.geometry(d, size)
    og = oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, (1, 1, 1) if d.remainder else d.coarsening)
    assert (og.xdim, og.ydim, og.zdim) == (geo.xdim, geo.ydim, geo.zdim)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
    mask = oracle.mask_seeded(6, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2 * d.tsteps)
    block = (64, d.max_threads // 64, 1)
    nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
    assert nblk[0] > 1 and nblk[1] > 1 and nblk[2] > 1
    got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk)
    assert bad == 0, f"{bad} LDS accesses outside the rings"
    ui = np.uint32 if dt == np.float32 else np.uint64
    diff = np.flatnonzero(got.view(ui) != ref.view(ui))
    assert diff.size == 0, f"{diff.size} points differ, first {diff[:5]}"


STAGED_CASES = manifest.staged_cases()


def _staged_run(d, size, block, edit=None):
    geo = spec.geometry(d, size)
    og = oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, d.coarsening)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
    mask = oracle.mask_seeded(6, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2)
    nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
    got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk, edit=edit)
    ui = np.uint32 if dt == np.float32 else np.uint64
    return np.flatnonzero(got.view(ui) != ref.view(ui)), bad, nblk, geo


@pytest.mark.parametrize("case", STAGED_CASES, ids=[c[0].key() for c in STAGED_CASES])
def test_emulated_staged_kernel_matches_oracle(case):
    """Kernels that stage each plane's rows in LDS (the waves of a workgroup run as host
    threads meeting at every barrier) equal the oracle bit for bit over 2 steps, with
    every LDS access inside the tile, of an element written before the last barrier and
    not written again before the next one."""
    d, size, block = case
    diff, bad, nblk, geo = _staged_run(d, size, block)
    # several blocks per axis, the last one partly outside the interior where a block has several threads
    assert all(n > 1 for n in nblk) and all(nblk[i] * block[i] > geo.threads[i] for i in range(3) if block[i] > 1)
    assert bad == 0, f"{bad} bad accesses"
    assert diff.size == 0, f"{diff.size} points differ, first {diff[:5]}"


STAGED_REMAINDER_CASES = manifest.staged_remainder_cases()


def _staged_remainder_run(d, size, block, edit=None):
    geo = spec.geometry(d, size)
    og = oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, (1, 1, 1))
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1, 4, a, halo=0.375)
    mask = oracle.mask_seeded(6, d.n)
    dt = np.float32 if d.dtype == F32C else np.float64
    a = a.astype(dt)
    run = oracle.run_f32c if d.dtype == F32C else oracle.run_f64
    ref = run(d, og, mask, a.copy(), a.copy(), 2)
    nblk = tuple((geo.threads[i] + block[i] - 1) // block[i] for i in range(3))
    got, bad = emulate.run(d, geo, mask.astype(dt), a.copy(), a.copy(), 2, block, nblk, edit=edit)
    ui = np.uint32 if dt == np.float32 else np.uint64
    return np.flatnonzero(got.view(ui) != ref.view(ui)), bad, nblk, geo


@pytest.mark.parametrize("case", STAGED_REMAINDER_CASES, ids=[c[0].key() for c in STAGED_REMAINDER_CASES])
def test_emulated_staged_remainder_kernel_matches_oracle(case):
    """LDS-staging kernels with remainder handling on grids whose interiors are multiples of
    neither the coarsening nor the block's tile: every interior point equals the untruncated
    oracle over 2 steps, every LDS access ordered by the barriers."""
    d, size, block = case
    diff, bad, nblk, geo = _staged_remainder_run(d, size, block)
    assert any(geo.elements[i] % (block[i] * d.coarsening[i]) for i in (1, 2))
    assert bad == 0, f"{bad} bad accesses"
    assert diff.size == 0, f"{diff.size} points differ, first {diff[:5]}"


def test_emulated_staged_remainder_needs_the_block_shift():
    """Without the shift of the last block of rows (its tile then runs past the interior),
    points differ from the oracle."""
    d, size, block = STAGED_REMAINDER_CASES[1]
    ty = block[1] * d.yc

    def edit(text):
        old = f"tcsg_min(byi * {ty}u, ydim - {2 * d.range + ty}u)"
        assert old in text
        return text.replace(old, f"(byi * {ty}u)")
    diff, bad, _, _ = _staged_remainder_run(d, size, block, edit=("noshift", edit))
    assert diff.size > 0


def _drop_barrier(k):
    def edit(text):
        lines = text.split("\n")
        at = [i for i, l in enumerate(lines) if l.strip() == "__syncthreads();"]
        del lines[at[k]]
        return "\n".join(lines)
    return edit


@pytest.mark.parametrize("k", [0, 3])
def test_emulated_staging_without_a_barrier_is_flagged(k):
    """With one barrier of a staging kernel removed, a plane's buffer is written while
    other waves still read it (or read before it is written): the emulation counts
    those accesses."""
    d, size, block = STAGED_CASES[0]
    diff, bad, _, _ = _staged_run(d, size, block, edit=(f"nobarrier{k}", _drop_barrier(k)))
    assert bad > 0


def test_stage_descriptor_limits():
    base = manifest.cfg2_desc(2, 4, 8, shape="dppw", max_threads=512).replace(prefetch=0)
    d = base.replace(stage=[128, 4, 1])
    assert d.stage == (128, 4, 1) and d.staged
    assert d.stage_lds_bytes == 2 * (4 * 4 + 8) * 128 * 2 * 8
    # bit 3 marks staging; the block in bits 16-20 (bx/64), 21-25, 26-30
    f = d.params_blob()[18]
    assert f & 8 and (f >> 16) & 31 == 2 and (f >> 21) & 31 == 4 and (f >> 26) & 31 == 1
    assert not base.staged and base.params_blob()[18] & 8 == 0
    # kernels without staging keep their names
    assert base.key() == base.replace(stage=(0, 0, 0)).key() != d.key()
    for bad in (dict(stage=(96, 2, 1)), dict(stage=(128, 4, 2)), dict(stage=(128, 4))):
        with pytest.raises(ValueError):
            base.replace(**bad)
    with pytest.raises(ValueError, match="LDS"):
        base.replace(stage=(256, 2, 1), yc=8)                 # 2 * 24 * 512 * 8 B > 160 KiB
    for other in (dict(xext="dpp"), dict(prefetch=1), dict(batched=True)):
        with pytest.raises(ValueError, match="staging"):
            base.replace(stage=(64, 2, 1), **other)
    with pytest.raises(ValueError, match="staging"):
        manifest.cfg1_desc("dppw").replace(stage=(64, 2, 1))   # 2D
    from threadcoarsenedstencilgenerator_amd.generator.emit import emit_source
    with pytest.raises(ValueError, match="block shape"):
        emit_source(d, plugin=True)
