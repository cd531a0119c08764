"""The kernels built ahead of time (in-tree code objects under kernels/).

Generated kernels are compile-time specialisations; the parity suite and the
benchmark workloads use the sets below, which ``__graft_entry__.build()``
compiles with hipcc for gfx950.
"""
from __future__ import annotations

import itertools

from .spec import (ALIGN_NONE, ALIGN_READ, ALIGN_WRITE, BOX, COLUMN, F32C, F64, JACOBI, PLAIN, ROW,
                   STAR, STREAM, TEMPORAL, VARIABLE, StencilDesc)

# code shapes every parity case is built in
SHAPES = {
    "plain": dict(variant=PLAIN),
    "plain_offset": dict(variant=PLAIN, order="offset"),
    "stream": dict(variant=STREAM, xext="load", xcd_remap=True, nt_store=True),
    "dpp": dict(variant=STREAM, xext="dpp", xcd_remap=True, nt_store=True),
    "dppw": dict(variant=STREAM, xext="dppw", xcd_remap=True, nt_store=True, zfirst=True),
}
# block shapes each code shape is tested with (dppw needs x extents that are multiples of 64)
PARITY_BLOCKS = {
    2: {"any": [(16, 4, 1), (64, 2, 1)], "dppw": [(64, 2, 1), (128, 1, 1)]},
    3: {"any": [(8, 4, 2), (64, 1, 2)], "dppw": [(64, 1, 2), (128, 2, 1)]},
}


def parity_blocks(desc):
    return PARITY_BLOCKS[desc.dimension]["dppw" if desc.xext == "dppw" else "any"]


def parity_cases():
    """(desc, grid size incl. halo) pairs covering the option matrix on small grids."""
    out = []
    for dim, nb, orient, coeff in itertools.product((2, 3), (BOX, STAR), (ROW, COLUMN), (JACOBI, VARIABLE)):
        for shape in SHAPES.values():
            # two parameter points per neighbourhood/orientation/coefficient cell
            pts = [
                dict(range=1, xc=1, yc=1, zc=1, alignment=ALIGN_NONE, cl_size=0),
                dict(range=2, xc=2, yc=3, zc=2, alignment=ALIGN_WRITE, cl_size=4),
            ]
            if nb == STAR:
                pts.append(dict(range=3, xc=1, yc=2, zc=3, alignment=ALIGN_READ, cl_size=8))
                # edge-lane columns left of the first interior column (r=3, xc=2: H*xc > r)
                pts.append(dict(range=3, xc=2, yc=1, zc=2, alignment=ALIGN_NONE, cl_size=0, prefetch=2))
            for p in pts:
                if dim == 2:
                    p = dict(p, zc=1)
                d = StencilDesc(dimension=dim, neighbourhood=nb, orientation=orient, coeff=coeff, **p, **shape)
                size = (37, 29) if dim == 2 else (23, 19, 17)
                out.append((d, size))
    return out


def f32c_cases():
    out = []
    for dim, nb, coeff, orient in ((3, STAR, VARIABLE, ROW), (2, BOX, VARIABLE, COLUMN), (3, BOX, JACOBI, ROW),
                                   (2, STAR, JACOBI, ROW)):
        for shape in ("plain", "plain_offset", "dpp", "dppw"):
            d = StencilDesc(dimension=dim, neighbourhood=nb, orientation=orient, coeff=coeff, range=2, xc=2,
                            yc=2, zc=2 if dim == 3 else 1, alignment=ALIGN_WRITE, cl_size=16, dtype=F32C,
                            **SHAPES[shape])
            out.append((d, (45, 33) if dim == 2 else (27, 21, 19)))
    return out


def remainder_cases():
    """Kernels with remainder handling on grids whose interiors are not
    multiples of the coarsening (every parity size is ragged for these
    coarsenings): both DPP code shapes and both PLAIN statement orders, plus
    FP32C cases."""
    out = []
    for d, size in parity_cases():
        if (d.variant == STREAM and d.xext == "load") or d.coarsening == (1, 1, 1):
            continue
        if d.range == 3 and d.xc == 1:
            continue
        out.append((d.replace(remainder=True), size))
    for d, size in f32c_cases():
        if d.variant == STREAM or d.order == "point":
            out.append((d.replace(remainder=True), (size[0] + 1, size[1] + 2) + tuple(s + 1 for s in size[2:])))
    # four FP64 columns per thread (two 16-byte stores per output row) and three FP64 columns
    out.append((StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=2, xc=4, yc=1,
                            zc=2, alignment=ALIGN_WRITE, cl_size=16, remainder=True, **SHAPES["dppw"]), (27, 9, 10)))
    out.append((StencilDesc(dimension=2, neighbourhood=BOX, orientation=COLUMN, coeff=JACOBI, range=1, xc=3, yc=2,
                            alignment=ALIGN_NONE, cl_size=0, remainder=True, **SHAPES["dpp"]), (33, 16)))
    return out


# ---- benchmark workloads ----
def temporal_cases():
    """Temporal kernels (tsteps per launch through LDS) with a small grid each:
    both neighbourhoods, coefficient kinds, orientations, dtypes and
    alignments; ranges 1-3; 1-4 steps per launch; grids of several tiles whose
    interiors are not multiples of the tile, with and without the reference's
    truncation."""
    base = dict(dimension=2, variant=TEMPORAL)
    return [
        (StencilDesc(neighbourhood=BOX, coeff=JACOBI, range=1, xc=1, yc=2, tsteps=3, max_threads=128,
                     remainder=True, **base), (150, 41)),
        (StencilDesc(neighbourhood=STAR, coeff=VARIABLE, range=2, xc=2, yc=1, tsteps=2, max_threads=256,
                     alignment=ALIGN_WRITE, cl_size=8, remainder=True, **base), (301, 23)),
        (StencilDesc(neighbourhood=BOX, orientation=COLUMN, coeff=VARIABLE, range=3, xc=1, yc=1, tsteps=4,
                     max_threads=64, alignment=ALIGN_READ, cl_size=16, **base), (97, 30)),
        (StencilDesc(neighbourhood=STAR, orientation=COLUMN, coeff=JACOBI, range=1, xc=1, yc=4, tsteps=1,
                     max_threads=192, remainder=True, **base), (70, 57)),
        (StencilDesc(neighbourhood=BOX, coeff=VARIABLE, range=2, xc=1, yc=3, tsteps=3, max_threads=128,
                     dtype=F32C, comp_group=7, remainder=True, **base), (140, 37)),
        (StencilDesc(neighbourhood=STAR, coeff=JACOBI, range=1, xc=2, yc=2, tsteps=4, max_threads=256,
                     dtype=F32C, **base), (260, 45)),
    ]


def temporal3d_cases():
    """3D temporal kernels (tsteps per launch, marching through z with a ring of
    planes per step) with a small grid each: both neighbourhoods, coefficient
    kinds, orientations, dtypes and alignments; ranges 1-2; 1-3 steps per launch;
    several tiles in x, y and z, with and without the reference's truncation."""
    base = dict(dimension=3, variant=TEMPORAL)
    return [
        (StencilDesc(neighbourhood=STAR, coeff=VARIABLE, range=1, xc=1, yc=2, zc=6, tsteps=2, max_threads=128,
                     remainder=True, **base), (150, 13, 20)),
        (StencilDesc(neighbourhood=BOX, orientation=COLUMN, coeff=JACOBI, range=1, xc=1, yc=1, zc=4, tsteps=2,
                     max_threads=256, alignment=ALIGN_WRITE, cl_size=8, **base), (140, 11, 15)),
        (StencilDesc(neighbourhood=STAR, coeff=JACOBI, range=1, xc=1, yc=2, zc=5, tsteps=3, max_threads=128,
                     alignment=ALIGN_READ, cl_size=16, remainder=True, **base), (135, 14, 23)),
        (StencilDesc(neighbourhood=BOX, coeff=VARIABLE, range=2, xc=1, yc=1, zc=3, tsteps=1, max_threads=64,
                     **base), (70, 9, 12)),
        (StencilDesc(neighbourhood=STAR, orientation=COLUMN, coeff=VARIABLE, range=2, xc=1, yc=2, zc=4, tsteps=2,
                     max_threads=128, remainder=True, **base), (145, 15, 17)),
        (StencilDesc(neighbourhood=STAR, coeff=VARIABLE, range=1, xc=1, yc=2, zc=4, tsteps=2, max_threads=128,
                     dtype=F32C, comp_group=3, remainder=True, **base), (130, 12, 14)),
        (StencilDesc(neighbourhood=BOX, orientation=COLUMN, coeff=VARIABLE, range=1, xc=1, yc=1, zc=3, tsteps=2,
                     max_threads=128, dtype=F32C, comp_group=9, **base), (100, 8, 10)),
    ]


def cfg1_desc(shape="plain"):
    """2D Box Jacobi r=1, coarsening 1x1, Row, alignment None."""
    return StencilDesc(dimension=2, neighbourhood=BOX, orientation=ROW, coeff=JACOBI, range=1, xc=1, yc=1,
                       alignment=ALIGN_NONE, cl_size=0, **SHAPES[shape])


def cfg2_desc(xc=1, yc=1, zc=1, shape="plain", dtype=F64, max_threads=1024, orientation=ROW):
    """3D Star Variable r=4, alignment Write at 16 points."""
    return StencilDesc(dimension=3, neighbourhood=STAR, orientation=orientation, coeff=VARIABLE, range=4,
                       xc=xc, yc=yc, zc=zc, alignment=ALIGN_WRITE, cl_size=16, dtype=dtype,
                       max_threads=max_threads, **SHAPES[shape])


def cfg2_sweep():
    """The coarsening sweep x{1,2} y{1,2,4} z{1,2,4,8} of the 512^3 workload (reference shape)."""
    return [cfg2_desc(x, y, z) for x in (1, 2) for y in (1, 2, 4) for z in (1, 2, 4, 8)]


def cfg2_tuned():
    """MI355X code shapes for the 512^3 workload."""
    out = []
    for (x, y, z) in ((1, 4, 16), (1, 2, 16), (2, 2, 16), (1, 4, 32), (2, 4, 16), (1, 8, 16), (4, 2, 16),
                      (2, 2, 32), (1, 2, 32)):
        for mt in (256, 512):
            out.append(cfg2_desc(x, y, z, shape="dpp", max_threads=mt))
    for (x, y, z) in ((1, 4, 32), (1, 4, 16), (1, 2, 32), (2, 2, 32), (1, 3, 32), (1, 4, 64), (2, 4, 32),
                      (1, 6, 32), (1, 8, 32)):
        for pf, wpe in ((0, 0), (1, 0), (2, 0), (1, 3), (1, 4)):
            out.append(cfg2_desc(x, y, z, shape="dppw", max_threads=256).replace(prefetch=pf, waves_per_eu=wpe))
    for (x, y, z) in ((2, 4, 32), (2, 4, 64), (2, 4, 128), (2, 2, 64), (2, 2, 128), (1, 4, 64), (1, 4, 128),
                      (1, 8, 64), (2, 3, 64), (2, 3, 128), (4, 2, 64), (1, 6, 64)):
        for zf in (False, True):
            out.append(cfg2_desc(x, y, z, shape="dppw", max_threads=256).replace(prefetch=0, zfirst=zf))
    return out


def cfg2_f32c_candidates():
    """FP32-storage compensated kernels for the 512^3 workload."""
    out = []
    for (x, y, z) in ((2, 4, 128), (2, 8, 64), (2, 6, 64)):
        for g in (1, 2, 5, 8, 13, 25):
            out.append(cfg2_desc(x, y, z, shape="dppw", max_threads=256, dtype=F32C).replace(prefetch=0,
                                                                                          comp_group=g))
    return out


def cfg3_desc(xc=1, yc=1, orientation=ROW, dtype=F64, shape="plain", max_threads=1024):
    """2D Box Variable r=8, alignment Read at 16 points."""
    return StencilDesc(dimension=2, neighbourhood=BOX, orientation=orientation, coeff=VARIABLE, range=8, xc=xc,
                       yc=yc, alignment=ALIGN_READ, cl_size=16, dtype=dtype, max_threads=max_threads,
                       **SHAPES[shape])


def cfg4_desc(dtype=F64, shape="dpp", batched=False, max_threads=256):
    """3D Box Jacobi r=2, coarsening 2x2x4 (the batched uneven-grids workload)."""
    d = StencilDesc(dimension=3, neighbourhood=BOX, orientation=ROW, coeff=JACOBI, range=2, xc=2, yc=2,
                    zc=4, alignment=ALIGN_NONE, cl_size=0, dtype=dtype, max_threads=max_threads,
                    **SHAPES[shape])
    return d.replace(batched=batched) if batched else d


def stream_offset_cases():
    """(desc, grid size incl. halo, blocks): 2D STREAM kernels with offset-major statements, every
    x-neighbour exchange mode, both neighbourhoods and orientations, Variable and Jacobi, FP64 and
    FP32C, remainder handling, and the 2D Box r=8 coarsening with paired edge loads."""
    base = dict(dimension=2, variant=STREAM, order="offset", prefetch=0, xcd_remap=True, nt_store=True)
    out = []
    for xext, blocks in (("dppw", [(64, 2, 1), (128, 1, 1)]), ("dpp", [(16, 4, 1), (96, 2, 1)]),
                         ("load", [(16, 4, 1)])):
        out += [
            (StencilDesc(neighbourhood=BOX, orientation=ROW, coeff=VARIABLE, range=3, xc=2, yc=3, alignment=ALIGN_READ,
                         cl_size=8, xext=xext, **base), (45, 23), blocks),
            (StencilDesc(neighbourhood=BOX, orientation=COLUMN, coeff=JACOBI, range=2, xc=4, yc=2, alignment=ALIGN_NONE,
                         cl_size=0, xext=xext, remainder=True, **base), (43, 21), blocks),
            (StencilDesc(neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=4, xc=2, yc=2,
                         alignment=ALIGN_WRITE, cl_size=16, dtype=F32C, comp_group=5, xext=xext, **base), (52, 20),
             blocks),
        ]
    out.append((StencilDesc(neighbourhood=BOX, orientation=ROW, coeff=VARIABLE, range=8, xc=4, yc=4, alignment=ALIGN_READ,
                            cl_size=16, xext="dppw", edge_pairs=True, **base), (96, 36), [(64, 1, 1)]))
    return out


def packed_star_desc(dtype=F64):
    """A batched 3D Star Variable r=3 kernel over packed tiles (x-neighbours by loads, remainder
    handling, Write alignment): the weights and a wider footprint through the packed tile map."""
    return StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=3, xc=2, yc=3, zc=5,
                       alignment=ALIGN_WRITE, cl_size=16, dtype=dtype, comp_group=7, max_threads=256,
                       remainder=True, batched=True, pack=True, **SHAPES["stream"])


def store_hazard_cases():
    """(name, desc, grid size incl. halo, block, steps): the STREAM kernels whose
    16-byte stores were once followed at once by a write of their data registers,
    at the workload sizes where they returned wrong points."""
    box = cfg1_desc("dppw").replace(xc=2, yc=4, max_threads=256)
    f32 = dict(shape="dppw", max_threads=256, dtype=F32C)
    f32o = dict(prefetch=0, comp_group=21, zfirst=False, cl_size=32)
    return [
        ("box2d_r1_2x4_dppw_256x1", box, (16386, 16386), (256, 1, 1), 8),
        ("box2d_r1_2x4_dppw_128x2", box, (16386, 16386), (128, 2, 1), 8),
        ("box2d_r1_2x4_load_256x1", cfg1_desc("stream").replace(xc=2, yc=4, max_threads=256), (16386, 16386),
         (256, 1, 1), 8),
        ("star3d_r4_f32c_4x4x64", cfg2_desc(4, 4, 64, **f32).replace(**f32o), (520, 520, 520), (128, 2, 1), 3),
        ("star3d_r4_f32c_4x2x128", cfg2_desc(4, 2, 128, **f32).replace(**f32o), (520, 520, 520), (128, 2, 1), 3),
    ]


def staged_cases():
    """(desc, grid size incl. halo, block) of STREAM kernels with LDS halo staging on small
    grids: several blocks per axis with partial last blocks, block rows of one and two waves,
    one and two z slices per block, thread rows that do not divide the tile's row count (the
    last staging load of some thread rows is not stored), FP32C, Box and Column cases, edge
    columns one or two rows per load."""
    dppw = dict(variant=STREAM, xext="dppw", xcd_remap=True, nt_store=True, prefetch=0)
    return [
        (StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=4, xc=2, yc=4, zc=4,
                     alignment=ALIGN_WRITE, cl_size=16, edge_pairs=True, stage=(64, 2, 2), max_threads=256, **dppw),
         (222, 28, 21), (64, 2, 2)),
        (StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=4, xc=2, yc=4, zc=3,
                     alignment=ALIGN_WRITE, cl_size=16, stage=(128, 4, 1), max_threads=512, **dppw),
         (300, 45, 18), (128, 4, 1)),
        (StencilDesc(dimension=3, neighbourhood=BOX, orientation=ROW, coeff=JACOBI, range=2, xc=2, yc=2, zc=3,
                     alignment=ALIGN_NONE, cl_size=0, stage=(64, 3, 1), max_threads=256, **dppw),
         (150, 19, 14), (64, 3, 1)),
        (StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=2, xc=4, yc=2, zc=2,
                     alignment=ALIGN_WRITE, cl_size=32, dtype=F32C, comp_group=3, edge_pairs=True, stage=(64, 2, 1),
                     max_threads=256, **dppw),
         (300, 15, 12), (64, 2, 1)),
        (StencilDesc(dimension=3, neighbourhood=STAR, orientation=COLUMN, coeff=JACOBI, range=1, xc=1, yc=1, zc=4,
                     alignment=ALIGN_READ, cl_size=8, stage=(128, 1, 1), max_threads=256, **dppw),
         (200, 9, 11), (128, 1, 1)),
        (StencilDesc(dimension=3, neighbourhood=BOX, orientation=ROW, coeff=VARIABLE, range=1, xc=2, yc=2, zc=2,
                     alignment=ALIGN_NONE, cl_size=0, stage=(64, 4, 2), max_threads=512, **dppw),
         (140, 21, 13), (64, 4, 2)),
    ]


def staged_remainder_cases():
    """The staged kernels generated with remainder handling, on grids whose interiors are not
    multiples of the coarsening nor of the block's tile in any axis (the last block of rows and
    of planes starts early, the last lanes in x drop their columns past the interior), each
    interior at least one tile in y and z."""
    sizes = [(223, 29, 21), (301, 45, 18), (151, 19, 14), (303, 15, 13), (201, 9, 11), (141, 21, 13)]
    out = []
    for (d, _, block), size in zip(staged_cases(), sizes):
        out.append((d.replace(remainder=True), size, block))
    return out


def smoke_desc():
    """3D Star Variable r=4, coarsening 1x4x16, DPP x exchange (the smoke run's kernel)."""
    return cfg2_desc(1, 4, 16, shape="dpp", max_threads=256)


def default_build_set():
    """Kernels the tests, smoke() and bench.py load.  Tuning candidate sets
    (cfg2_sweep, cfg2_tuned, cfg2_f32c_candidates, ...) are compiled on
    demand by the sweeps and cached in the same directory."""
    descs = [d for d, _ in parity_cases()] + [d for d, _ in f32c_cases()] + [d for d, _ in remainder_cases()]
    descs += [d for d, _ in temporal_cases()] + [d for d, _ in temporal3d_cases()]
    descs += [cfg1_desc("plain"), cfg1_desc("stream"), cfg1_desc("dppw")]
    descs += [cfg2_desc(1, 1, 1), cfg2_desc(2, 4, 8)]            # reference-shaped kernels
    descs += [smoke_desc()]
    descs += [cfg4_desc(F64, "plain"), cfg4_desc(F64, "dppw"), cfg4_desc(F64, "plain", max_threads=1024)]
    descs += [cfg4_desc(F64, "plain", batched=True, max_threads=1024),
              cfg4_desc(F32C, "plain", batched=True, max_threads=1024)]
    descs += [cfg4_desc(F64, "dppw", batched=True), cfg4_desc(F32C, "dppw", batched=True),
              cfg4_desc(F32C, "plain")]
    # batched kernels with remainder handling (grids not rounded to the coarsening)
    descs += [cfg4_desc(dt, "plain", batched=True, max_threads=1024).replace(remainder=True) for dt in (F64, F32C)]
    descs += [cfg4_desc(dt, "dppw", batched=True).replace(remainder=True) for dt in (F64, F32C)]
    # ... and over packed tiles
    descs += [cfg4_desc(dt, shape, batched=True, max_threads=1024 if shape == "plain" else 256).replace(
        remainder=True, pack=True) for dt in (F64, F32C) for shape in ("plain", "stream")]
    descs += [packed_star_desc(dt) for dt in (F64, F32C)]
    descs += [d for d, _, _ in stream_offset_cases()]
    descs += [c[1] for c in store_hazard_cases()]
    descs += [d for d, _, _ in staged_cases()] + [d for d, _, _ in staged_remainder_cases()]
    from . import tuned, tuner
    descs += tuned.build_set()
    descs += [d for d, _ in tuner.live_configs()]   # the tuner's live check
    return list(dict.fromkeys(descs))
