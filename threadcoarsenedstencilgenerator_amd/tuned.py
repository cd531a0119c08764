"""Best measured kernel configuration per benchmark workload on MI355X.

Each entry is the code shape, coarsening and block shape that the block-shape
and coarsening sweeps found fastest (the measurements are summarised under
profiles/).  bench.py runs these.  Tiles are ordered x-fastest within each
XCD's contiguous range (zfirst=False): an XCD then owns whole rows of a
z-chunk, so no 128-byte line at a tile's x edge is fetched by two XCDs
(HBM reads 1.20x the interior's instead of 1.34x with z-fastest order).
"""
from __future__ import annotations

from . import manifest
from .spec import ALIGN_NONE, F32C, F64, ROW, STAR, STREAM, TEMPORAL, VARIABLE, StencilDesc

# 3D Star Variable r=4, 512^3 interior, FP64, Write alignment at 16 points.  LDS halo staging
# for blocks of 128x4 threads (a 256x16-point tile per 128-plane z chunk, 98 KiB of LDS): each
# plane's 24 rows are loaded once per block instead of 12 rows per thread for its 4 output rows,
# 15.8% faster than the same coarsening loading its rows itself (346.9 against 299.5 GLup/s in an
# interleaved timing; blocks of 128x2 threads: 322, 256x2: 325, 64x8: 310; 64-plane chunks 318;
# the wave-edge columns one row per load 340; without nontemporal stores 329; without the XCD
# remap 324).  The wave-edge columns of two rows per load, x-fastest tile order.
CFG2_F64 = dict(desc=manifest.cfg2_desc(2, 4, 128, shape="dppw", max_threads=512).replace(
    prefetch=0, zfirst=False, edge_pairs=True, stage=(128, 4, 1)), block=(128, 4, 1))
# the same workload in FP32 storage with compensated accumulation: coarsening 4x4x64 with LDS
# staging for blocks of 64x4 threads (a 256x16-point tile, 48 KiB of LDS; 615.6 GLup/s against
# 563.2 for the unstaged 2x8x64 kernel; 4x4x64 staged for 64x8 threads: 592.5, 64x2: 590.8;
# 2x4x64 staged for 128x4: 591.0; 2x4x128 kernels: 482-532).  FP32 rows are aligned to 32 points =
# 128 bytes, the line size the FP64 grid's 16 points make; the first group sums the 21 terms up to
# the output's own plane, the second the 4 terms of the planes after it, and TwoSum joins them.
CFG2_F32C = dict(desc=manifest.cfg2_desc(4, 4, 64, shape="dppw", max_threads=256, dtype=F32C).replace(
    prefetch=0, comp_group=21, zfirst=False, cl_size=32, edge_pairs=True, stage=(64, 4, 1)), block=(64, 4, 1))
# the kernels that load their own rows (no LDS staging), fastest before staging (the staged
# kernels' baselines in the measurements)
CFG2_F64_UNSTAGED = dict(desc=CFG2_F64["desc"].replace(stage=(0, 0, 0)), block=(128, 2, 2))
CFG2_F32C_UNSTAGED = dict(desc=manifest.cfg2_desc(2, 8, 64, shape="dppw", max_threads=256, dtype=F32C).replace(
    prefetch=0, comp_group=21, zfirst=False, cl_size=32, edge_pairs=True), block=(128, 2, 1))


def remainder_cfg(dtype: int = F64, staged: bool = True):
    """The 512^3 workload's kernel generated with remainder handling: the staged production
    kernel (its last block of rows and of planes starts early enough to end at the interior's
    end), or with staged=False the kernel that loads its own rows (its last thread of each
    axis starts early instead)."""
    if staged:
        c = CFG2_F32C if dtype == F32C else CFG2_F64
    else:
        c = CFG2_F32C_UNSTAGED if dtype == F32C else CFG2_F64_UNSTAGED
    return dict(desc=c["desc"].replace(remainder=True), block=c["block"])


# 2D Box Variable r=8 (289 weights), 16384x8192 interior, Read alignment at 16 points
BOX2D_R8_GRID = (16384 + 16, 8192 + 16)
# (FP64-bound: 577 flops per update.  The reference's code shape with a 4x4 register tile,
# its statements in offset order: the 289 weights stream through the scalar registers once
# per thread, each used for all 16 outputs, and each footprint point is loaded once)
BOX2D_R8_F64 = dict(desc=manifest.cfg3_desc(4, 4, shape="plain", max_threads=256).replace(order="offset"),
                    block=(64, 2, 1))
BOX2D_R8_F32C = dict(desc=manifest.cfg3_desc(4, 4, shape="plain", max_threads=256, dtype=F32C).replace(
    order="offset", comp_group=17), block=(128, 2, 1))


# 2D Box Jacobi r=1 (the 1024^2 workload's stencil) on a 16384x16384 interior, where it is
# HBM-bound: the temporal kernel advances several steps per launch through LDS; the one-step
# kernel is the fastest single-step shape measured on the same grid
BOX2D_R1_GRID = (16384 + 2, 16384 + 2)
# (sweep over 3-8 steps per launch and tiles of 64x16, 64x32, 128x8 and 128x16 points: 4 steps on
# 128x16-point tiles with 512 threads was fastest, 670 GLup/s against 303 for the best one-step shape)
BOX2D_R1_TEMPORAL = dict(desc=manifest.cfg1_desc("plain").replace(
    variant=TEMPORAL, tsteps=4, xc=2, yc=2, max_threads=512, remainder=True), block=(64, 8, 1))
# (one-step: STREAM 2x4 with x-neighbours by loads, 324 GLup/s; the segment-DPP form ran at 317,
# the wave-aligned DPP form at 301)
# FP32 storage (the F32C dtype; comp_group 9 puts all 9 terms of the Box r=1 stencil in one group,
# a plain FP32 chain with no TwoSum): half the LDS per point, so 64x64-point tiles fit; 4 steps
# per launch, 1169 GLup/s in the sweep
BOX2D_R1_TEMPORAL_F32C = dict(desc=manifest.cfg1_desc("plain").replace(
    variant=TEMPORAL, tsteps=4, xc=1, yc=8, max_threads=512, remainder=True, dtype=F32C, comp_group=9),
    block=(64, 8, 1))
BOX2D_R1_ONE_STEP = dict(desc=manifest.cfg1_desc("stream").replace(xc=2, yc=4, max_threads=256),
                         block=(256, 1, 1))


# 3D Star Variable r=1 (7 weights) on a 512^3 interior (halo 1), where one step is HBM-bound: the
# 3D temporal kernel marches each 64x16-point tile through 128 planes, 2 steps per launch, with a ring
# of 3 planes per step in LDS (61 KB, 512 threads).  Sweep over 1-4 steps per launch, tiles 64x8 to
# 128x16 and 64x32, 32-128 planes, 256-1024 threads: 2 steps on 64x16x128 was fastest.  The one-step
# baseline is the fastest one-step shape measured on the same grid: STREAM 2x4x128 with LDS halo
# staging for blocks of 256x2 threads (323.1 GLup/s; staged for 128x4: 318.8, 128x2: 307.2; the
# same kernel loading its own rows: 301.9; x-neighbours by loads: 298.1).
STAR3D_R1_GRID = (514, 514, 514)
STAR3D_R1_TEMPORAL = dict(desc=StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=1,
                                           xc=1, yc=2, zc=128, alignment=ALIGN_NONE, cl_size=0, variant=TEMPORAL,
                                           tsteps=2, max_threads=512, remainder=True), block=(64, 8, 1))
STAR3D_R1_ONE_STEP = dict(desc=StencilDesc(dimension=3, neighbourhood=STAR, orientation=ROW, coeff=VARIABLE, range=1,
                                           xc=2, yc=4, zc=128, alignment=ALIGN_NONE, cl_size=0, variant=STREAM,
                                           xext="dppw", xcd_remap=True, nt_store=True, prefetch=0, edge_pairs=True,
                                           max_threads=512, stage=(256, 2, 1)),
                          block=(256, 2, 1))


# 64 uneven 3D Box Jacobi r=2 grids (interior axes log-uniform in [32, 256], rounded to the
# workload's 2x2x4 coarsening), one persistent launch per step.  The z-march code shape (STREAM)
# with x-neighbours by loads, coarsening 2x4x16 and remainder handling, over packed tiles: a tile is
# a run of 256 consecutive threads of a grid's xy thread space, so the 32-256-point grids leave no
# lanes idle at their x and y ends (rectangular 32x8 tiles cover 0.65 of their lanes with interior
# points).  FP64 152.8 GLup/s against 128.0 for the same kernel over rectangular tiles and 106.9 for
# the earlier tuned kernel (the reference's code shape, 2x2x8 at 32x8x2); FP32C (groups of 5) 202.8
# against 170.2 and 154.2.  Sweep over coarsenings 2x2x8-2x6x16, 4x2x16, 1x4x16, 2x4x32, x-neighbours by loads or
# by DPP, 128-512 threads, packed and rectangular tiles (profiles: batched_bench.md).
BATCH_ROUNDING = (2, 2, 4)
BATCHED_F64 = dict(desc=manifest.cfg4_desc(shape="stream", batched=True, max_threads=256).replace(
    xc=2, yc=4, zc=16, remainder=True, zfirst=False, pack=True), block=(32, 8, 1))
# FP32C: TwoSum groups of 25 terms, one z plane of the 5x5x5 box each (fewer live error terms: 108
# VGPRs at 2x4x16 against 208 for groups of 5, 218.2 GLup/s in the sweep against 202.8), which leaves
# room for 4 columns per lane: 4x4x8 over packed runs of 256 threads, 305.4 GLup/s in the sweep
# (4x4x12 304.6, 4x4x16 287.3, 2x8x16 271.2; 315.1 in the bench line).  Normwise error against the FP64 kernel after 20 steps 1.2e-6 (groups of 5:
# 1.07e-6), set by the FP32 rounding of the stored U[0,1) state at every step.
BATCHED_F32C = dict(desc=manifest.cfg4_desc(shape="stream", batched=True, dtype=F32C, max_threads=256).replace(
    xc=4, yc=4, zc=8, remainder=True, zfirst=False, nt_store=False, pack=True, comp_group=25), block=(16, 16, 1))
# the tuned kernels before packing: the reference's code shape (PLAIN) with a z chunk of 8 planes
# per thread (12 planes loaded for 8 outputs instead of 8 for 4) over rectangular tiles; FP64 2x2x8
# at 32x8x2, FP32C 1x2x8 at 32x8x2
BATCHED_PLAIN_F64 = dict(desc=manifest.cfg4_desc(shape="plain", batched=True, max_threads=1024).replace(
    zc=8, remainder=True), block=(32, 8, 2))
BATCHED_PLAIN_F32C = dict(desc=manifest.cfg4_desc(shape="plain", batched=True, dtype=F32C, max_threads=1024).replace(
    xc=1, zc=8, remainder=True), block=(32, 8, 2))


def best(workload: str, dtype: int = F64):
    if workload == "star3d_r4_512":
        return CFG2_F32C if dtype == F32C else CFG2_F64
    if workload == "box2d_r8_16384x8192":
        return BOX2D_R8_F32C if dtype == F32C else BOX2D_R8_F64
    if workload == "batched":
        return BATCHED_F32C if dtype == F32C else BATCHED_F64
    raise KeyError(workload)


def build_set():
    return [CFG2_F64["desc"], CFG2_F32C["desc"], BOX2D_R8_F64["desc"], BOX2D_R8_F32C["desc"],
            CFG2_F64_UNSTAGED["desc"], CFG2_F32C_UNSTAGED["desc"],
            remainder_cfg(F64)["desc"], remainder_cfg(F32C)["desc"],
            remainder_cfg(F64, staged=False)["desc"], remainder_cfg(F32C, staged=False)["desc"],
            BOX2D_R1_TEMPORAL["desc"], BOX2D_R1_ONE_STEP["desc"], BOX2D_R1_TEMPORAL_F32C["desc"],
            STAR3D_R1_TEMPORAL["desc"], STAR3D_R1_ONE_STEP["desc"], BATCHED_F64["desc"], BATCHED_F32C["desc"],
            BATCHED_PLAIN_F64["desc"], BATCHED_PLAIN_F32C["desc"]]
