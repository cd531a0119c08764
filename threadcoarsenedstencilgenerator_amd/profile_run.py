"""Workload for counter profiling (rocprofv3 --pmc).

Runs, in one process, a plain device copy of 2^30 bytes (the calibration
kernel: FETCH_SIZE / WRITE_SIZE read for it are compared with the bytes it
moves) and then a few steps of the tuned 512^3 3D Star r=4 kernel, so one
counter pass measures both.

    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc --output-format csv -- \
        python -m threadcoarsenedstencilgenerator_amd.profile_run [--dtype f64|f32c] [--steps 3]
        (--workload box2d_r1_temporal: the temporal 2D Box r=1 kernel on 16384^2 instead;
         --workload all --meta FILE: every measured workload in one pass, see all_workloads)

With --batched-sweep it instead times the batched workload's kernel candidates
(code shape, coarsening, tile packing and block shape, each checked bit for bit
against the plain 2x2x4 kernel).

With --ceiling it instead times the HBM stream probes (copy, read-only,
write-only; plain and nontemporal; 1-16 loads of 16 bytes in flight per lane,
256-1024 threads per workgroup; one pass and grid-stride launches) over
2 GiB buffers (past every cache level, about twice the 512^3 workload's grid)
and prints one JSON line per probe: the practical bandwidth ceiling the
stencil kernels are compared with.
"""
from __future__ import annotations

import argparse


def ceiling(reps: int = 20):
    import json

    import torch

    from . import runtime as rt
    torch.cuda.set_device(0)
    nbytes = 2 * 2 ** 30 - (2 * 2 ** 30) % (1 << 20)   # 2 GiB per buffer: past every cache level
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    out = []
    for mode in ("copy", "read", "write"):
        moved = 2 * nbytes if mode == "copy" else nbytes
        for nt in (False, True):
            for unroll, block in ((1, 256), (2, 256), (4, 256), (8, 256), (16, 256), (4, 512), (4, 1024),
                                  (8, 512), (16, 1024)):
                for wg in (0, 256 * 4, 256 * 8, 256 * 16, 256 * 32):
                    ms = rt.hbm_probe(mode, a.data_ptr(), b.data_ptr(), nbytes, nt, wg, reps, unroll=unroll,
                                      block=block)
                    line = dict(probe=mode, nontemporal=nt, unroll=unroll, block=block,
                                workgroups=wg or "one pass", bytes_per_launch=moved, ms=round(ms, 4),
                                tbs=round(moved / ms / 1e9, 3))
                    out.append(line)
                    print(json.dumps(line), flush=True)
    best = {m: max(x["tbs"] for x in out if x["probe"] == m) for m in ("copy", "read", "write")}
    print(json.dumps({"hbm_ceiling_tbs": best, "peak_tbs": 8.0}), flush=True)


def batched_candidates(dtype):
    """(name, desc, block) of the batched persistent kernels the code-shape, coarsening and block
    sweep times, all with remainder handling (any coarsening runs the workload's 2x2x4-rounded
    grids untruncated): the reference's code shape (PLAIN), the z-march shape (STREAM) with
    x-neighbours by loads ("load"), by lane shifts within block rows ("dpp") or within whole waves
    ("dppw"), over rectangular tiles or packed ones ("pack": runs of block x * block y consecutive
    threads of a grid's xy thread space).  PLAIN kernels have a launch bound of 1024 threads,
    STREAM ones 256 unless named ("_mt")."""
    from . import manifest
    from .spec import F64
    plain = manifest.cfg4_desc(shape="plain", batched=True, dtype=dtype, max_threads=1024)
    out = [("plain 2x2x4", plain, (32, 8, 2))]
    for co, blocks, pack in (((2, 2, 8), ((32, 8, 2), (16, 16, 2), (32, 4, 4)), False),
                             ((1, 2, 8), ((32, 8, 2),), False), ((2, 2, 12), ((32, 8, 2),), False),
                             ((2, 2, 16), ((32, 8, 2),), False), ((2, 4, 4), ((32, 8, 2),), False),
                             ((4, 2, 4), ((16, 8, 2),), False),
                             ((2, 2, 8), ((32, 8, 2),), True), ((1, 2, 8), ((32, 8, 2),), True)):
        d = plain.replace(xc=co[0], yc=co[1], zc=co[2], remainder=True, pack=pack)
        for blk in blocks:
            out.append((f"plain {co[0]}x{co[1]}x{co[2]}" + (" pack" if pack else ""), d, blk))
    shape = {"load": "stream", "dpp": "dpp", "dppw": "dppw"}
    nt = dtype == F64
    table = (("dppw", (2, 2, 16), 256, ((64, 4, 1),), False, 1, nt),
             ("dpp", (2, 2, 16), 256, ((32, 8, 1), (16, 16, 1)), False, 1, nt),
             ("dpp", (2, 4, 16), 256, ((32, 8, 1),), False, 1, nt),
             ("load", (2, 2, 16), 256, ((32, 8, 1),), False, 1, nt),
             ("load", (2, 4, 16), 256, ((32, 8, 1), (16, 16, 1)), False, 1, nt),
             ("load", (2, 4, 16), 256, ((32, 8, 1), (32, 4, 2)), True, 1, nt),
             ("load", (2, 4, 16), 256, ((32, 8, 1),), True, 1, not nt),
             ("load", (2, 4, 16), 256, ((32, 8, 1),), True, 0, nt),
             ("load", (2, 4, 16), 256, ((32, 8, 1),), True, 2, nt),
             ("load", (2, 4, 16), 128, ((32, 4, 1),), True, 1, nt),
             ("load", (2, 4, 16), 512, ((32, 16, 1),), True, 1, nt),
             ("load", (2, 2, 16), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (2, 2, 8), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (2, 4, 8), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (2, 4, 32), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (2, 3, 16), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (2, 6, 16), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (1, 4, 16), 256, ((32, 8, 1),), True, 1, nt),
             ("load", (4, 2, 16), 256, ((16, 8, 1),), True, 1, nt))
    for xext, co, mt, blocks, pack, pf, ntst in table:
        d = manifest.cfg4_desc(shape=shape[xext], batched=True, dtype=dtype, max_threads=mt).replace(
            xc=co[0], yc=co[1], zc=co[2], remainder=True, zfirst=False, pack=pack, prefetch=pf, nt_store=ntst)
        for blk in blocks:
            out.append((f"{xext} {co[0]}x{co[1]}x{co[2]}" + (" pack" if pack else "") + (f" mt{mt}" if mt != 256 else "")
                        + (f" prefetch{pf}" if pf != 1 else "") + (" nt" if ntst else ""), d, blk))
    if dtype != F64:
        # TwoSum groups (FP32C): 5 terms by default; 13, and 25 (one z plane of the box each)
        d = manifest.cfg4_desc(shape="stream", batched=True, dtype=dtype, max_threads=256).replace(
            xc=2, yc=4, zc=16, remainder=True, zfirst=False, pack=True)
        for cg in (13, 25):
            out.append((f"plain 2x2x4 group{cg}", plain.replace(comp_group=cg), (32, 8, 2)))
            out.append((f"load 2x4x16 pack group{cg}", d.replace(comp_group=cg), (32, 8, 1)))
        # with groups of 25 (fewer live registers): four columns per lane
        for co in ((4, 4, 8), (4, 4, 12), (4, 4, 16), (2, 8, 16), (4, 2, 16)):
            out.append((f"load {co[0]}x{co[1]}x{co[2]} pack group25",
                        d.replace(xc=co[0], yc=co[1], zc=co[2], comp_group=25), (16, 16, 1)))
    return out


def batched_sweep(dtype_name: str, steps: int = 50):
    """The batched workload (64 grids, interiors log-uniform in [32, 256] rounded to 2x2x4)
    through each candidate of batched_candidates: 2 steps from the seeded state compared bit
    for bit with the plain 2x2x4 kernel's of the same TwoSum group size (the first such
    candidate), then GLup/s over `steps` graph-replayed steps.  One JSON line per candidate.
    Kernels not in the tree are compiled first."""
    import json
    import time

    import numpy as np
    import torch

    from . import build as B
    from . import runtime as rt
    from . import spec, tuned
    torch.cuda.set_device(0)
    dt = spec.F32C if dtype_name == "f32c" else spec.F64
    ui = np.uint32 if dt == spec.F32C else np.uint64
    stream = torch.cuda.current_stream()
    refs = {}   # per TwoSum group size: the plain 2x2x4 kernel's result
    warm = True
    for name, d, block in batched_candidates(dt):
        B.build_kernel(d, force=False)
        k = rt.Kernel(d)
        sizes = rt.batch_sizes(64, 1, 32, 256, tuned.BATCH_ROUNDING, d.range)
        geoms = [k.geometry(sz) for sz in sizes]
        grids = [_setup(k, g, rt.INIT_U01, 100 + i) for i, g in enumerate(geoms)]
        bt = rt.Batch(k, geoms, [g.ptr(0) for g in grids], [g.ptr(1) for g in grids], block)
        pts, tiles, nwg = bt.info()
        fin = bt.run(2, graph=False, stream=stream)
        torch.cuda.synchronize()
        res = [g.host(fin) for g in grids]
        ref = refs.setdefault(d.comp_group, res)
        same = all(np.array_equal(a.view(ui), c.view(ui)) for a, c in zip(ref, res))
        # the first candidate also brings the GPU to its working clocks
        bt.run(10 * steps if warm else steps, graph=True, stream=stream)
        warm = False
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bt.run(steps, graph=True, stream=stream)
        torch.cuda.synchronize()
        glups = pts * steps / (time.perf_counter() - t0) / 1e9
        print(json.dumps(dict(dtype=dtype_name, candidate=name, block=list(block), kernel=d.key(),
                              glups=round(glups, 2), tiles=tiles, workgroups=nwg, same_as_plain_2x2x4=bool(same))),
              flush=True)
        bt.close()
        del grids
        k.close()
        torch.cuda.empty_cache()


def _setup(k, g, mode, seed):
    from . import runtime as rt
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), mode, seed)
    return grids


def all_workloads(steps: int, meta: str):
    """One pass over every measured workload (the counter report's dispatches):
    the 2^30-byte calibration copy, then `steps` launches each of the 512^3 FP64
    and FP32C production kernels, the 2D Box r=8 16384x8192 FP64 kernel, the
    2D temporal kernel on 16384^2, the 3D temporal and one-step 3D Star r=1
    kernels on 512^3 and the batched persistent kernel over the 64 uneven grids
    (no graph).  Writes `meta`: per workload the kernel symbol,
    launch grid and workgroup (threads), interior points, time steps per launch,
    flops per update (reference count) and algorithmic bytes per update."""
    import json

    import torch

    from . import manifest, spec, tuned
    from . import runtime as rt
    torch.cuda.set_device(0)
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    out = [dict(name="copy_2^30B", symbol="__amd_rocclr_copyBuffer", bytes_per_launch=1 << 30)]

    def record(name, k, d, g, block, kind="single"):
        blocks = rt.grid_blocks(g, block)
        es = 4 if d.dtype == spec.F32C else 8
        out.append(dict(name=name, kernel=d.key(), symbol="tcsg_update", kind=kind,
                        grid_threads=[blocks[i] * block[i] for i in range(3)], workgroup=list(block),
                        interior_points=g.interior_points, steps_per_launch=k.steps_per_launch,
                        flops_per_update=d.flops_per_update, bytes_per_update=2 * es))
        return blocks

    for name, cfg, grid, mode in (("star3d_r4_512_f64", tuned.CFG2_F64, (520, 520, 520), rt.INIT_NORMAL),
                                  ("star3d_r4_512_f32c", tuned.CFG2_F32C, (520, 520, 520), rt.INIT_NORMAL),
                                  ("box2d_r8_16384x8192_f64", tuned.BOX2D_R8_F64, tuned.BOX2D_R8_GRID, rt.INIT_UPM1),
                                  ("box2d_r1_16384_temporal_f64", tuned.BOX2D_R1_TEMPORAL, tuned.BOX2D_R1_GRID,
                                   rt.INIT_U01),
                                  ("star3d_r1_512_temporal_f64", tuned.STAR3D_R1_TEMPORAL, tuned.STAR3D_R1_GRID,
                                   rt.INIT_NORMAL),
                                  ("star3d_r1_512_one_step_f64", tuned.STAR3D_R1_ONE_STEP, tuned.STAR3D_R1_GRID,
                                   rt.INIT_NORMAL)):
        d, block = cfg["desc"], tuple(cfg["block"])
        k = rt.Kernel(d)
        if d.coeff == spec.VARIABLE:
            k.set_mask(rt.mask_seeded(2, d.n))
        g = k.geometry(grid)
        grids = _setup(k, g, mode, 1)
        blocks = record(name, k, d, g, block)
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), steps)
        torch.cuda.synchronize()
        print(f"{name}: {d.key()} block {block} grid {blocks}", flush=True)
        del grids
        k.close()
        torch.cuda.empty_cache()
    # batched: 64 uneven 3D Box Jacobi r=2 grids, one persistent launch per step
    d, block = tuned.BATCHED_F64["desc"], tuple(tuned.BATCHED_F64["block"])
    k = rt.Kernel(d)
    sizes = rt.batch_sizes(64, 1, 32, 256, tuned.BATCH_ROUNDING, d.range)
    geoms = [k.geometry(sz) for sz in sizes]
    gr = [_setup(k, g, rt.INIT_U01, 100 + i) for i, g in enumerate(geoms)]
    bt = rt.Batch(k, geoms, [x.ptr(0) for x in gr], [x.ptr(1) for x in gr], block)
    pts, tiles, nwg = bt.info()
    out.append(dict(name="batched_64_box3d_r2_f64", kernel=d.key(), symbol="tcsg_update_batched", kind="batched",
                    grid_threads=[nwg * block[0] * block[1] * block[2], 1, 1], workgroup=list(block),
                    interior_points=pts, steps_per_launch=1, flops_per_update=d.flops_per_update,
                    bytes_per_update=16, tiles=tiles))
    bt.run(steps, graph=False)
    torch.cuda.synchronize()
    print(f"batched: {d.key()} {tiles} tiles, {nwg} workgroups", flush=True)
    bt.close()
    with open(meta, "w") as f:
        json.dump(out, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--dtype", default="f64", choices=("f64", "f32c"))
    ap.add_argument("--ceiling", action="store_true")
    ap.add_argument("--workload", default="star3d_r4_512", choices=("star3d_r4_512", "box2d_r1_temporal", "all"))
    ap.add_argument("--meta", default="profile_workloads.json", help="--workload all: where to write the "
                    "dispatch table")
    ap.add_argument("--batched-sweep", action="store_true", help="time the batched kernel's code shape, "
                    "coarsening and block candidates (--dtype f64|f32c) instead")
    args = ap.parse_args()
    if args.ceiling:
        return ceiling()
    if args.batched_sweep:
        return batched_sweep(args.dtype)
    if args.workload == "all":
        return all_workloads(args.steps, args.meta)
    import torch

    from . import runtime as rt
    from . import spec, tuned
    torch.cuda.set_device(0)
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("calibration copy: bytes read per launch", n * 8, flush=True)
    del a, b
    if args.workload == "box2d_r1_temporal":
        best = tuned.BOX2D_R1_TEMPORAL
        d, block = best["desc"], best["block"]
        k = rt.Kernel(d)
        g = k.geometry(tuned.BOX2D_R1_GRID)
    else:
        best = tuned.best("star3d_r4_512", spec.F32C if args.dtype == "f32c" else spec.F64)
        d, block = best["desc"], best["block"]
        k = rt.Kernel(d)
        g = k.geometry((520, 520, 520))
        k.set_mask(rt.mask_seeded(2, d.n))
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), rt.INIT_NORMAL, 1)
    k.run(g, block, rt.grid_blocks(g, block), grids.ptr(0), grids.ptr(1), args.steps)
    torch.cuda.synchronize()
    es = 4 if d.dtype == spec.F32C else 8
    print(f"{d.key()} block {tuple(block)}: interior points {g.interior_points}, steps per launch "
          f"{k.steps_per_launch}, algorithmic bytes per launch {2 * es * g.interior_points}", flush=True)


if __name__ == "__main__":
    main()
