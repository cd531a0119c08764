#!/usr/bin/env python3
"""Headline benchmark: FP64 3D Star Variable stencil, range 4, 512^3.

One "step" is one time step of the stencil over the whole 512^3 interior
(halo 4, rows padded for Write alignment at 16 points = 128 B), i.e. one
launch of the generated kernel reading one buffer and writing the other.
Inputs are synthetic (seeded N(0,1) grid, seeded mask U[-1,1] normalised to
sum |w| = 1) and resident in HBM before timing starts.

Prints one JSON line: GLup/s (whole job), the roofline of the stencil
kernel (algorithmic 16 B per lattice update = one FP64 read + one FP64
write, against 8 TB/s), and the OpenMP CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]
    python bench.py --workload batched     # 64 uneven 3D Box grids, persistent kernel + hipGraph
    python bench.py --workload box2d_r8    # 2D Box r=8 16384x8192, FP64 and FP32-compensated
    python bench.py --workload box2d_r1_temporal   # 2D Box r=1 16384^2, several steps per launch
    python bench.py --workload star3d_r4_large     # the 512^3 kernels on 1024x1024x512 (4.5 GB buffers)
    python bench.py --workload star3d_r1_temporal  # 3D Star r=1 512^3, 2 steps per launch (z-march, LDS rings)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GRID = (520, 520, 520)     # 512^3 interior + 4-point halo per side
INPUT_SEED, MASK_SEED = 1, 2


def _traffic(desc_key: str, block) -> dict | None:
    """HBM bytes per launch measured with rocprofv3 counters (profiles/)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    e = t.get(desc_key)
    if not e or tuple(e.get("block", ())) != tuple(block):
        return None
    return e


def cpu_baseline(desc, budget_s: float):
    """The OpenMP C oracle (oracle/stencil_oracle.c) on this host: same grid,
    same seeded inputs, as many whole steps as fit in about budget_s."""
    import numpy as np

    import oracle
    g = oracle.Geometry(3, desc.range, desc.alignment, desc.cl_size, GRID, desc.coarsening)
    a = np.empty(g.n_points)
    b = np.empty(g.n_points)
    oracle.init_seeded(g, 2, INPUT_SEED, a)
    b[:] = a
    mask = oracle.mask_seeded(MASK_SEED, desc.n)
    t0 = time.perf_counter()
    oracle.step_f64(desc, g, mask, a, b)
    one = time.perf_counter() - t0
    steps = max(1, min(400, int(budget_s / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    oracle.run_f64(desc, g, mask, b, a, steps)
    dt = time.perf_counter() - t0
    lups = g.interior_points * steps / dt / 1e9
    return {"value": round(lups, 4), "unit": "GLup/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{steps} FP64 time steps of the same 512^3 3D Star r=4 Variable grid "
                      f"({dt:.1f} s), OpenMP C oracle"}


def f32c_companion(k64, g, g64, steps, stream):
    """The same workload in FP32 storage with compensated accumulation: the error of
    its state after 20 steps against the FP64 state (both started from the same FP64
    initial grid), reduced in FP64 on the GPU: normwise max |u32 - u64| / max |u64|
    (the headline), normwise RMS, and pointwise max |u32 - u64| / max(|u64|, 1e-3 max
    |u64|).  Returns (that record, timed(warmup) -> (GLup/s, kernel ms per step) over the
    same number of steps as the FP64 headline, close()); the caller times it after the
    headline, so both kernels are timed on a GPU at its working clocks, with the same
    warm-up and on states of the same age (steps warmup+1 .. warmup+steps of the seeded
    grid)."""
    import numpy as np
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import spec, tuned
    c32 = tuned.best("star3d_r4_512", spec.F32C)
    d32, b32 = c32["desc"], tuple(c32["block"])
    k32 = rt.Kernel(d32)
    k32.set_mask(rt.mask_seeded(MASK_SEED, d32.n))
    gf = k32.geometry(GRID)   # FP32 rows padded to their own 128-byte lines
    g32 = rt.Grids(gf, np.float32)
    for i in range(2):
        # the seeded fill computes each value in FP64, then stores it in the
        # grid's type: the FP32 grid holds the FP32 rounding of the FP64 grid
        k64.init(g, g64.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
        k32.init(gf, g32.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
    blocks = rt.grid_blocks(gf, b32)
    f32 = k32.run(gf, b32, blocks, g32.ptr(0), g32.ptr(1), 20, stream)
    desc64 = k64.desc
    b64 = tuple(tuned.best("star3d_r4_512")["block"])
    f64 = k64.run(g, b64, rt.grid_blocks(g, b64), g64.ptr(0), g64.ptr(1), 20, stream)
    torch.cuda.synchronize()
    mx, mb, s2, b2, _ = rt.error_f32_vs_f64(gf, g32.ptr(f32), g, g64.ptr(f64), 1.0)
    floor = 1e-3 * mb   # pointwise errors relative to |u64|, floored at 1e-3 of the field's magnitude
    prel = rt.error_f32_vs_f64(gf, g32.ptr(f32), g, g64.ptr(f64), floor)[4]
    info = {"glups": None,
            "coarsening": list(d32.coarsening), "block": list(b32), "comp_group": d32.comp_group,
            "bytes_per_update": 8, "error_after_steps": 20, "normwise_max_rel": float(mx / mb),
            "normwise_rms_rel": float(np.sqrt(s2 / b2)),
            "pointwise_max_rel": float(prel), "pointwise_floor": float(floor),
            "pointwise_floor_rule": "1e-3 * max|u64|", "f64_kernel": desc64.key()}

    def timed(warmup):
        """(GLup/s of `steps` FP32C steps after `warmup` untimed ones, wall clock, synchronised;
        HIP-event kernel ms per step of the same steps), each from the seeded grid, as the FP64
        headline is timed: the FP32C kernel's time depends on the age of the state it steps."""
        def fresh():
            for i in range(2):
                k32.init(gf, g32.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
            return k32.run(gf, b32, blocks, g32.ptr(0), g32.ptr(1), warmup, stream)
        fin = fresh()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k32.run(gf, b32, blocks, g32.ptr(fin), g32.ptr(1 - fin), steps, stream)
        torch.cuda.synchronize()
        glups = gf.interior_points * steps / (time.perf_counter() - t0) / 1e9
        fin = fresh()
        ms = k32.time_steps(gf, b32, blocks, g32.ptr(fin), g32.ptr(1 - fin), 0, steps, stream) / steps
        return glups, ms

    def close():
        k32.close()
    return info, timed, close


def bench_batched(args):
    """64 independent 3D Box Jacobi r=2 grids, each interior axis log-uniform
    in [32, 256] (rounded to the workload's 2x2x4 coarsening), 100 time steps,
    by the tuned kernel (z march with x-neighbours by loads, coarsening 2x4x16,
    remainder handling, packed tiles: runs of 256 consecutive threads of a
    grid's xy thread space).  Modes: one persistent kernel per step (as many
    workgroups per CU as fit, taking tiles from a work queue with one range of
    the tile table per XCD) over a captured hipGraph of the step loop (the
    headline of this workload), the same without the graph, and 64 separate
    launches per step, of the same kernel and of the plain 2x2x4 kernel (the
    workload's named baseline); also the previous tuned kernel (the
    reference's code shape, z chunks of 8, rectangular tiles) persistent +
    graph.  The *_unrounded entries run the same draws without the rounding."""
    import torch

    from threadcoarsenedstencilgenerator_amd import manifest, tuned
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd.spec import F32C, F64
    out = {}
    steps = 100
    err_steps = 20
    ref64 = None

    def timed(run):
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    for name, dt, rounded in (("f64", F64, True), ("f32c", F32C, True), ("f64_unrounded", F64, False),
                              ("f32c_unrounded", F32C, False)):
        cfg = tuned.best("batched", dt)
        d, block = cfg["desc"], tuple(cfg["block"])
        k = rt.Kernel(d)
        sizes = rt.batch_sizes(64, 1, 32, 256, tuned.BATCH_ROUNDING if rounded else (1, 1, 1), d.range)
        geoms, grids = [], []
        for i, size in enumerate(sizes):
            g = k.geometry(size)
            gr = rt.Grids(g, k.dtype)
            for j in range(2):
                k.init(g, gr.ptr(j), rt.INIT_U01, 100 + i)
            geoms.append(g)
            grids.append(gr)
        stream = torch.cuda.current_stream()
        res, info = {}, {}
        b = rt.Batch(k, geoms, [g.ptr(0) for g in grids], [g.ptr(1) for g in grids], block)
        pts, tiles, nwg = b.info()
        info["persistent"] = nwg
        err = None
        if rounded:
            # FP32C against FP64: both from the same seeded values, after err_steps steps
            fin = b.run(err_steps, graph=False, stream=stream)
            torch.cuda.synchronize()
            if dt == F64:
                ref64 = [gr.bufs[fin].clone() for gr in grids]
            else:
                err = max(float((gr.bufs[fin].double() - r).abs().max() / r.abs().max()) for gr, r in zip(grids, ref64))
                del ref64
        res["persistent_graph"] = round(pts * steps / timed(lambda: b.run(steps, graph=True, stream=stream)) / 1e9, 2)
        res["persistent"] = round(pts * steps / timed(lambda: b.run(steps, graph=False, stream=stream)) / 1e9, 2)
        res["separate_launches"] = round(pts * steps / timed(lambda: b.run_separate(steps, stream=stream)) / 1e9, 2)
        b.close()
        k.close()
        if rounded:
            # the named baseline: 64 launches per step of the plain kernel at the workload's coarsening
            d0 = manifest.cfg4_desc(shape="plain", batched=True, dtype=dt, max_threads=1024)
            k0 = rt.Kernel(d0)
            b0 = rt.Batch(k0, [k0.geometry(sz) for sz in sizes], [g.ptr(0) for g in grids],
                          [g.ptr(1) for g in grids], (32, 8, 2))
            res["separate_launches_2x2x4"] = round(
                pts * steps / timed(lambda: b0.run_separate(steps, stream=stream)) / 1e9, 2)
            b0.close()
            k0.close()
            prev = tuned.BATCHED_PLAIN_F32C if dt == F32C else tuned.BATCHED_PLAIN_F64
            k1 = rt.Kernel(prev["desc"])
            b1 = rt.Batch(k1, [k1.geometry(sz) for sz in sizes], [g.ptr(0) for g in grids],
                          [g.ptr(1) for g in grids], tuple(prev["block"]))
            res["persistent_graph_previous_plain"] = round(
                pts * steps / timed(lambda: b1.run(steps, graph=True, stream=stream)) / 1e9, 2)
            b1.close()
            k1.close()
        out[name] = {"glups": res, "points": pts, "tiles": tiles, "workgroups": info, "kernel": d.key(),
                     "block": list(block),
                     "speedup_persistent_graph_vs_separate": round(res["persistent_graph"] / res["separate_launches"], 3)}
        if rounded:
            out[name]["speedup_vs_separate_2x2x4"] = round(res["persistent_graph"] / res["separate_launches_2x2x4"], 3)
        if err is not None:
            out[name]["normwise_max_rel_error_vs_f64"] = {"steps": err_steps, "value": err,
                                                          "comp_group": d.comp_group}
        del grids
        torch.cuda.empty_cache()
    # HBM roofline of the FP64 persistent + graph mode: one read and one write of every interior
    # point per step, against the step rate; traffic from the counters where profiles/ has them
    glups = out["f64"]["glups"]["persistent_graph"]
    d64, block = tuned.BATCHED_F64["desc"], tuple(tuned.BATCHED_F64["block"])
    tr = _traffic(d64.key(), block)
    alg = 16.0 * out["f64"]["points"]
    roof = {"bound": "hbm", "achieved": round(16.0 * glups, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(16.0 * glups / HBM_PEAK_GBS, 4),
            "traffic": (round(tr["hbm_bytes_per_launch"] / tr["algorithmic_bytes_per_launch"] * alg) if tr else None),
            "algorithmic_bytes_per_launch": int(alg)}
    line = {"metric": "GLup/s, 64 uneven 3D Box Jacobi r=2 grids, 100 steps", "value": glups, "roofline": roof,
            "unit": "GLup/s", "n_gpus": 1, "steps": steps, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "64 grids, interior axes log-uniform in [32,256] rounded "
                                            "to coarsening 2x2x4, U[0,1) values", "block": list(block),
                                            "code_shape": "z march (x-neighbours by loads), coarsening 2x4x16, "
                                                          "remainder handling, packed tiles (persistent "
                                                          "batched kernel)"},
            "detail": out}
    return line


def bench_box2d_r8(args):
    """2D Box Variable r=8 (289 weights, normalised to sum |w| = 1), 16384x8192
    interior, Read alignment at 16 points, values U[-1,1): FP64 (the drop-in)
    and FP32 storage with compensated accumulation, 50 steps each (after 10
    untimed warm-up steps of each kernel and a fresh seeded fill), and the
    FP32 state's error against the FP64 state after the 50 steps.  FP64-bound:
    577 flops per update against the 78.6 TFLOP/s FP64 vector peak."""
    import numpy as np
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import spec, tuned
    steps = 50
    c64, c32 = tuned.best("box2d_r8_16384x8192"), tuned.best("box2d_r8_16384x8192", spec.F32C)
    k64, k32 = rt.Kernel(c64["desc"]), rt.Kernel(c32["desc"])
    b64, b32 = tuple(c64["block"]), tuple(c32["block"])
    g = k64.geometry(tuned.BOX2D_R8_GRID)
    gf = k32.geometry(tuned.BOX2D_R8_GRID)
    mask = rt.mask_seeded(MASK_SEED, c64["desc"].n)
    k64.set_mask(mask)
    k32.set_mask(mask)
    g64 = rt.Grids(g, np.float64)
    for i in range(2):
        k64.init(g, g64.ptr(i), rt.INIT_UPM1, INPUT_SEED)
    g32 = rt.Grids(gf, np.float32)
    for i in range(2):
        k32.init(gf, g32.ptr(i), rt.INIT_UPM1, INPUT_SEED)   # the FP32 rounding of the FP64 grid
    stream = torch.cuda.current_stream()
    # untimed warm-up steps of both kernels (the first steps on an idle GPU run up to 15% slower),
    # then both grids back to the seeded state
    k64.run(g, b64, rt.grid_blocks(g, b64), g64.ptr(0), g64.ptr(1), 10, stream)
    k32.run(gf, b32, rt.grid_blocks(gf, b32), g32.ptr(0), g32.ptr(1), 10, stream)
    for i in range(2):
        k64.init(g, g64.ptr(i), rt.INIT_UPM1, INPUT_SEED)
        k32.init(gf, g32.ptr(i), rt.INIT_UPM1, INPUT_SEED)
    torch.cuda.synchronize()
    t64 = t32 = 0.0
    errs = {}
    done = 0
    for part in (40, steps - 40):   # error after 40 steps and after all 50
        t0 = time.perf_counter()
        k64.run(g, b64, rt.grid_blocks(g, b64), g64.ptr(0), g64.ptr(1), part, stream)
        torch.cuda.synchronize()
        t64 += time.perf_counter() - t0
        t0 = time.perf_counter()
        k32.run(gf, b32, rt.grid_blocks(gf, b32), g32.ptr(0), g32.ptr(1), part, stream)
        torch.cuda.synchronize()
        t32 += time.perf_counter() - t0
        mx, mb, s2, b2, _ = rt.error_f32_vs_f64(gf, g32.ptr(0), g, g64.ptr(0), 1e-3)   # even counts end in buffer 0
        done += part
        errs[done] = {"normwise_max_rel": float(mx / mb), "normwise_rms_rel": float(np.sqrt(s2 / b2)),
                      "max_abs_f64": float(mb)}
    ms64 = k64.time_steps(g, b64, rt.grid_blocks(g, b64), g64.ptr(0), g64.ptr(1), 0, 10, stream) / 10
    pts = g.interior_points
    flops = c64["desc"].flops_per_update * pts
    achieved = flops / (ms64 * 1e-3) / 1e12
    line = {"metric": "GLup/s, FP64 2D Box r=8 16384x8192 (% of FP64 vector peak in roofline.frac)",
            "value": round(pts * steps / t64 / 1e9, 3), "unit": "GLup/s", "n_gpus": 1, "steps": steps,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "2D Box Variable r=8 (289 weights, sum|w|=1), 16384x8192 interior + 8-point "
                                   "halo, Read alignment 16 points, U[-1,1) values",
                       "coarsening": list(c64["desc"].coarsening[:2]), "block": list(b64)},
            "roofline": {"bound": "fp64 vector", "achieved": round(achieved, 2), "peak": 78.6, "unit": "TFLOP/s",
                         "frac": round(achieved / 78.6, 4), "flops_per_update": c64["desc"].flops_per_update,
                         "kernel_ms": round(ms64, 4)},
            "f32c": {"glups": round(pts * steps / t32 / 1e9, 3), "speedup_vs_f64": round(t64 / t32, 3),
                     "coarsening": list(c32["desc"].coarsening[:2]), "comp_group": c32["desc"].comp_group,
                     "error_vs_f64_after_steps": errs,
                     "note": "the field shrinks ~6x per step (sum w = -0.004): after 50 steps max|u| < 1.2e-38, "
                             "so the FP32 state is subnormal"}}
    return line


def bench_remainder(args):
    """The 512^3 workload's production kernel (LDS staging) generated with
    remainder handling, on a 511^3 interior (no axis a multiple of its
    coarsening or of its block's tile): every point updated, against the same
    kernel without the option, which truncates that grid as the reference does
    (z to a multiple of its 128-plane chunk), and against itself on the 512^3
    grid (the cost of the row-group stores); the remainder-handling kernel that
    loads its own rows is timed on the 511^3 grid too."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import tuned
    best = tuned.CFG2_F64
    block = tuple(best["block"])
    unstaged = tuned.remainder_cfg(staged=False)
    kr, kt = rt.Kernel(tuned.remainder_cfg()["desc"]), rt.Kernel(best["desc"])
    ku = rt.Kernel(unstaged["desc"])
    mask = rt.mask_seeded(MASK_SEED, best["desc"].n)
    stream = torch.cuda.current_stream()
    out = {}
    for name, k, grid, blk in (("remainder_511", kr, (519, 519, 519), block),
                               ("truncating_511", kt, (519, 519, 519), block),
                               ("remainder_512", kr, GRID, block), ("truncating_512", kt, GRID, block),
                               ("remainder_unstaged_511", ku, (519, 519, 519), tuple(unstaged["block"]))):
        k.set_mask(mask)
        g = k.geometry(grid)
        grids = rt.Grids(g, k.dtype)
        for i in range(2):
            k.init(g, grids.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
        blocks = rt.grid_blocks(g, blk)
        k.run(g, blk, blocks, grids.ptr(0), grids.ptr(1), args.warmup, stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k.run(g, blk, blocks, grids.ptr(0), grids.ptr(1), args.steps, stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms = k.time_steps(g, blk, blocks, grids.ptr(0), grids.ptr(1), 0, args.steps, stream) / args.steps
        user = (grid[0] - 2 * g.range) * (grid[1] - 2 * g.range) * (grid[2] - 2 * g.range)
        out[name] = {"glups": round(g.interior_points * args.steps / dt / 1e9, 2), "kernel_ms": round(ms, 4),
                     "interior": list(g.elements.tup()), "coverage": round(g.interior_points / user, 4),
                     "hbm_gbs": round(16.0 * g.interior_points / (ms * 1e-3) / 1e9, 1)}
        del grids
        torch.cuda.empty_cache()
    r = out["remainder_511"]
    line = {"metric": "GLup/s, FP64 3D Star r=4 511^3 with remainder handling", "value": r["glups"],
            "unit": "GLup/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "3D Star Variable r=4, 511^3 interior + 4-point halo, FP64, Write alignment "
                                   "16 points, N(0,1) values",
                       "coarsening": list(best["desc"].coarsening), "block": list(block)},
            "roofline": {"bound": "hbm", "achieved": r["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(r["hbm_gbs"] / HBM_PEAK_GBS, 4), "traffic": None},
            "runs": out}
    return line


def bench_large(args):
    """The 512^3 production kernel on a 1024x1024x512 interior (halo 4, Write
    alignment: 4.53 GB per FP64 buffer, past 4 GiB), where each block addresses
    its own window of planes with 32-bit offsets; FP64 and the FP32-storage
    kernel (2.3 GB buffers)."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import spec, tuned
    grid = (1032, 1032, 520)
    stream = torch.cuda.current_stream()
    out = {}
    for name, cfg in (("f64", tuned.best("star3d_r4_512")), ("f32c", tuned.best("star3d_r4_512", spec.F32C))):
        k = rt.Kernel(cfg["desc"])
        block = tuple(cfg["block"])
        k.set_mask(rt.mask_seeded(MASK_SEED, cfg["desc"].n))
        g = k.geometry(grid)
        grids = rt.Grids(g, k.dtype)
        for i in range(2):
            k.init(g, grids.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
        blocks = rt.grid_blocks(g, block)
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), args.warmup, stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), args.steps, stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms = k.time_steps(g, block, blocks, grids.ptr(0), grids.ptr(1), 0, args.steps, stream) / args.steps
        es = 4 if cfg["desc"].dtype == spec.F32C else 8
        pts = g.interior_points
        out[name] = {"glups": round(pts * args.steps / dt / 1e9, 2), "kernel_ms": round(ms, 4),
                     "buffer_bytes": int(g.n_points * es),
                     "hbm_gbs": round(2 * es * pts / (ms * 1e-3) / 1e9, 1), "kernel": cfg["desc"].key()}
        del grids
        k.close()
        torch.cuda.empty_cache()
    r = out["f64"]
    line = {"metric": "GLup/s, FP64 3D Star r=4 1024x1024x512 (4.5 GB buffers)", "value": r["glups"],
            "unit": "GLup/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "3D Star Variable r=4, 1024x1024x512 interior + 4-point halo, Write alignment "
                                   "16 points, N(0,1) values", "coarsening": list(tuned.CFG2_F64["desc"].coarsening),
                       "block": list(tuned.CFG2_F64["block"])},
            "roofline": {"bound": "hbm", "achieved": r["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(r["hbm_gbs"] / HBM_PEAK_GBS, 4), "traffic": None},
            "runs": out}
    return line


def bench_temporal(args):
    """2D Box Jacobi r=1 (the 1024^2 workload's stencil) on a 16384x16384
    interior, FP64, U[0,1) values: the temporal kernel (several steps per
    launch through LDS) against the fastest one-step kernel on the same grid.
    Both run the same number of time steps; the one-step kernel's roofline is
    its 16 B per update of HBM traffic, which the temporal kernel spends once
    per launch instead of once per step."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import tuned
    tcfg, ocfg = tuned.BOX2D_R1_TEMPORAL, tuned.BOX2D_R1_ONE_STEP
    stream = torch.cuda.current_stream()
    out = {}
    for name, cfg in (("temporal", tcfg), ("one_step", ocfg)):
        k = rt.Kernel(cfg["desc"])
        block = tuple(cfg["block"])
        K = k.steps_per_launch
        launches = max(1, args.steps // K)
        g = k.geometry(tuned.BOX2D_R1_GRID)
        grids = rt.Grids(g, k.dtype)
        for i in range(2):
            k.init(g, grids.ptr(i), rt.INIT_U01, INPUT_SEED)
        blocks = rt.grid_blocks(g, block)
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), max(1, args.warmup // K), stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), launches, stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms = k.time_steps(g, block, blocks, grids.ptr(0), grids.ptr(1), 0, launches, stream) / launches
        pts = g.interior_points
        out[name] = {"glups": round(pts * K * launches / dt / 1e9, 2), "steps": K * launches,
                     "steps_per_launch": K, "kernel_ms_per_launch": round(ms, 4),
                     "kernel_glups": round(pts * K / (ms * 1e-3) / 1e9, 2), "block": list(block),
                     "kernel": cfg["desc"].key()}
        del grids
        k.close()
        torch.cuda.empty_cache()
    # FP32 storage (F32C): the same number of steps from the FP32 rounding
    # of the same grid, and its error against the FP64 temporal kernel's state
    import numpy as np
    c32 = tuned.BOX2D_R1_TEMPORAL_F32C
    k64, k32 = rt.Kernel(tcfg["desc"]), rt.Kernel(c32["desc"])
    g64, g32 = k64.geometry(tuned.BOX2D_R1_GRID), k32.geometry(tuned.BOX2D_R1_GRID)
    G64, G32 = rt.Grids(g64, np.float64), rt.Grids(g32, np.float32)
    for i in range(2):
        k64.init(g64, G64.ptr(i), rt.INIT_U01, INPUT_SEED)
        k32.init(g32, G32.ptr(i), rt.INIT_U01, INPUT_SEED)
    steps = out["temporal"]["steps"]
    b64, b32 = tuple(tcfg["block"]), tuple(c32["block"])
    f64i = k64.run(g64, b64, rt.grid_blocks(g64, b64), G64.ptr(0), G64.ptr(1), steps // k64.steps_per_launch, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f32i = k32.run(g32, b32, rt.grid_blocks(g32, b32), G32.ptr(0), G32.ptr(1), steps // k32.steps_per_launch, stream)
    torch.cuda.synchronize()
    dt32 = time.perf_counter() - t0
    mx, mb, s2, b2, _ = rt.error_f32_vs_f64(g32, G32.ptr(f32i), g64, G64.ptr(f64i), 1e-3)
    f32c = {"glups": round(g32.interior_points * steps / dt32 / 1e9, 2), "steps": steps,
            "steps_per_launch": k32.steps_per_launch, "tile": list(c32["desc"].tile[:2]), "block": list(b32),
            "comp_group": c32["desc"].comp_group, "normwise_max_rel_vs_f64": float(mx / mb),
            "normwise_rms_rel_vs_f64": float(np.sqrt(s2 / b2)), "kernel": c32["desc"].key()}
    del G64, G32
    k64.close()
    k32.close()
    t, o = out["temporal"], out["one_step"]
    pts = (tuned.BOX2D_R1_GRID[0] - 2) * (tuned.BOX2D_R1_GRID[1] - 2)
    alg = 16.0 * pts / (t["kernel_ms_per_launch"] * 1e-3) / 1e9   # one read + one write per point per launch
    tr = _traffic(tcfg["desc"].key(), tcfg["block"])
    line = {"metric": "GLup/s, FP64 2D Box Jacobi r=1 16384^2, temporal blocking", "value": t["glups"],
            "unit": "GLup/s", "n_gpus": 1, "steps": t["steps"], "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "2D Box Jacobi r=1, 16384x16384 interior + 1-point halo, FP64, alignment None, "
                                   "U[0,1) values",
                       "steps_per_launch": t["steps_per_launch"], "tile": list(tcfg["desc"].tile[:2]),
                       "block": t["block"]},
            "roofline": {"bound": "hbm", "achieved": round(alg, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / HBM_PEAK_GBS, 4),
                         "traffic": (tr["hbm_bytes_per_launch"] if tr else None),
                         "one_step_equivalent_gbs": round(16.0 * t["kernel_glups"], 1),
                         "note": "achieved: 16 B per point per launch (each launch advances steps_per_launch "
                                 "steps); one_step_equivalent_gbs: the HBM rate a one-step kernel would need "
                                 "for the same updates per second"},
            "vs_one_step": round(t["glups"] / o["glups"], 3), "runs": out,
            "f32c": dict(f32c, speedup_vs_f64=round(f32c["glups"] / t["glups"], 3))}
    return line


def bench_temporal3d(args):
    """3D Star Variable r=1 (7 weights) on a 512^3 interior, FP64, N(0,1) values:
    the 3D temporal kernel (2 steps per launch, z-march with plane rings in LDS)
    against the fastest one-step kernel on the same grid, the same number of
    steps each, and the OpenMP oracle on this host."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import tuned
    tcfg, ocfg = tuned.STAR3D_R1_TEMPORAL, tuned.STAR3D_R1_ONE_STEP
    stream = torch.cuda.current_stream()
    mask = rt.mask_seeded(MASK_SEED, tcfg["desc"].n)
    out = {}
    for name, cfg in (("temporal", tcfg), ("one_step", ocfg)):
        k = rt.Kernel(cfg["desc"])
        k.set_mask(mask)
        block = tuple(cfg["block"])
        K = k.steps_per_launch
        launches = max(1, args.steps // K)
        g = k.geometry(tuned.STAR3D_R1_GRID)
        grids = rt.Grids(g, k.dtype)
        for i in range(2):
            k.init(g, grids.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
        blocks = rt.grid_blocks(g, block)
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), max(1, args.warmup // K), stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), launches, stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms = k.time_steps(g, block, blocks, grids.ptr(0), grids.ptr(1), 0, launches, stream) / launches
        pts = g.interior_points
        out[name] = {"glups": round(pts * K * launches / dt / 1e9, 2), "steps": K * launches,
                     "steps_per_launch": K, "kernel_ms_per_launch": round(ms, 4),
                     "kernel_glups": round(pts * K / (ms * 1e-3) / 1e9, 2), "block": list(block),
                     "kernel": cfg["desc"].key()}
        del grids
        k.close()
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_grid(tcfg["desc"], tuned.STAR3D_R1_GRID, min(args.cpu_seconds, 10.0),
                                "3D Star r=1 Variable 512^3")
    t, o = out["temporal"], out["one_step"]
    pts = 512 ** 3
    alg = 16.0 * pts / (t["kernel_ms_per_launch"] * 1e-3) / 1e9   # one read + one write per point per launch
    line = {"metric": "GLup/s, FP64 3D Star r=1 512^3, temporal blocking", "value": t["glups"],
            "unit": "GLup/s", "n_gpus": 1, "steps": t["steps"], "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "3D Star Variable r=1 (7 weights, sum|w|=1), 512^3 interior + 1-point halo, "
                                   "FP64, alignment None, N(0,1) values",
                       "steps_per_launch": t["steps_per_launch"],
                       "tile": [64 * tcfg["desc"].xc, tcfg["desc"].max_threads // 64 * tcfg["desc"].yc,
                                tcfg["desc"].zc], "block": t["block"]},
            "roofline": {"bound": "hbm", "achieved": round(alg, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / HBM_PEAK_GBS, 4), "traffic": None,
                         "one_step_equivalent_gbs": round(16.0 * t["kernel_glups"], 1),
                         "note": "achieved: 16 B per point per launch (each launch advances steps_per_launch "
                                 "steps); one_step_equivalent_gbs: the HBM rate a one-step kernel would need "
                                 "for the same updates per second"},
            "vs_one_step": round(t["glups"] / o["glups"], 3), "runs": out, "cpu_baseline": cpu}
    return line


def cpu_baseline_grid(desc, grid, budget_s: float, what: str):
    """The OpenMP C oracle on this host for `desc` on `grid` (seeded N(0,1)
    interior): as many whole steps as fit in about budget_s."""
    import numpy as np

    import oracle
    g = oracle.Geometry(desc.dimension, desc.range, desc.alignment, desc.cl_size, grid, (1, 1, 1))
    a = np.empty(g.n_points)
    b = np.empty(g.n_points)
    oracle.init_seeded(g, 2, INPUT_SEED, a)
    b[:] = a
    mask = oracle.mask_seeded(MASK_SEED, desc.n)
    t0 = time.perf_counter()
    oracle.step_f64(desc, g, mask, a, b)
    one = time.perf_counter() - t0
    steps = max(1, min(400, int(budget_s / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    oracle.run_f64(desc, g, mask, b, a, steps)
    dt = time.perf_counter() - t0
    return {"value": round(g.interior_points * steps / dt / 1e9, 4), "unit": "GLup/s",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{steps} FP64 time steps of the same {what} grid ({dt:.1f} s), OpenMP C oracle"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the summaries of the other measured workloads in the headline line")
    ap.add_argument("--workload", default="star3d_r4_512", choices=("star3d_r4_512", "batched", "box2d_r8",
                                                                    "star3d_r4_511_remainder", "box2d_r1_temporal",
                                                                    "star3d_r4_large", "star3d_r1_temporal"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              f"the single grid would shard by z-slabs with a halo exchange per step, which is not built.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import spec, tuned

    torch.cuda.set_device(0)
    others = {"batched": bench_batched, "box2d_r8": bench_box2d_r8, "star3d_r4_511_remainder": bench_remainder,
              "box2d_r1_temporal": bench_temporal, "star3d_r4_large": bench_large,
              "star3d_r1_temporal": bench_temporal3d}
    if args.workload in others:
        print(json.dumps(others[args.workload](args)), flush=True)
        return
    best = tuned.best("star3d_r4_512")
    desc, block = best["desc"], tuple(best["block"])
    k = rt.Kernel(desc)
    g = k.geometry(GRID)
    k.set_mask(rt.mask_seeded(MASK_SEED, desc.n))
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), rt.INIT_NORMAL, INPUT_SEED)
    blocks = rt.grid_blocks(g, block)
    stream = torch.cuda.current_stream()
    pts = g.interior_points

    # the FP32C companion's error run first: its GPU work (initialisation, 40 steps, the error
    # reduction) also brings the GPU to its working clocks before the headline's warm-up and
    # timed steps; its own timing follows the headline's, with the same warm-up
    f32c, f32c_timed, f32c_close = f32c_companion(k, g, grids, args.steps, stream)
    for i in range(2):
        k.init(g, grids.ptr(i), rt.INIT_NORMAL, INPUT_SEED)

    fin = k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), args.warmup, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k.run(g, block, blocks, grids.ptr(fin), grids.ptr(1 - fin), args.steps, stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    glups = pts * args.steps / dt / 1e9
    g32, ms32 = f32c_timed(args.warmup)
    f32c_close()
    f32c["glups"] = round(g32, 3)
    f32c["ratio_vs_f64"] = round(f32c["glups"] / glups, 3)

    # kernel duration from HIP events recorded on the launch stream
    ms = k.time_steps(g, block, blocks, grids.ptr(0), grids.ptr(1), 0, args.steps, stream)
    f32c["kernel_ms"] = round(ms32, 5)
    f32c["ratio_vs_f64_kernel_time"] = round(ms / args.steps / ms32, 3)
    kern_s = ms * 1e-3 / args.steps
    bytes_alg = 16.0 * pts
    achieved = bytes_alg / kern_s / 1e9
    tr = _traffic(desc.key(), block)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (tr["hbm_bytes_per_launch"] if tr else None),
            "kernel_ms": round(ms / args.steps, 5), "algorithmic_bytes_per_launch": int(bytes_alg)}


    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, args.cpu_seconds)

    line = {
        "metric": "GLup/s, FP64 3D Star r=4 512^3 (% of HBM roofline in roofline.frac)",
        "value": round(glups, 3), "unit": "GLup/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D Star Variable stencil r=4, 512x512x512 interior + 4-point halo, FP64, "
                               "Write alignment 16 points, seeded N(0,1) grid and U[-1,1] mask (sum|w|=1)",
                   "coarsening": list(desc.coarsening), "block": list(block),
                   "code_shape": {"variant": "stream" if desc.variant == spec.STREAM else "plain",
                                  "xext": desc.xext, "xcd_remap": desc.xcd_remap, "nt_store": desc.nt_store,
                                  "max_threads": desc.max_threads, "edge_pairs": desc.edge_pairs,
                                  "lds_stage_block": list(desc.stage) if desc.staged else None},
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "f32c": f32c,
    }
    if not args.headline_only:
        line["other_workloads"] = other_workloads(args, others)
    print(json.dumps(line), flush=True)


def cpu_rate(desc, sizes, budget_s: float, max_steps: int = 50):
    """GLup/s of the OpenMP C oracle on this host over grids of `sizes` (incl. halo), same kind
    of seeded inputs: whole steps of every grid, as many as fit in about budget_s (at least
    one).  Returns (GLup/s, steps, seconds)."""
    import numpy as np

    import oracle
    grids = []
    for i, size in enumerate(sizes):
        g = oracle.Geometry(desc.dimension, desc.range, desc.alignment, desc.cl_size, size, desc.coarsening)
        a = np.empty(g.n_points)
        oracle.init_seeded(g, 1, INPUT_SEED + i, a)
        grids.append((g, a, a.copy()))
    mask = oracle.mask_seeded(MASK_SEED, desc.n)
    pts = sum(g.interior_points for g, _, _ in grids)
    t0 = time.perf_counter()
    for g, a, b in grids:
        oracle.step_f64(desc, g, mask, a, b)
    one = time.perf_counter() - t0
    steps = max(1, min(max_steps, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for g, a, b in grids:
        oracle.run_f64(desc, g, mask, a, b, steps)
    dt = time.perf_counter() - t0
    return pts * steps / dt / 1e9, steps, dt


def other_cpu_baseline(name, budget_s: float = 3.0):
    """The oracle's rate for another measured workload's stencil, on its grid or a bounded
    part of it (FP64; the same per-point work as the GPU run's)."""
    import oracle
    from threadcoarsenedstencilgenerator_amd import manifest, tuned
    from threadcoarsenedstencilgenerator_amd.runtime import batch_sizes
    if name in ("star3d_r4_large", "star3d_r4_511_remainder"):
        return {"same_as": "headline cpu_baseline (same stencil, per-point work independent of the grid)"}
    if name == "box2d_r1_temporal":
        d, sizes, what = manifest.cfg1_desc("plain"), [tuned.BOX2D_R1_GRID], "2D Box Jacobi r=1 16384^2"
    elif name == "star3d_r1_temporal":
        d, sizes, what = tuned.STAR3D_R1_ONE_STEP["desc"], [tuned.STAR3D_R1_GRID], "3D Star Variable r=1 512^3"
    elif name == "box2d_r8":
        # a quarter of the rows: one step of the whole grid is ~77 GFLOP
        d, sizes, what = manifest.cfg3_desc(1, 1), [(16384 + 16, 2048 + 16)], \
            "2D Box Variable r=8, 16384x2048 (a quarter of the 16384x8192 grid)"
    elif name == "batched":
        d = tuned.BATCHED_PLAIN_F64["desc"].replace(batched=False)
        sizes, what = batch_sizes(64, 1, 32, 256, tuned.BATCH_ROUNDING, d.range), "the 64 uneven 3D Box Jacobi r=2 grids"
    else:
        return None
    rate, steps, dt = cpu_rate(d, sizes, budget_s)
    return {"value": round(rate, 4), "unit": "GLup/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{steps} FP64 step(s) of {what} ({dt:.1f} s), OpenMP C oracle"}


def other_workloads(args, others):
    """The other measured workloads, each with its own fixed step count, summarised
    (GLup/s, roofline fraction and what each compares against); a workload that
    fails is reported with its error instead."""
    import torch
    out = {}
    for name, steps in (("box2d_r1_temporal", 100), ("star3d_r1_temporal", 100), ("box2d_r8", 50),
                        ("batched", 100), ("star3d_r4_large", 10), ("star3d_r4_511_remainder", 10)):
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.no_cpu_baseline = steps, 2, True
        t0 = time.perf_counter()
        try:
            ln = others[name](sub)
        except Exception as e:  # noqa: BLE001 -- reported, the headline line still prints
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            continue
        finally:
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
        r = ln.get("roofline") or {}
        summ = {"metric": ln["metric"], "value": ln["value"], "unit": ln["unit"], "steps": ln.get("steps"),
                "roofline_frac": r.get("frac"), "roofline_bound": r.get("bound"), "seconds": round(time.perf_counter() - t0, 1)}
        if "vs_one_step" in ln:
            summ["vs_one_step"] = ln["vs_one_step"]
            summ["one_step_glups"] = ln["runs"]["one_step"]["glups"]
            summ["kernel_glups"] = ln["runs"]["temporal"]["kernel_glups"]
        if not args.no_cpu_baseline:
            summ["cpu_baseline"] = other_cpu_baseline(name)
        if name == "batched":
            d = ln["detail"]
            summ["f32c"] = d["f32c"]["glups"]["persistent_graph"]
            summ["separate_launches"] = d["f64"]["glups"]["separate_launches"]
            summ["separate_launches_2x2x4"] = d["f64"]["glups"]["separate_launches_2x2x4"]
            summ["previous_plain_kernel"] = d["f64"]["glups"]["persistent_graph_previous_plain"]
            summ["f32c_normwise_max_rel_error"] = d["f32c"]["normwise_max_rel_error_vs_f64"]["value"]
        if name == "box2d_r8":
            summ["f32c"] = ln["f32c"]["glups"]
        if name == "star3d_r4_large":
            summ["f32c"] = ln["runs"]["f32c"]["glups"]
            summ["buffer_bytes"] = ln["runs"]["f64"]["buffer_bytes"]
        out[name] = summ
    return out


if __name__ == "__main__":
    main()
