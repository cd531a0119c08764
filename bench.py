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


def bench_batched(args):
    """64 independent 3D Box Jacobi r=2 grids, each interior axis log-uniform
    in [32, 256] (rounded to the 2x2x4 coarsening), 100 time steps: one
    persistent kernel per step over a captured hipGraph, against the same
    steps as 64 separate launches per step."""
    import numpy as np
    import torch

    from threadcoarsenedstencilgenerator_amd import manifest
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd.spec import F32C, F64
    out = {}
    steps = 100
    for name, dt in (("f64", F64), ("f32c", F32C)):
        d = manifest.cfg4_desc(shape="dppw", batched=True, dtype=dt)
        k = rt.Kernel(d)
        sizes = rt.batch_sizes(64, 1, 32, 256, d.coarsening, d.range)
        geoms, grids = [], []
        for i, size in enumerate(sizes):
            g = k.geometry(size)
            gr = rt.Grids(g, k.dtype)
            for j in range(2):
                k.init(g, gr.ptr(j), rt.INIT_U01, 100 + i)
            geoms.append(g)
            grids.append(gr)
        b = rt.Batch(k, geoms, [g.ptr(0) for g in grids], [g.ptr(1) for g in grids], (64, 2, 1))
        pts, tiles, wg = b.info()
        stream = torch.cuda.current_stream()
        res = {}
        for mode in ("graph", "persistent", "separate"):
            run = (lambda: b.run(steps, graph=True, stream=stream)) if mode == "graph" else \
                  (lambda: b.run(steps, graph=False, stream=stream)) if mode == "persistent" else \
                  (lambda: b.run_separate(steps, stream=stream))
            run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            dt_s = time.perf_counter() - t0
            res[mode] = round(pts * steps / dt_s / 1e9, 2)
        out[name] = {"glups": res, "points": pts, "tiles": tiles, "workgroups": wg,
                     "speedup_graph_vs_separate": round(res["graph"] / res["separate"], 3)}
        b.close()
        del grids
        torch.cuda.empty_cache()
    line = {"metric": "GLup/s, 64 uneven 3D Box Jacobi r=2 grids, 100 steps", "value": out["f64"]["glups"]["graph"],
            "unit": "GLup/s", "n_gpus": 1, "steps": steps, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "64 grids, interior axes log-uniform in [32,256] rounded "
                                            "to coarsening 2x2x4, U[0,1) values", "block": [64, 2, 1]},
            "detail": out}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", default="star3d_r4_512", choices=("star3d_r4_512", "batched"))
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
    if args.workload == "batched":
        return bench_batched(args)
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

    fin = k.run(g, block, blocks, grids.ptr(0), grids.ptr(1), args.warmup, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k.run(g, block, blocks, grids.ptr(fin), grids.ptr(1 - fin), args.steps, stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    glups = pts * args.steps / dt / 1e9

    # kernel duration from HIP events recorded on the launch stream
    ms = k.time_steps(g, block, blocks, grids.ptr(0), grids.ptr(1), 0, args.steps, stream)
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
                                  "max_threads": desc.max_threads},
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
