"""Workload for counter profiling (rocprofv3 --pmc).

Runs, in one process, a plain device copy of 2^30 bytes (the calibration
kernel: FETCH_SIZE / WRITE_SIZE read for it are compared with the bytes it
moves) and then a few steps of the tuned 512^3 3D Star r=4 kernel, so one
counter pass measures both.

    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc --output-format csv -- \
        python -m threadcoarsenedstencilgenerator_amd.profile_run [--dtype f64|f32c] [--steps 3]
"""
from __future__ import annotations

import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--dtype", default="f64", choices=("f64", "f32c"))
    args = ap.parse_args()
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
    print(f"{d.key()} block {tuple(block)}: interior points {g.interior_points}, "
          f"algorithmic bytes per launch {2 * es * g.interior_points}", flush=True)


if __name__ == "__main__":
    main()
