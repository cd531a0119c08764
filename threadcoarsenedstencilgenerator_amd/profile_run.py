"""Workload for counter profiling (rocprofv3 --pmc).

Runs, in one process, a plain device copy of 2^30 bytes (the calibration
kernel: FETCH_SIZE / WRITE_SIZE read for it are compared with the bytes it
moves) and then a few steps of the tuned 512^3 3D Star r=4 kernel, so one
counter pass measures both.

    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc --output-format csv -- \
        python -m threadcoarsenedstencilgenerator_amd.profile_run [--dtype f64|f32c] [--steps 3]
        (--workload box2d_r1_temporal: the temporal 2D Box r=1 kernel on 16384^2 instead)

With --ceiling it instead times the HBM stream probes (copy, read-only,
write-only; plain and nontemporal; one pass and grid-stride launches) over
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
            for wg in (0, 256 * 4, 256 * 8, 256 * 16, 256 * 32):
                ms = rt.hbm_probe(mode, a.data_ptr(), b.data_ptr(), nbytes, nt, wg, reps)
                line = dict(probe=mode, nontemporal=nt, workgroups=wg or "one pass", bytes_per_launch=moved,
                            ms=round(ms, 4), tbs=round(moved / ms / 1e9, 3))
                out.append(line)
                print(json.dumps(line), flush=True)
    best = {m: max(x["tbs"] for x in out if x["probe"] == m) for m in ("copy", "read", "write")}
    print(json.dumps({"hbm_ceiling_tbs": best, "peak_tbs": 8.0}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--dtype", default="f64", choices=("f64", "f32c"))
    ap.add_argument("--ceiling", action="store_true")
    ap.add_argument("--workload", default="star3d_r4_512", choices=("star3d_r4_512", "box2d_r1_temporal"))
    args = ap.parse_args()
    if args.ceiling:
        return ceiling()
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
