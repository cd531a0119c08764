"""Writes tests/golden/oracle_parity.json: SHA-256 digests of the CPU
oracle's final grids for the small-grid parity matrix (every case of
manifest.parity_cases(), manifest.f32c_cases() and manifest.remainder_cases(),
3 steps) and for the
2D Box Jacobi r=1 1024x1024 workload (10 steps), from the seeded inputs the
GPU parity tests use.

The reference ships no golden vectors, so these digests are regression pins
of the oracle (whose arithmetic the other CPU tests tie to closed forms and
exact-arithmetic brute force): the CPU suite recomputes them, and the GPU
parity tests compare the generated kernels' final grids with them.

    python tests/golden/make_oracle_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from threadcoarsenedstencilgenerator_amd import manifest, spec  # noqa: E402

OUT = os.path.join(HERE, "oracle_parity.json")
INPUT_SEED, MASK_SEED = 1, 3   # tests/parity.py run_case defaults


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def final_grid(desc, size, steps, f32_mode=False):
    """Oracle final grid from the runtime's seeded fill: U[0,1) for FP64 cases,
    U[-1,1) rounded to FP32 for FP32-storage cases (the fills tests/parity.py
    asks the GPU for), halo and padding 0; Variable masks from MASK_SEED.
    Remainder-handling kernels keep every interior point: the oracle runs on
    the untruncated grid (coarsening 1)."""
    c = (1, 1, 1) if desc.remainder else desc.coarsening
    og = oracle.Geometry(desc.dimension, desc.range, desc.alignment, desc.cl_size, size, c)
    a = np.zeros(og.n_points)
    oracle.init_seeded(og, 1 if f32_mode else 0, INPUT_SEED, a)
    mask = oracle.mask_seeded(MASK_SEED, desc.n) if desc.coeff == spec.VARIABLE else None
    if desc.dtype == spec.F32C:
        a = a.astype(np.float32)
        return oracle.run_f32c(desc, og, mask, a.copy(), a.copy(), steps)
    return oracle.run_f64(desc, og, mask, a.copy(), a.copy(), steps)


def cases():
    """(name, desc, grid size, steps, FP32 fill) of every pinned run."""
    out = [(d.key(), d, size, 3, False) for d, size in manifest.parity_cases()]
    out += [(d.key(), d, size, 3, True) for d, size in manifest.f32c_cases()]
    out += [(d.key(), d, size, 3, d.dtype == spec.F32C) for d, size in manifest.remainder_cases()]
    out.append(("box2d_jacobi_r1_1024_10steps", manifest.cfg1_desc("plain"), (1026, 1026), 10, False))
    return out


def compute():
    return {name: {"size": list(size), "steps": steps, "sha256": digest(final_grid(d, size, steps, f32))}
            for name, d, size, steps, f32 in cases()}


if __name__ == "__main__":
    oracle.build(ref=False)
    with open(OUT, "w") as f:
        json.dump(compute(), f, indent=1, sort_keys=True)
    print(f"wrote {OUT}")
