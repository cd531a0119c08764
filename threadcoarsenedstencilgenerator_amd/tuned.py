"""Best measured kernel configuration per benchmark workload on MI355X.

Each entry is the code shape, coarsening and block shape that the block-shape
and coarsening sweeps found fastest (the measurements are summarised under
profiles/).  bench.py runs these.
"""
from __future__ import annotations

from . import manifest
from .spec import F32C, F64

# 3D Star Variable r=4, 512^3 interior, FP64, Write alignment at 16 points
CFG2_F64 = dict(desc=manifest.cfg2_desc(2, 4, 128, shape="dppw", max_threads=256).replace(prefetch=0),
                block=(128, 2, 1))
# the same workload in FP32 storage with compensated accumulation
# (x coarsening 2: FP32 kernels with 4 points per lane in x returned wrong values
# in lanes 12-15 of each 16-lane group on the test machines; the xc=2 kernels and all
# FP64 kernels are bit-exact over the full grid)
CFG2_F32C = dict(desc=manifest.cfg2_desc(2, 8, 64, shape="dppw", max_threads=256, dtype=F32C).replace(
    prefetch=0, comp_group=13), block=(128, 1, 1))


def best(workload: str, dtype: int = F64):
    if workload == "star3d_r4_512":
        return CFG2_F32C if dtype == F32C else CFG2_F64
    raise KeyError(workload)


def build_set():
    return [CFG2_F64["desc"], CFG2_F32C["desc"]]
