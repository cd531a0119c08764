"""The error reduction's entry point that allocates nothing
(tcsg_error_f32_vs_f64_noalloc: partial results in the library's static device
scratch) against the allocating one (tcsg_error_f32_vs_f64), through the C ABI:
two calls in a row, the second on a smaller grid, each equal to the allocating
entry point's results (maxima exactly, sums to 1e-12 relative: the allocating
one adds its partial sums with atomics in no fixed order)."""
import ctypes as C

import numpy as np
import pytest

from threadcoarsenedstencilgenerator_amd import spec, tuned

pytestmark = pytest.mark.gpu


def _states(size, seed):
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d64, d32 = tuned.CFG2_F64["desc"], tuned.CFG2_F32C["desc"]
    k64, k32 = rt.Kernel(d64), rt.Kernel(d32)
    g, gf = k64.geometry(size), k32.geometry(size)
    s64, s32 = rt.Grids(g, np.float64), rt.Grids(gf, np.float32)
    k64.init(g, s64.ptr(0), rt.INIT_NORMAL, seed)
    k32.init(gf, s32.ptr(0), rt.INIT_NORMAL, seed + 1)   # a different state: nonzero errors
    torch.cuda.synchronize()
    return g, gf, s64, s32, (k64, k32)


@pytest.mark.parametrize("floor", [1e-3, 1.0])
def test_noalloc_matches_allocating_entry_point_twice(floor):
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    L = rt.library()
    out = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
    for size, seed in (((136, 72, 264), 3), ((40, 24, 136), 9)):   # the second call on a smaller grid
        g, gf, s64, s32, ks = _states(size, seed)
        print(f"noalloc error reduction: grid {size}, interior {g.interior_points} points, floor {floor}", flush=True)
        want = rt.error_f32_vs_f64(gf, s32.ptr(0), g, s64.ptr(0), floor)
        rt.error_f32_vs_f64_noalloc(gf, s32.ptr(0), g, s64.ptr(0), out.data_ptr(), floor)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert got[0] == want[0] and got[1] == want[1] and got[4] == want[4]
        assert got[2] == pytest.approx(want[2], rel=1e-12) and got[3] == pytest.approx(want[3], rel=1e-12)
        assert got[0] > 0 and got[3] > 0
        # deterministic: a repeat gives the same bits
        rt.error_f32_vs_f64_noalloc(gf, s32.ptr(0), g, s64.ptr(0), out.data_ptr(), floor)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().view(np.uint64), got.view(np.uint64))
        del s64, s32


def test_noalloc_refuses_interiors_beyond_its_scratch():
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    L = rt.library()
    d = tuned.CFG2_F64["desc"]
    g = rt.geometry(d, (1032, 1032, 1160))          # 1024 x 1024 x 1152 > 2^30 interior points
    assert g.interior_points > 1 << 30
    rc = L.tcsg_error_f32_vs_f64_noalloc(C.byref(g), C.c_void_p(16), C.byref(g), C.c_void_p(16), 1e-3,
                                         C.c_void_p(16), None)
    assert rc == 5   # TCSG_ERR_TOO_LARGE, before any launch
