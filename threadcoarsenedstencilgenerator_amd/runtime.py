"""Host-side mirror of the reference's launch interface over libtcsg.so.

The reference harness (hip/benchmark.cpp) drives its kernels through
hip/launchUpdateGrid.h: copyMask, maxPotentialBlockSize, launchInit and
benchmarkUpdateGrid.  :class:`Kernel` exposes the same operations (plus
one-step and event-timed launches) on a loaded code object; errors raise
:class:`TcsgError` with the library's status.

Device memory comes from torch (plumbing only); every computation runs in
the native library and the generated kernels.  If libtcsg.so is missing the
import fails: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import spec
from .build import LIB, kernel_path

OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "parameter mismatch",
          4: "block shape", 5: "too large", 6: "no mask", 7: "file"}
INIT_REFERENCE, INIT_U01, INIT_UPM1, INIT_NORMAL = 0, 1, 2, 3
ERR_BLOCK_SHAPE = 4


class TcsgError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {_lib().tcsg_status_string(status).decode()} (status {status}, "
                         f"hip {_lib().tcsg_last_hip_error()})")


class Dim3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]

    def tup(self):
        return (self.x, self.y, self.z)


class CDesc(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("dimension", "neighbourhood", "orientation", "coeff", "range",
                                          "xc", "yc", "zc", "alignment", "cl_size", "dtype", "variant")]


class CGeometry(C.Structure):
    _fields_ = [("size", Dim3), ("elements", Dim3), ("threads", Dim3), ("left_offset", C.c_uint32),
                ("right_offset", C.c_uint32), ("dimension", C.c_uint32), ("range", C.c_uint32),
                ("n_points", C.c_uint64)]

    @property
    def interior_points(self):
        e = self.elements
        return e.x * e.y * e.z


_L = None


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} is not built; run __graft_entry__.build()")
        # torch brings its own HIP runtime library; loaded first, it is the one libtcsg binds
        # to as well.  Loaded after libtcsg (which would bring the system's), torch finds no
        # GPU.  One runtime serves both, so device pointers and streams pass between them.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB)
        p, u32, i, u64 = C.c_void_p, C.c_uint32, C.c_int, C.c_uint64
        L.tcsg_geometry_compute.argtypes = [C.POINTER(CDesc), Dim3, C.POINTER(CGeometry)]
        L.tcsg_geometry_check.argtypes = [C.POINTER(CDesc), C.POINTER(CGeometry)]
        L.tcsg_kernel_load.argtypes = [C.POINTER(p), C.c_char_p, C.POINTER(CDesc)]
        L.tcsg_kernel_free.argtypes = [p]
        L.tcsg_kernel_desc.argtypes = [p, C.POINTER(CDesc)]
        L.tcsg_kernel_geometry.argtypes = [p, Dim3, C.POINTER(CGeometry)]
        L.tcsg_kernel_remainder.argtypes = [p, C.POINTER(i)]
        L.tcsg_kernel_steps_per_launch.argtypes = [p, C.POINTER(i)]
        L.tcsg_set_mask.argtypes = [p, p, u32]
        L.tcsg_max_block_size.argtypes = [p, C.POINTER(i), C.POINTER(i)]
        L.tcsg_check_block.argtypes = [p, Dim3]
        L.tcsg_kernel_resources.argtypes = [p, u32, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        L.tcsg_grid_blocks.argtypes = [C.POINTER(CGeometry), Dim3]
        L.tcsg_grid_blocks.restype = Dim3
        L.tcsg_launch_init.argtypes = [p, C.POINTER(CGeometry), i, u64, p, p]
        L.tcsg_update_grid.argtypes = [p, C.POINTER(CGeometry), Dim3, Dim3, p, p, p]
        L.tcsg_run.argtypes = [p, C.POINTER(CGeometry), Dim3, Dim3, p, p, i, p, C.POINTER(i)]
        L.tcsg_benchmark_update_grid.argtypes = [p, C.POINTER(CGeometry), Dim3, Dim3, C.c_double, p, p,
                                                 C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_size_t)]
        L.tcsg_time_steps.argtypes = [p, C.POINTER(CGeometry), Dim3, Dim3, p, p, i, i, p,
                                      C.POINTER(C.c_float)]
        L.tcsg_batch_create.argtypes = [C.POINTER(p), p, u32, C.POINTER(CGeometry), C.POINTER(p), C.POINTER(p),
                                        Dim3, u32]
        L.tcsg_batch_free.argtypes = [p]
        L.tcsg_batch_run.argtypes = [p, i, i, p, C.POINTER(i)]
        L.tcsg_batch_run_separate.argtypes = [p, i, p, C.POINTER(i)]
        L.tcsg_batch_info.argtypes = [p, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32)]
        L.tcsg_error_f32_vs_f64.argtypes = [C.POINTER(CGeometry), p, C.POINTER(CGeometry), p, C.c_double,
                                            C.POINTER(C.c_double)]
        L.tcsg_error_f32_vs_f64_noalloc.argtypes = [C.POINTER(CGeometry), p, C.POINTER(CGeometry), p, C.c_double,
                                                    p, p]
        L.tcsg_convert_f64_to_f32.argtypes = [p, p, u64, p]
        L.tcsg_hbm_probe.argtypes = [i, i, p, p, u64, u32, i, p, C.POINTER(C.c_float)]
        L.tcsg_hbm_probe_ex.argtypes = [i, i, u32, u32, p, p, u64, u32, i, p, C.POINTER(C.c_float)]
        L.tcsg_mlp_num_params.argtypes = [u32, u32]
        L.tcsg_mlp_num_params.restype = C.c_size_t
        L.tcsg_mlp_create.argtypes = [C.POINTER(p), u32, u32, u32, u64]
        L.tcsg_mlp_free.argtypes = [p]
        L.tcsg_mlp_get_params.argtypes = [p, p, C.c_size_t]
        L.tcsg_mlp_set_params.argtypes = [p, p, C.c_size_t]
        L.tcsg_mlp_forward.argtypes = [p, p, u32, p, p]
        L.tcsg_mlp_gradients.argtypes = [p, p, p, u32, p, p]
        L.tcsg_mlp_train_step.argtypes = [p, p, p, u32, C.c_float, p, p]
        L.tcsg_status_string.argtypes = [i]
        L.tcsg_status_string.restype = C.c_char_p
        L.tcsg_last_hip_error.argtypes = []
        _L = L
    return _L


def library():
    """The loaded libtcsg.so (ctypes handle)."""
    return _lib()


def _check(rc, what):
    if rc != OK:
        raise TcsgError(rc, what)


def to_cdesc(d: spec.StencilDesc) -> CDesc:
    return CDesc(d.dimension, d.neighbourhood, d.orientation, d.coeff, d.range, d.xc, d.yc, d.zc,
                 d.alignment, d.cl_size, d.dtype, d.variant)


def dim3(t) -> Dim3:
    t = tuple(t) + (1,) * (3 - len(tuple(t)))
    return Dim3(*t)


def geometry(desc: spec.StencilDesc, size) -> CGeometry:
    """Padded geometry for a user grid size incl. halo (hip/benchmark.cpp:133-182)."""
    g = CGeometry()
    _check(_lib().tcsg_geometry_compute(C.byref(to_cdesc(desc)), dim3(size), C.byref(g)), "geometry")
    return g


def grid_blocks(g: CGeometry, block) -> tuple:
    return _lib().tcsg_grid_blocks(C.byref(g), dim3(block)).tup()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def mask_seeded(seed: int, n: int) -> np.ndarray:
    """Variable-mask weights U[-1,1) from SplitMix64 (counter = index),
    rescaled so that sum |w| = 1, in FP64."""
    M = (1 << 64) - 1
    w = np.empty(n, dtype=np.float64)
    for k in range(n):
        z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        w[k] = float(z >> 11) * 2.0 ** -52 - 1.0
    s = 0.0
    for k in range(n):
        s += abs(w[k])
    return w / s


class Kernel:
    """A loaded generated kernel (one stencil parameter set)."""

    def __init__(self, desc: spec.StencilDesc, path: str | None = None):
        self.desc = desc
        self.path = path or kernel_path(desc)
        h = C.c_void_p()
        _check(_lib().tcsg_kernel_load(C.byref(h), self.path.encode(), C.byref(to_cdesc(desc))),
               f"load {self.path}")
        self._h = h
        self.dtype = np.float32 if desc.dtype == spec.F32C else np.float64

    def close(self):
        if self._h:
            _lib().tcsg_kernel_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- launchUpdateGrid.h equivalents ---
    def set_mask(self, mask):
        """copyMask (hip/generateLaunchUpdateGrid.py:77-82)."""
        m = np.ascontiguousarray(mask, dtype=np.float64)
        _check(_lib().tcsg_set_mask(self._h, m.ctypes.data_as(C.c_void_p), m.size), "set_mask")

    def max_block_size(self):
        """(occupancy-optimal size as in maxPotentialBlockSize, largest legal size)."""
        a, b = C.c_int(), C.c_int()
        _check(_lib().tcsg_max_block_size(self._h, C.byref(a), C.byref(b)), "max_block_size")
        return a.value, b.value

    def block_ok(self, block) -> bool:
        return _lib().tcsg_check_block(self._h, dim3(block)) == OK

    def resources(self, threads: int = 0):
        """(VGPRs per lane, scratch bytes per lane, workgroups of `threads`
        threads one CU holds)."""
        v, s, w = C.c_int(), C.c_int(), C.c_int()
        _check(_lib().tcsg_kernel_resources(self._h, threads, C.byref(v), C.byref(s), C.byref(w)), "resources")
        return v.value, s.value, w.value

    def geometry(self, size) -> CGeometry:
        """Geometry this kernel launches on: the reference's truncated one, or
        with remainder handling every point of `size`."""
        g = CGeometry()
        _check(_lib().tcsg_kernel_geometry(self._h, dim3(size), C.byref(g)), "kernel_geometry")
        return g

    @property
    def remainder(self) -> bool:
        v = C.c_int()
        _check(_lib().tcsg_kernel_remainder(self._h, C.byref(v)), "kernel_remainder")
        return bool(v.value)

    @property
    def steps_per_launch(self) -> int:
        """Time steps one launch advances (temporal kernels: their tsteps)."""
        v = C.c_int()
        _check(_lib().tcsg_kernel_steps_per_launch(self._h, C.byref(v)), "kernel_steps_per_launch")
        return v.value

    def init(self, g: CGeometry, ptr: int, mode=INIT_REFERENCE, seed=1, stream=None):
        """launchInit (hip/generateLaunchUpdateGrid.py:150-158) and seeded fills."""
        _check(_lib().tcsg_launch_init(self._h, C.byref(g), mode, seed, ptr, _stream(stream)), "init")

    def update(self, g: CGeometry, block, blocks, src: int, dst: int, stream=None):
        _check(_lib().tcsg_update_grid(self._h, C.byref(g), dim3(block), dim3(blocks), src, dst,
                                       _stream(stream)), "update_grid")

    def run(self, g: CGeometry, block, blocks, d0: int, d1: int, steps: int, stream=None) -> int:
        fin = C.c_int()
        _check(_lib().tcsg_run(self._h, C.byref(g), dim3(block), dim3(blocks), d0, d1, steps,
                               _stream(stream), C.byref(fin)), "run")
        return fin.value

    def benchmark(self, g: CGeometry, block, blocks, max_time: float, d0: int, d1: int):
        """benchmarkUpdateGrid (hip/generateLaunchUpdateGrid.py:103-147): per-step durations [s]."""
        cap = 1 << 16
        buf = (C.c_double * cap)()
        n = C.c_size_t()
        _check(_lib().tcsg_benchmark_update_grid(self._h, C.byref(g), dim3(block), dim3(blocks), max_time,
                                                 d0, d1, buf, cap, C.byref(n)), "benchmark_update_grid")
        return [buf[i] for i in range(min(n.value, cap))]

    def time_steps(self, g: CGeometry, block, blocks, d0: int, d1: int, warmup: int, steps: int,
                   stream=None) -> float:
        ms = C.c_float()
        _check(_lib().tcsg_time_steps(self._h, C.byref(g), dim3(block), dim3(blocks), d0, d1, warmup, steps,
                                      _stream(stream), C.byref(ms)), "time_steps")
        return ms.value


def error_f32_vs_f64(ga: CGeometry, a_ptr: int, gb: CGeometry, b_ptr: int, floor: float = 1e-3):
    """FP64 error reduction on the GPU of FP32 state a (layout ga) against FP64
    state b (layout gb): (max|a-b|, max|b|, sum (a-b)^2, sum b^2, max pointwise rel)."""
    out = (C.c_double * 5)()
    _check(_lib().tcsg_error_f32_vs_f64(C.byref(ga), a_ptr, C.byref(gb), b_ptr, floor, out), "error_f32_vs_f64")
    return tuple(out)


def error_f32_vs_f64_noalloc(ga: CGeometry, a_ptr: int, gb: CGeometry, b_ptr: int, out_ptr: int,
                             floor: float = 1e-3, stream=None):
    """The same reduction without allocating: deterministic two passes through the
    library's static scratch, asynchronous on `stream`, the 5 results into the
    device array at out_ptr."""
    _check(_lib().tcsg_error_f32_vs_f64_noalloc(C.byref(ga), a_ptr, C.byref(gb), b_ptr, floor, out_ptr,
                                                _stream(stream)), "error_f32_vs_f64_noalloc")


def convert_f64_to_f32(src_ptr: int, dst_ptr: int, n: int, stream=None):
    _check(_lib().tcsg_convert_f64_to_f32(src_ptr, dst_ptr, n, _stream(stream)), "convert")


def hbm_probe(mode: str, src: int, dst: int, nbytes: int, nontemporal: bool = False, workgroups: int = 0,
              reps: int = 10, stream=None, unroll: int = 4, block: int = 256) -> float:
    """Milliseconds per launch of the copy / read / write bandwidth probe
    (`unroll` 16-byte elements per lane in flight, `block` threads per workgroup)."""
    ms = C.c_float()
    _check(_lib().tcsg_hbm_probe_ex({"copy": 0, "read": 1, "write": 2}[mode], int(nontemporal), unroll, block, src,
                                    dst, nbytes, workgroups, reps, _stream(stream), C.byref(ms)), "hbm_probe")
    return ms.value


class Grids:
    """Two device buffers for a geometry (ping-pong), allocated with torch."""

    def __init__(self, g: CGeometry, dtype, device="cuda"):
        import torch
        tdt = torch.float32 if dtype == np.float32 else torch.float64
        self.g = g
        self.bufs = [torch.empty(g.n_points, dtype=tdt, device=device) for _ in range(2)]

    def ptr(self, i):
        return self.bufs[i].data_ptr()

    def host(self, i) -> np.ndarray:
        return self.bufs[i].cpu().numpy()

    def upload(self, i, arr: np.ndarray):
        import torch
        self.bufs[i].copy_(torch.from_numpy(np.ascontiguousarray(arr)))


class Batch:
    """Many independent grids advanced by one persistent kernel per step
    (the kernel must be generated with ``batched=True``)."""

    def __init__(self, kernel: Kernel, geoms, bufs0, bufs1, block, workgroups: int = 0):
        n = len(geoms)
        G = (CGeometry * n)(*geoms)
        P = C.c_void_p * n
        h = C.c_void_p()
        _check(_lib().tcsg_batch_create(C.byref(h), kernel._h, n, G, P(*bufs0), P(*bufs1), dim3(block),
                                        workgroups), "batch_create")
        self._h = h
        self.kernel = kernel  # keeps the module alive

    def run(self, steps: int, graph: bool = True, stream=None) -> int:
        fin = C.c_int()
        _check(_lib().tcsg_batch_run(self._h, steps, int(graph), _stream(stream), C.byref(fin)), "batch_run")
        return fin.value

    def run_separate(self, steps: int, stream=None) -> int:
        fin = C.c_int()
        _check(_lib().tcsg_batch_run_separate(self._h, steps, _stream(stream), C.byref(fin)), "batch_run_separate")
        return fin.value

    def info(self):
        pts, tiles, wg = C.c_uint64(), C.c_uint32(), C.c_uint32()
        _check(_lib().tcsg_batch_info(self._h, C.byref(pts), C.byref(tiles), C.byref(wg)), "batch_info")
        return pts.value, tiles.value, wg.value

    def close(self):
        if self._h:
            _lib().tcsg_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_sizes(n: int, seed: int, lo: int, hi: int, coarsening, r: int):
    """Grid sizes (incl. halo) of the batched workload: each interior axis
    length log-uniform in [lo, hi] from SplitMix64(seed), rounded to a
    multiple of the coarsening factor."""
    import math
    M = (1 << 64) - 1
    out = []
    k = 0
    for _ in range(n):
        size = []
        for a in range(3):
            z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & M
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
            z ^= z >> 31
            k += 1
            u = (z >> 11) * 2.0 ** -53
            length = math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
            c = coarsening[a]
            length = max(c, int(round(length / c)) * c)
            size.append(length + 2 * r)
        out.append(tuple(size))
    return out
