"""CPU oracle for the thread-coarsened stencil path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package, and only as the reference the
results are compared with or the timed CPU baseline.  The product (``threadcoarsenedstencilgenerator_amd``)
never imports it.

The arithmetic lives in ``stencil_oracle.c`` (plain C + OpenMP, compiled with
``-ffp-contract=off``); each C function cites the reference file:line it
restates.  This module is a ctypes wrapper with numpy buffers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

BOX, STAR = 0, 1
ROW, COLUMN = 0, 1
JACOBI, VARIABLE = 0, 1
ALIGN_NONE, ALIGN_READ, ALIGN_WRITE = 0, 1, 2

_lib = None
_ref = None


def _host_has_v3() -> bool:
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return " avx2" in flags and " fma" in flags and " bmi2" in flags


def build(ref: bool = True) -> None:
    """Compile the oracle (and, when the reference tree is present, the
    reference's own primes.cpp, and its harness over the compiled-in launch
    units the package build wrote, into oracle/_ref)."""
    for target in ("all", "ref") if ref and os.path.isdir("/root/reference/hip") else ("all",):
        r = subprocess.run(["make", "-s", "-j4", "-C", HERE, target], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"make {target} in {HERE} failed:\n{r.stdout}\n{r.stderr}")


def lib():
    global _lib
    if _lib is None:
        name = "liboracle_v3.so" if _host_has_v3() else "liboracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        u64, ul, i, p = C.c_uint64, C.c_ulong, C.c_int, C.c_void_p
        L.oracle_stencil_size.argtypes = [i, i, i]
        L.oracle_mask_index.argtypes = [i, i, i, i, i, i]
        L.oracle_offsets.argtypes = [i, i, i, i, p, p]
        L.oracle_geometry.argtypes = [i, i, i, i, p, p, p]
        L.oracle_init_reference.argtypes = [i, i, ul, ul, ul, ul, ul, p]
        L.oracle_init_seeded.argtypes = [i, i, i, u64, C.c_double, ul, ul, ul, ul, ul, p]
        L.oracle_mask_seeded.argtypes = [u64, i, p]
        L.oracle_u01.argtypes = [u64, u64]
        L.oracle_u01.restype = C.c_double
        L.oracle_upm1.argtypes = [u64, u64]
        L.oracle_upm1.restype = C.c_double
        L.oracle_step_f64.argtypes = [i, i, i, i, i, ul, ul, ul, ul, ul, p, p, p]
        L.oracle_run_f64.argtypes = [i, i, i, i, i, ul, ul, ul, ul, ul, p, p, p, i]
        L.oracle_step_f32c.argtypes = [i, i, i, i, i, i, ul, ul, ul, ul, ul, p, p, p]
        L.oracle_divisors.argtypes = [C.c_uint, p]
        L.oracle_block_shapes.argtypes = [i, C.c_uint, C.c_uint, p, p, i]
        L.oracle_num_threads.argtypes = []
        _lib = L
    return _lib


def ref_lib():
    """The reference's own primes.cpp built into oracle/_ref (None if absent)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref_primes.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_prime_combinations.argtypes = [C.c_uint, C.c_void_p, C.c_int]
        L.ref_prime_combinations2d.argtypes = [C.c_uint, C.c_void_p, C.c_int]
        L.ref_prime_combinations3d.argtypes = [C.c_uint, C.c_void_p, C.c_int]
        _ref = L
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def stencil_size(dim, nb, r) -> int:
    return lib().oracle_stencil_size(dim, nb, r)


def mask_index(dim, nb, r, dx, dy, dz=0) -> int:
    return lib().oracle_mask_index(dim, nb, r, dx, dy, dz)


def offsets(dim, nb, orient, r):
    """(offsets[n,3] in summation order, mask index per offset)."""
    n = stencil_size(dim, nb, r)
    off = np.zeros((n, 3), dtype=np.int32)
    midx = np.zeros(n, dtype=np.int32)
    got = lib().oracle_offsets(dim, nb, orient, r, _ptr(off), _ptr(midx))
    assert got == n
    return off, midx


class Geometry:
    """Padded geometry: xdim, ydim, zdim, lo, ro, interior elements, threads."""

    def __init__(self, dim, r, align, cl, size, coarsening):
        sz = np.array(list(size) + [1] * (3 - len(size)), dtype=np.uint32)
        c = np.array(list(coarsening) + [1] * (3 - len(coarsening)), dtype=np.uint32)
        g = np.zeros(11, dtype=np.uint64)
        rc = lib().oracle_geometry(dim, r, align, cl, _ptr(sz), _ptr(c), _ptr(g))
        if rc != 0:
            raise ValueError("grid too small for the halo")
        self.dim, self.r = dim, r
        self.xdim, self.ydim, self.zdim = int(g[0]), int(g[1]), int(g[2])
        self.lo, self.ro = int(g[3]), int(g[4])
        self.elements = tuple(int(v) for v in g[5:8])
        self.threads = tuple(int(v) for v in g[8:11])

    @property
    def n_points(self):
        return self.xdim * self.ydim * self.zdim

    @property
    def interior_points(self):
        e = self.elements
        return e[0] * e[1] * e[2]

    def args(self):
        return (self.xdim, self.ydim, self.zdim, self.lo, self.ro)


def init_reference(geom: Geometry, a: np.ndarray):
    lib().oracle_init_reference(geom.dim, geom.r, *geom.args(), _ptr(a))


def init_seeded(geom: Geometry, mode: int, seed: int, a: np.ndarray, halo=0.0):
    """mode 0: U[0,1), 1: U[-1,1), 2: N(0,1) (Box-Muller)."""
    lib().oracle_init_seeded(geom.dim, geom.r, mode, seed, halo, *geom.args(), _ptr(a))


def mask_seeded(seed: int, n: int) -> np.ndarray:
    w = np.zeros(n, dtype=np.float64)
    lib().oracle_mask_seeded(seed, n, _ptr(w))
    return w


def step_f64(desc, geom: Geometry, mask, src: np.ndarray, dst: np.ndarray):
    m = np.ascontiguousarray(mask if mask is not None else np.zeros(1), dtype=np.float64)
    lib().oracle_step_f64(desc.dimension, desc.neighbourhood, desc.orientation,
                          desc.coeff, desc.range, *geom.args(), _ptr(m), _ptr(src), _ptr(dst))


def run_f64(desc, geom: Geometry, mask, a: np.ndarray, b: np.ndarray, steps: int) -> np.ndarray:
    """Ping-pong `steps` steps; returns the array holding the final state."""
    m = np.ascontiguousarray(mask if mask is not None else np.zeros(1), dtype=np.float64)
    k = lib().oracle_run_f64(desc.dimension, desc.neighbourhood, desc.orientation,
                             desc.coeff, desc.range, *geom.args(), _ptr(m), _ptr(a), _ptr(b), steps)
    return a if k == 0 else b


def step_f32c(desc, geom: Geometry, mask, src: np.ndarray, dst: np.ndarray):
    """One FP32-storage compensated step (group size desc.comp_group)."""
    m = np.ascontiguousarray(mask if mask is not None else np.zeros(1), dtype=np.float64)
    lib().oracle_step_f32c(desc.dimension, desc.neighbourhood, desc.orientation,
                           desc.coeff, desc.range, desc.comp_group, *geom.args(),
                           _ptr(m), _ptr(src), _ptr(dst))


def run_f32c(desc, geom: Geometry, mask, a: np.ndarray, b: np.ndarray, steps: int) -> np.ndarray:
    bufs = [a, b]
    for s in range(steps):
        step_f32c(desc, geom, mask, bufs[s & 1], bufs[(s + 1) & 1])
    return bufs[steps & 1]


def divisors(T: int):
    out = np.zeros(4096, dtype=np.uint32)
    n = lib().oracle_divisors(T, _ptr(out))
    return [int(v) for v in out[:n]]


def block_shapes(dim, max_threads=1024, warp=64, max_dims=(1024, 1024, 1024)):
    md = np.array(max_dims, dtype=np.uint32)
    cap = 1 << 16
    out = np.zeros(3 * cap, dtype=np.uint32)
    n = lib().oracle_block_shapes(dim, max_threads, warp, _ptr(md), _ptr(out), cap)
    return [tuple(int(v) for v in out[3 * i:3 * i + 3]) for i in range(n)]


def num_threads() -> int:
    return lib().oracle_num_threads()
