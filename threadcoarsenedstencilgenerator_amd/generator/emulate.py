"""Run generated STREAM kernels on the host through the wave-level CPU
emulation (cpu_emu.h): the same kernel text as the gfx950 build, with every
lane of a wave, its DPP moves and its range-checked buffer accesses
emulated.  Used by the CPU test suite to check the emitted code against the
oracle without a GPU; the GPU suite checks the compiled code objects."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

from .. import spec
from ..spec import F32C
from .emit import emit_source

HERE = os.path.dirname(os.path.abspath(__file__))
# host-only build cache (the emulation libraries never run on the GPU)
EMU_DIR = os.path.join(tempfile.gettempdir(), f"tcsg_emu_{os.getuid()}")


class _KArgs(C.Structure):
    _fields_ = [("in_", C.c_void_p), ("out", C.c_void_p), ("mask", C.c_void_p),
                ("xdim", C.c_uint32), ("ydim", C.c_uint32), ("zdim", C.c_uint32), ("right_offset", C.c_uint32),
                ("in_bytes", C.c_uint64), ("out_bytes", C.c_uint64), ("rec_limit", C.c_uint32), ("oob", C.c_uint32)]


def build(desc: spec.StencilDesc, edit=None) -> str:
    """Build (or reuse) the emulation library of ``desc``; ``edit``, a (name, function)
    pair, builds a separate library from the emitted text as the function rewrites it
    (tests use it to check that the emulation detects a broken kernel)."""
    os.makedirs(EMU_DIR, exist_ok=True)
    so = os.path.join(EMU_DIR, desc.key() + (f"_{edit[0]}" if edit else "") + ".so")
    src_text = emit_source(desc, "cpu")
    if edit:
        src_text = edit[1](src_text)
    # the stamp covers the emulation header too, so a changed header rebuilds
    with open(os.path.join(HERE, "cpu_emu.h")) as f:
        stamp_text = src_text + "\n// cpu_emu.h\n" + f.read()
    stamp = so + ".src"
    if os.path.exists(so) and os.path.exists(stamp) and open(stamp).read() == stamp_text:
        return so
    with tempfile.TemporaryDirectory(prefix="tcsg_emu_") as td:
        src = os.path.join(td, "k.cpp")
        open(src, "w").write(src_text)
        tmp = so + f".{os.getpid()}.tmp"
        r = subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-fPIC", "-shared", "-ffp-contract=off", f"-I{HERE}",
                            src, "-o", tmp], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"emulation build failed for {desc.key()}:\n{r.stderr[-4000:]}")
        os.replace(tmp, so)
    open(stamp, "w").write(stamp_text)
    return so


def stream_limits(desc: spec.StencilDesc, geom, bext: int):
    """(record-count cap, out-of-range base offset) of a STREAM block's buffer window
    for block extent `bext` along z (3D) or y (2D), as the runtime sets them
    (0xFFFFFFFF for other shapes)."""
    if desc.variant != spec.STREAM:
        return 0xFFFFFFFF, 0xFFFFFFFF
    es = 4 if desc.dtype == F32C else 8
    sx = geom.xdim * es
    sxy = sx * geom.ydim
    r = desc.range
    h = (r + desc.xc - 1) // desc.xc
    cols = (2 * h * desc.xc + desc.xc + 2 * r) * es
    # rows a lane's loads reach from its base: its footprint's, or with LDS staging the block's
    rows = (desc.stage[1] if desc.staged else 1) * desc.yc + 2 * r
    span = (desc.zc + 2 * r) * (sxy if desc.dimension == 3 else 0) + rows * sx + cols
    window = (bext * desc.zc + 2 * r) * sxy if desc.dimension == 3 else (bext * desc.yc + 2 * r) * sx
    slack = sx + cols
    assert window + 2 * slack + span < 1 << 32
    return window + slack, window + 2 * slack


def run(desc: spec.StencilDesc, geom, mask, a: np.ndarray, b: np.ndarray, steps: int, block, blocks, edit=None):
    """`steps` ping-pong steps of the emulated kernel on host arrays a/b
    (geometry as from spec.geometry).  Returns (final array, bad accesses of this
    call)."""
    L = C.CDLL(build(desc, edit))
    bad0 = C.c_ulong.in_dll(L, "tcsg_emu_bad_access").value
    L.tcsg_emulate.argtypes = [C.c_void_p] + [C.c_uint32] * 6
    dt = np.float32 if desc.dtype == F32C else np.float64
    m = np.ascontiguousarray(mask if mask is not None else np.zeros(1), dtype=dt)
    bufs = [a, b]
    nbytes = a.nbytes
    for s in range(steps):
        k = _KArgs(bufs[s & 1].ctypes.data, bufs[(s + 1) & 1].ctypes.data, m.ctypes.data,
                   geom.xdim, geom.ydim, geom.zdim, geom.ro, nbytes, nbytes,
                   *stream_limits(desc, geom, block[2] if desc.dimension == 3 else block[1]))
        rc = L.tcsg_emulate(C.byref(k), *blocks, *block)
        if rc != 0:
            raise ValueError("block shape must be whole waves")
    bad = C.c_ulong.in_dll(L, "tcsg_emu_bad_access").value - bad0
    return bufs[steps & 1], bad


def run_batched(desc: spec.StencilDesc, geoms, arrays0, arrays1, mask, steps: int, block, nwg: int):
    """Emulate tcsg_update_batched over several grids (tile table built as the
    runtime builds it: grid-major, x fastest; with desc.pack, runs of
    block x * block y threads over each grid's xy thread space per block z
    planes).  Returns (final arrays, bad)."""
    L = C.CDLL(build(desc))
    L.tcsg_emulate_batched.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32]
    dt = np.float32 if desc.dtype == F32C else np.float64
    m = np.ascontiguousarray(mask if mask is not None else np.zeros(1), dtype=dt)
    tiles = []
    for gi, g in enumerate(geoms):
        nb = [(g.threads[a] + block[a] - 1) // block[a] for a in range(3)]
        if desc.pack:
            per = block[0] * block[1]
            for z in range(nb[2]):
                for x in range((g.threads[0] * g.threads[1] + per - 1) // per):
                    tiles.append((gi, x, 0, z))
            continue
        for z in range(nb[2]):
            for y in range(nb[1]):
                for x in range(nb[0]):
                    tiles.append((gi, x, y, z))
    tiles = np.array(tiles, dtype=np.uint32)
    bufs = [list(arrays0), list(arrays1)]
    queue = np.zeros(16, dtype=np.uint32)   # the kernel's work queue, zero before the first launch
    for s in range(steps):
        ka = (_KArgs * len(geoms))()
        for i, g in enumerate(geoms):
            src, dst = bufs[s & 1][i], bufs[(s + 1) & 1][i]
            ka[i] = _KArgs(src.ctypes.data, dst.ctypes.data, m.ctypes.data, g.xdim, g.ydim, g.zdim, g.ro,
                           src.nbytes, dst.nbytes,
                           *stream_limits(desc, g, block[2] if desc.dimension == 3 else block[1]))
        if L.tcsg_emulate_batched(C.cast(ka, C.c_void_p), tiles.ctypes.data, len(tiles), queue.ctypes.data, nwg,
                                  *block) != 0:
            raise ValueError("block shape must be whole waves")
        if queue.any():
            raise AssertionError(f"work queue not cleared by the launch's last workgroup: {queue}")
    bad = C.c_ulong.in_dll(L, "tcsg_emu_bad_access").value
    return bufs[steps & 1], bad
