"""Stencil descriptor, neighbourhood, mask layout and geometry.

These are the parameters of the reference's generator CLIs
(hip/generate2DStencil_hip.py:28-60, hip/generate3DStencil_hip.py:29-63) and
harness (hip/benchmark.cpp:62-81,118-218).  The generator uses them to emit
kernels; the C runtime restates the geometry part in C for the launch path.
"""
from __future__ import annotations

import dataclasses
import hashlib

BOX, STAR = 0, 1
ROW, COLUMN = 0, 1
JACOBI, VARIABLE = 0, 1
ALIGN_NONE, ALIGN_READ, ALIGN_WRITE = 0, 1, 2
F64, F32C = 0, 1
PLAIN, STREAM, TEMPORAL = 0, 1, 2

NB_NAMES = {"Box": BOX, "Star": STAR}
ORIENT_NAMES = {"Row": ROW, "Column": COLUMN}
COEFF_NAMES = {"Jacobi": JACOBI, "Variable": VARIABLE}
ALIGN_NAMES = {"None": ALIGN_NONE, "Read": ALIGN_READ, "Write": ALIGN_WRITE}
DTYPE_NAMES = {"f64": F64, "f32c": F32C}
VARIANT_NAMES = {"plain": PLAIN, "stream": STREAM, "temporal": TEMPORAL}


def _inv(d):
    return {v: k for k, v in d.items()}


@dataclasses.dataclass(frozen=True)
class StencilDesc:
    """One generated kernel's compile-time parameters.

    dimension/neighbourhood/orientation/coeff/range/coarsening/alignment/
    cl_size mirror the reference generator argv.  dtype, variant and the
    tuning knobs (max_threads, xext, xcd_remap, nt_store, comp_group,
    prefetch, waves_per_eu, zfirst, batched, nt_load, order, remainder,
    edge_pairs, tsteps, stage, pack) have no reference counterpart: they select the MI355X
    code shape.  The TEMPORAL variant advances tsteps time steps per launch:
    each workgroup stages its tile of (64 xc) x (max_threads/64 yc) points
    plus tsteps*range halo points per side in LDS and runs the steps there
    (3D: marching through zc planes with a ring of 2r+1 planes per step).
    """

    dimension: int = 3
    neighbourhood: int = STAR
    orientation: int = ROW
    coeff: int = VARIABLE
    range: int = 1
    xc: int = 1
    yc: int = 1
    zc: int = 1
    alignment: int = ALIGN_NONE
    cl_size: int = 0
    dtype: int = F64
    variant: int = PLAIN
    max_threads: int = 1024
    xext: str = "load"       # how x-neighbours outside the thread's columns are obtained:
    #                          "load" global loads, "dpp" lane shifts (any block shape),
    #                          "dppw" lane shifts for blocks whose x extent is a multiple of 64
    xcd_remap: bool = False  # remap block ids so neighbouring tiles share an XCD's L2
    nt_store: bool = False   # nontemporal stores of the outputs
    comp_group: int = 5      # F32C: terms summed per group before the TwoSum combine
    prefetch: int = 1        # STREAM: planes whose loads are issued ahead of the arithmetic
    waves_per_eu: int = 0    # STREAM: occupancy the compiler must allow (0: its own choice)
    zfirst: bool = False     # with xcd_remap: order tiles z-fastest, so z-neighbour tiles (which
    #                          share halo planes) run together on one XCD
    batched: bool = False    # also emit tcsg_update_batched, a persistent kernel over a
    #                          tile table spanning many grids
    nt_load: bool = False    # STREAM: nontemporal hint on the input loads
    remainder: bool = False  # grids whose interior is not a multiple of the coarsening keep
    #                          every point: the last thread of an axis starts early enough to
    #                          end at the interior's end (STREAM kernels in x instead drop the
    #                          columns past the interior); off, axes are truncated as the
    #                          reference does
    edge_pairs: bool = False  # dppw: the edge columns of two adjacent rows in one load (the
    #                          second row's in the next edge lanes, moved by a DPP row shift)
    tsteps: int = 1          # TEMPORAL: time steps per launch
    order: str = "point"     # statement order: "point" loads each footprint point once and adds it
    #                          to every output that uses it; "offset" (PLAIN, and 2D STREAM without
    #                          prefetch) walks the mask in summation order, every output per offset
    #                          (weights used in order, each loaded point kept until its last use)
    stage: tuple = (0, 0, 0)  # STREAM 3D: LDS halo staging for the block shape (bx, by, bz) the
    #                          kernel is then launched with: each z-slice of the block loads the
    #                          rows of every plane once, with the y halo, into LDS (two buffers),
    #                          and each thread reads its footprint rows from there instead of
    #                          loading the y-arm rows its neighbours also load; (0, 0, 0): off
    pack: bool = False       # batched 3D: a tile is a run of bx*by consecutive threads of the grid's
    #                          xy thread space, taken x-fastest (a run continues on the next thread
    #                          row), times bz thread planes, instead of a bx x by rectangle: grids
    #                          whose thread rows the block's x extent does not divide leave no
    #                          lanes idle.  Lanes must not exchange values (PLAIN, or STREAM with
    #                          x-neighbours by loads)

    def __post_init__(self):
        object.__setattr__(self, "stage", tuple(self.stage))
        if self.dimension not in (2, 3):
            raise ValueError("dimension must be 2 or 3")
        if self.neighbourhood not in (BOX, STAR):
            raise ValueError("unknown stencil neighbourhood")
        if self.orientation not in (ROW, COLUMN):
            raise ValueError("unknown access orientation")
        if self.coeff not in (JACOBI, VARIABLE):
            raise ValueError("unknown coefficient type")
        if self.range < 1:
            raise ValueError("range has to be at least 1")
        if self.xc < 1 or self.yc < 1 or self.zc < 1:
            raise ValueError("coarsening factors have to be at least 1")
        if self.dimension == 2 and self.zc != 1:
            raise ValueError("2D stencils have no z coarsening")
        if self.alignment not in (ALIGN_NONE, ALIGN_READ, ALIGN_WRITE):
            raise ValueError("unknown alignment")
        if self.alignment != ALIGN_NONE and self.cl_size <= 0:
            raise ValueError("size of alignment has to be > 0")
        if self.xext not in ("load", "dpp", "dppw"):
            raise ValueError("xext must be 'load', 'dpp' or 'dppw'")
        if self.prefetch < 0 or self.waves_per_eu < 0 or self.waves_per_eu > 8:
            raise ValueError("prefetch must be >= 0 and waves_per_eu in [0, 8]")
        if self.max_threads % 64 or not 64 <= self.max_threads <= 1024:
            raise ValueError("max_threads must be a multiple of 64 in [64, 1024]")
        if self.comp_group < 1:
            raise ValueError("comp_group must be >= 1")
        if self.order not in ("point", "offset"):
            raise ValueError("order must be 'point' or 'offset'")
        if self.order == "offset" and self.variant != PLAIN and not (
                self.variant == STREAM and self.dimension == 2 and not self.staged and self.prefetch == 0):
            raise ValueError("the offset order is a statement order of PLAIN kernels and of 2D STREAM "
                             "kernels without LDS staging or register prefetch")
        if self.edge_pairs and (self.variant != STREAM or self.xext != "dppw"):
            raise ValueError("edge_pairs is an option of the wave-aligned DPP code shape")
        if self.variant not in (PLAIN, STREAM, TEMPORAL):
            raise ValueError("unknown code variant")
        if self.variant != TEMPORAL and self.tsteps != 1:
            raise ValueError("tsteps > 1 needs the temporal variant")
        if len(self.stage) != 3 or any(not isinstance(v, int) or v < 0 for v in self.stage):
            raise ValueError("stage is a block shape (bx, by, bz), or (0, 0, 0)")
        if self.staged:
            bx, by, bz = self.stage
            if (self.variant != STREAM or self.dimension != 3 or self.xext != "dppw"
                    or self.batched or self.prefetch):
                raise ValueError("LDS staging is an option of the 3D wave-aligned DPP STREAM shape "
                                 "(without batching or register prefetch)")
            if bx % 64 or not 64 <= bx <= 1024 or not 1 <= by <= 31 or not 1 <= bz <= 31:
                raise ValueError("staged block: bx a multiple of 64 up to 1024, by and bz in [1, 31]")
            if bx * by * bz > self.max_threads:
                raise ValueError("staged block has more threads than the launch bound")
            if self.stage_lds_bytes > 163840:
                raise ValueError(f"staged block needs {self.stage_lds_bytes} bytes of LDS (gfx950: at most 160 KiB)")
        if self.pack and (not self.batched or self.dimension != 3 or self.staged
                          or not (self.variant == PLAIN or (self.variant == STREAM and self.xext == "load"))):
            raise ValueError("pack is an option of batched 3D kernels whose lanes exchange no values "
                             "(PLAIN, or STREAM with x-neighbours by loads)")
        if self.variant == TEMPORAL:
            if not 1 <= self.tsteps <= 16:
                raise ValueError("tsteps must be in [1, 16]")
            if (self.xext != "load" or self.xcd_remap or self.nt_store or self.nt_load or self.zfirst
                    or self.batched or self.order != "point"):
                raise ValueError("the temporal variant takes none of the STREAM/PLAIN code-shape options")
            if self.lds_bytes > 163840:
                raise ValueError(f"temporal tile needs {self.lds_bytes} bytes of LDS (gfx950: at most 160 KiB)")

    # ---- derived quantities ----
    @property
    def n(self) -> int:
        """Stencil size, hip/benchmark.cpp:62-73."""
        r = self.range
        if self.neighbourhood == BOX:
            return (2 * r + 1) ** self.dimension
        return 2 * self.dimension * r + 1

    @property
    def flops_per_update(self) -> int:
        """hip/benchmark.cpp:75-80: Jacobi n, Variable 2n-1."""
        return self.n if self.coeff == JACOBI else 2 * self.n - 1

    @property
    def left_offset(self) -> int:
        """Write alignment pads rows on the left so the first interior point
        starts a line: hip/benchmark.h:110-116, hip/generate2DStencil_hip.py:83-85."""
        if self.alignment == ALIGN_WRITE and self.range % self.cl_size:
            return self.cl_size - self.range % self.cl_size
        return 0

    @property
    def tile(self):
        """TEMPORAL: output points per workgroup in x and y, and the LDS tile
        (outputs plus tsteps*range halo points per side).  3D workgroups march
        through zc planes of such tiles."""
        h = self.tsteps * self.range
        tx, ty = 64 * self.xc, self.max_threads // 64 * self.yc
        return tx, ty, tx + 2 * h, ty + 2 * h

    @property
    def lds_bytes(self) -> int:
        """TEMPORAL: LDS per workgroup.  2D: two buffers of the LDS tile.  3D: a ring of
        2r+1 planes per level s < tsteps, level s covering the tile plus (tsteps - s) r
        points per side."""
        tx, ty, lw, lh = self.tile
        es = 4 if self.dtype == F32C else 8
        if self.dimension == 2:
            return 2 * lw * lh * es
        r, k = self.range, self.tsteps
        return sum((2 * r + 1) * (tx + 2 * (k - s) * r) * (ty + 2 * (k - s) * r) for s in range(k)) * es

    @property
    def staged(self) -> bool:
        return tuple(self.stage) != (0, 0, 0)

    @property
    def stage_lds_bytes(self) -> int:
        """STREAM with LDS staging: two buffers per z-slice of the block, each the
        block's rows plus 2r halo rows by its columns."""
        if not self.staged:
            return 0
        bx, by, bz = self.stage
        es = 4 if self.dtype == F32C else 8
        return bz * 2 * (by * self.yc + 2 * self.range) * bx * self.xc * es

    @property
    def coarsening(self):
        return (self.xc, self.yc, self.zc)

    def key(self) -> str:
        """Stable short name used for code-object files."""
        d = dataclasses.asdict(self)
        for f, default in _LATER_FIELDS.items():   # keep the names of kernels that predate these knobs
            if d[f] == default:
                del d[f]
        h = hashlib.sha1(repr(sorted(d.items())).encode()).hexdigest()[:10]
        nb = "box" if self.neighbourhood == BOX else "star"
        return (f"{self.dimension}d_{nb}_r{self.range}_{_inv(COEFF_NAMES)[self.coeff].lower()}"
                f"_{_inv(ORIENT_NAMES)[self.orientation].lower()}_c{self.xc}x{self.yc}x{self.zc}"
                f"_{_inv(ALIGN_NAMES)[self.alignment].lower()}{self.cl_size}"
                f"_{_inv(DTYPE_NAMES)[self.dtype]}_{_inv(VARIANT_NAMES)[self.variant]}"
                f"{'' if self.variant == PLAIN else '_' + self.xext if self.variant == STREAM else f'_k{self.tsteps}'}"
                f"_{h}")

    def params_blob(self):
        """Integers baked into the code object and checked by the runtime
        before launch (the reference's compiled-in sanity check,
        hip/generateLaunchUpdateGrid.py:106-119)."""
        return [0x54435347, 1, self.dimension, self.neighbourhood, self.orientation,
                self.coeff, self.range, self.xc, self.yc, self.zc, self.alignment,
                self.cl_size, self.dtype, self.variant, self.n, self.max_threads,
                {"load": 0, "dpp": 1, "dppw": 2}[self.xext], int(self.xcd_remap),
                int(self.nt_store) | int(self.nt_load) << 1 | int(self.remainder) << 2 | int(self.staged) << 3
                | (self.tsteps if self.variant == TEMPORAL else 0) << 8
                | (self.stage[0] // 64) << 16 | self.stage[1] << 21 | self.stage[2] << 26,
                self.comp_group, self.prefetch, self.waves_per_eu, int(self.zfirst), int(self.batched) | int(self.pack) << 1]

    def replace(self, **kw) -> "StencilDesc":
        return dataclasses.replace(self, **kw)


# knobs added after kernel names were first recorded (measured corpora refer to
# kernels by name): left out of the name's hash while at their default
_LATER_FIELDS = {"nt_load": False, "order": "point", "remainder": False, "edge_pairs": False, "tsteps": 1,
                 "stage": (0, 0, 0), "pack": False}


def mask_index(desc: StencilDesc, dx: int, dy: int, dz: int = 0):
    """Position of offset (dx,dy,dz) in the mask: rank in (dz,dy,dx)
    lexicographic order over the neighbourhood.  Box: row-major
    (hip/generate2DStencil_hip.py:124-128, hip/generate3DStencil_hip.py:137-142);
    Star: the arm blocks of calcMaskAddress (2D :129-144, 3D :143-169).
    Returns None for offsets outside the neighbourhood."""
    r = desc.range
    if desc.dimension == 2:
        if dz:
            return None
    if max(abs(dx), abs(dy), abs(dz)) > r:
        return None
    w = 2 * r + 1
    if desc.neighbourhood == BOX:
        if desc.dimension == 2:
            return (dx + r) + (dy + r) * w
        return (dx + r) + (dy + r) * w + (dz + r) * w * w
    if (dx != 0) + (dy != 0) + (dz != 0) > 1:
        return None
    if desc.dimension == 2:
        if dy < 0:
            return dy + r
        if dy == 0:
            return r + dx + r
        return dy + r + 2 * r
    if dz < 0:
        return dz + r
    if dz == 0 and dy < 0:
        return dy + 2 * r
    if dz == 0 and dy == 0:
        return dx + 3 * r
    if dz == 0:
        return dy + 4 * r
    return dz + 5 * r


def in_neighbourhood(desc: StencilDesc, dx, dy, dz=0) -> bool:
    return mask_index(desc, dx, dy, dz) is not None


def sort_key(desc: StencilDesc, x, y, z=0):
    """Traversal order key.  Row visits (z, y, x) lexicographically, Column
    (z, x, y): the order of the reference's loop nests
    (hip/generate2DStencil_hip.py:245-298, hip/generate3DStencil_hip.py:266-361).
    Translating by an output position keeps lexicographic order, so this is
    also each output's summation order over its offsets."""
    return (z, y, x) if desc.orientation == ROW else (z, x, y)


def offsets(desc: StencilDesc):
    """Neighbourhood offsets (dx,dy,dz) in per-output summation order."""
    r = desc.range
    zr = r if desc.dimension == 3 else 0
    pts = [(dx, dy, dz) for dz in range(-zr, zr + 1) for dy in range(-r, r + 1)
           for dx in range(-r, r + 1) if in_neighbourhood(desc, dx, dy, dz)]
    pts.sort(key=lambda p: sort_key(desc, *p))
    assert len(pts) == desc.n
    return pts


def jacobi_c(desc: StencilDesc) -> float:
    """fl(1/n), the literal the reference prints via Python repr
    (hip/generate2DStencil_hip.py:300-307, hip/generate3DStencil_hip.py:363-370)."""
    return 1.0 / desc.n


@dataclasses.dataclass
class Geometry:
    """Padded grid geometry (hip/benchmark.cpp:140-177, hip/benchmark.h:100-123).

    size      user grid size including the halo, per axis
    xdim..    padded allocation extents (x includes left/right padding)
    lo, ro    left / right padding in points
    elements  interior (updated) points per axis
    threads   threads per axis = elements / coarsening, rounded up (only
              remainder-handling kernels see interiors that are not multiples)
    """

    dimension: int
    range: int
    xdim: int
    ydim: int
    zdim: int
    lo: int
    ro: int
    elements: tuple
    threads: tuple

    @property
    def n_points(self) -> int:
        return self.xdim * self.ydim * self.zdim

    @property
    def interior_points(self) -> int:
        e = self.elements
        return e[0] * e[1] * e[2]


def geometry(desc: StencilDesc, size) -> Geometry:
    r = desc.range
    size = list(size) + [1] * (3 - len(size))
    c = desc.coarsening
    t = [1, 1, 1]
    e = [1, 1, 1]
    for a in range(desc.dimension):
        if size[a] <= 2 * r:
            raise ValueError(f"grid axis {a} ({size[a]}) too small for range {r}")
        t[a] = size[a] if desc.remainder else size[a] - (size[a] - 2 * r) % c[a]
        e[a] = t[a] - 2 * r
        if desc.remainder and (a > 0 or desc.variant == PLAIN) and desc.variant != TEMPORAL and e[a] < c[a]:
            raise ValueError("remainder handling needs an interior of at least one coarsening "
                             + ("per axis" if desc.variant == PLAIN else "in y and z"))
    lo = ro = 0
    xmem = t[0]
    cl = desc.cl_size
    if desc.alignment == ALIGN_READ:
        if xmem % cl:
            ro = cl - xmem % cl
            xmem += ro
    elif desc.alignment == ALIGN_WRITE:
        lo = desc.left_offset
        xmem += lo
        if xmem % cl:
            ro = cl - xmem % cl
        xmem += ro
    thr = tuple(-(-e[a] // c[a]) if a < desc.dimension else 1 for a in range(3))
    return Geometry(desc.dimension, r, xmem, t[1], t[2] if desc.dimension == 3 else 1,
                    lo, ro, tuple(e), thr)
