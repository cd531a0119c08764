"""HIP (gfx950) source emitter for one thread-coarsened stencil kernel.

One call of :func:`emit_source` produces a translation unit with

* ``tcsg_update`` -- one time step over the interior (the role of the
  reference's generated ``updateGrid``, hip/generate2DStencil_hip.py:228-319,
  hip/generate3DStencil_hip.py:242-384), taking a ``tcsg_kargs`` struct;
* ``tcsg_params`` -- the compile-time parameters, which the runtime compares
  with the caller's descriptor before any launch (the reference's compiled-in
  sanity check, hip/generateLaunchUpdateGrid.py:106-119);
* optionally ``tcsg_update_batched`` (the persistent kernel over a tile
  table) and the reference kernel file's ``updateGrid``, ``init``, ``d_mask``
  and ``range``.

Arithmetic contract, shared by every variant:  each output point sums its
neighbourhood in the reference's traversal order (Row (dz,dy,dx), Column
(dz,dx,dy), both lexicographic), Variable as ``acc = v0*w0`` then
``acc = fma(v, w, acc)``, Jacobi as ``acc = v0``, ``acc = acc + v``,
``out = acc * fl(1/n)``.  FMAs are explicit and the file is compiled with
``-ffp-contract=off``, so the result does not depend on compiler contraction.

Variants
--------
PLAIN   the reference's code shape: each thread owns an xc*yc*zc block,
        loads every footprint point once into a register (global loads,
        64-bit addressing), accumulates in traversal order, writes back.
STREAM  the MI355X shape of the same computation:
        * branch-free: buffer-resource loads/stores with 32-bit byte offsets
          (loads of lanes outside the interior return 0, their stores are
          dropped by the range check), so the compiler can pipeline the
          loads of the next plane across the stores of the previous one;
        * planes are consumed front to back and every output is stored as
          soon as its last contributor has been added (z-marching register
          window of 2r+1 planes);
        * ``xext="dpp"``: x-neighbours outside the thread's columns come from
          neighbouring lanes through wave64 DPP ``wave_shr:1``/``wave_shl:1``
          lane shifts.  A run of lanes in one block row and one wave (a
          segment) needs r columns beyond each of its ends; the H = ceil(r/xc)
          lanes at each end load those with one extra load per row, and a
          second DPP chain hands them to the lanes that need them, so no lane
          loads an x-neighbour individually.  ``xext="dppw"`` (block rows of
          whole waves) fetches both ends' columns with one load per row, each
          end's lanes using their own offset;
        * optional XCD-aware block remapping (consecutive tiles on one XCD's
          L2) and nontemporal stores.
"""
from __future__ import annotations

from collections import defaultdict

from .. import spec
from ..spec import (ALIGN_NONE, ALIGN_WRITE, BOX, F32C, F64, PLAIN, ROW,
                    STREAM, TEMPORAL, VARIABLE, StencilDesc)


DPP_WAVE_SHL1 = 0x130  # lane i receives lane i+1
DPP_WAVE_SHR1 = 0x138  # lane i receives lane i-1
DPP_ROW_SHL = 0x100     # + n: lane i receives lane i+n within its row of 16 lanes
DPP_ROW_SHR = 0x110     # + n: lane i receives lane i-n within its row of 16 lanes


def _n(v: int) -> str:
    """Identifier-safe signed integer."""
    return f"m{-v}" if v < 0 else str(v)


def _hexf(x: float, f32: bool) -> str:
    if f32:
        import struct
        xf = struct.unpack("f", struct.pack("f", x))[0]
        return float(xf).hex() + "f"
    return float(x).hex()


class _Emitter:
    def __init__(self, d: StencilDesc, target: str = "gfx950"):
        self.d = d
        self.target = target
        self.batched_body = False
        self.f32 = d.dtype == F32C
        self.T = "float" if self.f32 else "double"
        self.R = d.range
        self.XC, self.YC, self.ZC = d.xc, d.yc, d.zc
        self.three = d.dimension == 3
        self.offs = spec.offsets(d)
        self.tindex = {o: t for t, o in enumerate(self.offs)}
        self.L = []

    # ------------------------------------------------------------------ util
    def w(self, s: str = ""):
        self.L.append(s)

    def outputs(self):
        zr = range(self.ZC) if self.three else range(1)
        return [(i, j, k) for k in zr for j in range(self.YC) for i in range(self.XC)]

    def footprint(self):
        """Thread-relative points with the outputs they feed, in traversal order."""
        contrib = defaultdict(list)
        for o in self.outputs():
            for (dx, dy, dz) in self.offs:
                p = (o[0] + dx, o[1] + dy, o[2] + dz)
                contrib[p].append((o, self.tindex[(dx, dy, dz)],
                                   spec.mask_index(self.d, dx, dy, dz)))
        pts = sorted(contrib, key=lambda p: spec.sort_key(self.d, *p))
        return pts, contrib

    def wname(self, m):
        return f"w{m}"

    # ----------------------------------------------------------- accumulate
    def acc_stmt(self, o, t, m, v):
        """Statement adding term t (value expr v, mask index m) to output o."""
        i, j, k = o
        a = f"r_{i}_{j}_{k}"
        var = self.d.coeff == VARIABLE
        n = self.d.n
        if not self.f32:
            if t == 0:
                return f"TV {a} = {v} * {self.wname(m)};" if var else f"TV {a} = {v};"
            return f"{a} = tcsg_fma({v}, {self.wname(m)}, {a});" if var else f"{a} = {a} + {v};"
        # F32C: groups of g terms summed with one FMA chain, groups combined by TwoSum
        g = self.d.comp_group
        gi = t // g
        q, s, c = f"q{gi}_{i}_{j}_{k}", f"s_{i}_{j}_{k}", f"c_{i}_{j}_{k}"
        out = []
        if t % g == 0:
            out.append(f"TV {q} = {v} * {self.wname(m)};" if var else f"TV {q} = {v};")
        else:
            out.append(f"{q} = tcsg_fma({v}, {self.wname(m)}, {q});" if var else f"{q} = {q} + {v};")
        if t % g == g - 1 or t == n - 1:
            if gi == 0:
                out.append(f"TV {s} = {q};")
            elif gi == 1:
                out.append(f"TV {c}; tcsg_two_sum({s}, {q}, {s}, {c});")
            else:
                out.append(f"{{ TV e_; tcsg_two_sum({s}, {q}, {s}, e_); {c} = {c} + e_; }}")
        return " ".join(out)

    def final_expr(self, o):
        i, j, k = o
        n = self.d.n
        var = self.d.coeff == VARIABLE
        if not self.f32:
            a = f"r_{i}_{j}_{k}"
            return a if var else f"{a} * {_hexf(1.0 / n, False)}"
        s, c = f"s_{i}_{j}_{k}", f"c_{i}_{j}_{k}"
        ngroups = (n + self.d.comp_group - 1) // self.d.comp_group
        tot = f"({s} + {c})" if ngroups > 1 else s
        return tot if var else f"{tot} * {_hexf(1.0 / n, True)}"

    # ------------------------------------------------------------ preamble
    def preamble(self):
        d = self.d
        T = self.T
        w = self.w
        cpu = self.target == "cpu"
        w("// Generated by threadcoarsenedstencilgenerator_amd.generator -- do not edit.")
        nbname = "Box" if d.neighbourhood == BOX else "Star"
        w(f"// {d.dimension}D {nbname} r={d.range} "
          f"{'Variable' if d.coeff == VARIABLE else 'Jacobi'} "
          f"{'Row' if d.orientation == ROW else 'Column'} coarsening {d.xc}x{d.yc}x{d.zc} "
          f"alignment {['None', 'Read', 'Write'][d.alignment]}/{d.cl_size} "
          f"{'f32c' if self.f32 else 'f64'} {'stream' if d.variant == STREAM else 'plain'} "
          f"xext={d.xext} xcd_remap={int(d.xcd_remap)} nt={int(d.nt_store)}"
          + (" nt_load=1" if d.nt_load else "") + (" order=offset" if d.order == "offset" else "")
          + (" remainder=1" if d.remainder else "") + (" pack=1" if d.pack else "")
          + (" (CPU emulation build)" if cpu else ""))
        if cpu:
            w('#include "cpu_emu.h"')
            w(f"typedef Lane<{T}> TV;")
        else:
            w("#include <hip/hip_runtime.h>")
            w("#include <stdint.h>")
            w("// sums are written with explicit fma; no other product may be fused into an add,")
            w("// whatever -ffp-contract the including build uses")
            w("#pragma clang fp contract(off)")
            w("typedef uint32_t u32;")
            w("typedef bool lbool;")
            w(f"typedef {T} TV;")
        w("")
        w("struct tcsg_kargs {")
        w(f"  const {T}* in;")
        w(f"  {T}* out;")
        w(f"  const {T}* mask;")
        w("  uint32_t xdim, ydim, zdim, right_offset;")
        w("  uint64_t in_bytes, out_bytes;  // buffer sizes")
        w("  uint32_t rec_limit, oob;       // STREAM: record count cap of a block's buffer window,")
        w("                                 // base offset of lanes outside the interior")
        w("};")
        blob = ", ".join(str(v) for v in d.params_blob())
        w(f'extern "C" __device__ const uint32_t tcsg_params[24] = {{{blob}}};')
        if not cpu:
            # the same block for host code that compiles this source in (a launch unit
            # that includes it, as the reference's launchUpdateGrid.cu includes kernels.cu)
            w(f"__attribute__((unused)) static const uint32_t tcsg_params_host[24] = {{{blob}}};")
        w("")
        if not cpu:
            w("template <class V> __device__ __forceinline__ V tcsg_sel(bool c, V a, V b) { return c ? a : b; }")
            w("__device__ __forceinline__ uint32_t tcsg_min(uint32_t a, uint32_t b) { return a < b ? a : b; }")
            w("template <class V> __device__ __forceinline__ V tcsg_uni(V v) { return v; }  // wave-uniform value")
            w("__device__ __forceinline__ double tcsg_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }")
            w("__device__ __forceinline__ float tcsg_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }")
        if self.f32:
            w("template <class F> __device__ __forceinline__ void tcsg_two_sum(F a, F b, F& s, F& e) {")
            w("  const F x = a + b;")
            w("  const F bp = x - a;")
            w("  const F ap = x - bp;")
            w("  e = (a - ap) + (b - bp);")
            w("  s = x;")
            w("}")
        if d.variant == STREAM and d.xext in ("dpp", "dppw") and not cpu:
            # tcsg_shr1/shl1: lane i receives lane i-1 / i+1 (0 where no such lane).
            # tcsg_shr1_or/shl1_or: the same, but lanes without a source lane keep `old`.
            for nm, ctrl in (("shr1", DPP_WAVE_SHR1), ("shl1", DPP_WAVE_SHL1)):
                if self.f32:
                    w(f"__device__ __forceinline__ float tcsg_{nm}(float v) {{")
                    w(f"  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp("
                      f"__builtin_bit_cast(int, v), {ctrl:#x}, 0xf, 0xf, true));")
                    w("}")
                    w(f"__device__ __forceinline__ float tcsg_{nm}_or(float old, float v) {{")
                    w(f"  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), "
                      f"__builtin_bit_cast(int, v), {ctrl:#x}, 0xf, 0xf, false));")
                    w("}")
                else:
                    w(f"__device__ __forceinline__ double tcsg_{nm}(double v) {{")
                    w("  const int lo = __double2loint(v), hi = __double2hiint(v);")
                    w(f"  return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, {ctrl:#x}, 0xf, 0xf, true),")
                    w(f"                          __builtin_amdgcn_mov_dpp(lo, {ctrl:#x}, 0xf, 0xf, true));")
                    w("}")
                    w(f"__device__ __forceinline__ double tcsg_{nm}_or(double old, double v) {{")
                    w("  const int lo = __double2loint(v), hi = __double2hiint(v);")
                    w("  const int olo = __double2loint(old), ohi = __double2hiint(old);")
                    w(f"  return __hiloint2double(__builtin_amdgcn_update_dpp(ohi, hi, {ctrl:#x}, 0xf, 0xf, false),")
                    w(f"                          __builtin_amdgcn_update_dpp(olo, lo, {ctrl:#x}, 0xf, 0xf, false));")
                    w("}")
        if d.edge_pairs:
            H = self.halo_lanes()
            for nm, ctrl in (("rowshl", DPP_ROW_SHL + H), ("rowshr", DPP_ROW_SHR + H)):
                if cpu:
                    w(f"inline TV tcsg_{nm}{H}(const TV& v) {{ return tcsg_emu_{nm}<{H}>(v); }}")
                elif self.f32:
                    w(f"__device__ __forceinline__ float tcsg_{nm}{H}(float v) {{")
                    w(f"  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp("
                      f"__builtin_bit_cast(int, v), {ctrl:#x}, 0xf, 0xf, true));")
                    w("}")
                else:
                    w(f"__device__ __forceinline__ double tcsg_{nm}{H}(double v) {{")
                    w("  const int lo = __double2loint(v), hi = __double2hiint(v);")
                    w(f"  return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, {ctrl:#x}, 0xf, 0xf, true),")
                    w(f"                          __builtin_amdgcn_mov_dpp(lo, {ctrl:#x}, 0xf, 0xf, true));")
                    w("}")
        if d.variant == STREAM and d.staged:
            bx, by, bz = d.stage
            n_lds = bz * 2 * (by * d.yc + 2 * d.range) * bx * d.xc
            if cpu:
                w(f"// the workgroup's LDS ({n_lds} elements); accesses are range- and race-checked")
                w(f"static tcsg_emu_lds<{T}> tcsg_emu_stage({n_lds});")
            else:
                w(f"__device__ __forceinline__ {T} tcsg_lds_ld(const {T}* p, uint32_t i) {{ return p[i]; }}")
                w(f"__device__ __forceinline__ void tcsg_lds_st({T}* p, uint32_t i, {T} v) {{ p[i] = v; }}")
        if d.variant == STREAM:
            w("// Buffer access: per-lane byte offset `off` (range-checked: lanes given an")
            w("// offset past the buffer read 0 and their stores are dropped) plus a")
            w("// wave-uniform non-negative row/plane offset `soff` in an SGPR.")
            if cpu:
                w("inline tcsg_emu_rsrc tcsg_rsrc(const void* p, uint32_t n) { return tcsg_make_rsrc(p, n); }")
                w(f"inline TV tcsg_ld(const tcsg_emu_rsrc& r, const u32& off, const u32& soff) {{ return tcsg_emu_ld<{T}>(r, off, soff); }}")
                w("inline void tcsg_st(const tcsg_emu_rsrc& r, const u32& off, const u32& soff, uint32_t imm, const TV& v) {")
                w("  tcsg_emu_st(r, off + imm, soff, v);")
                w("}")
                if d.remainder and d.xc > 1:
                    # a wide store is range-checked per dword on the hardware, as its elements are here
                    es_ = 4 if self.f32 else 8
                    for n in ((2, 4) if self.f32 else (2,)):
                        if n > d.xc:
                            continue
                        args = ", ".join(f"const TV& v{c}" for c in range(n))
                        body = " ".join(f"tcsg_emu_st(r, off + (imm + {c * es_}u), soff, v{c});" for c in range(n))
                        w(f"inline void tcsg_st{n}(const tcsg_emu_rsrc& r, const u32& off, const u32& soff, uint32_t imm, "
                          f"{args}) {{ {body} }}")
            else:
                bits = 32 if self.f32 else 64
                ity = "int" if self.f32 else "__attribute__((ext_vector_type(2))) int"
                w("__device__ __forceinline__ __amdgpu_buffer_rsrc_t tcsg_rsrc(const void* p, uint32_t n) {")
                w("  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)n, 0x00020000);")
                w("}")
                laux = 2 if d.nt_load else 0
                w(f"__device__ __forceinline__ {T} tcsg_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff) {{")
                w(f"  return __builtin_bit_cast({T}, __builtin_amdgcn_raw_buffer_load_b{bits}(r, off, soff, {laux}));")
                w("}")
                # nontemporal stores (gfx950 cache policy nt = 2), at device scope (sc1 = 16) where
                # every store instruction writes whole 128-byte lines once: Write alignment at a
                # multiple of 128 bytes (interior rows start on a line), at most 16 bytes per lane
                # and row (one instruction per row), and no remainder handling (whose shifted last
                # tiles write lines their neighbours also write).  The L2 then writes the lines
                # through instead of holding them dirty: 0.3-0.5% faster on the 512^3 Star r=4
                # kernels; where instructions write parts of lines, write-through costs 6-42%
                es_ = 4 if self.f32 else 8
                whole_lines = (d.alignment == ALIGN_WRITE and (d.cl_size * es_) % 128 == 0 and not d.remainder
                               and es_ * self.XC <= 16)
                aux = (18 if whole_lines else 2) if d.nt_store else 0
                # a row's columns merge into one store of es*XC bytes (explicit below with
                # remainder handling); only stores wider than 64 bits need the soffset-0 form
                wide = (4 if self.f32 else 8) * self.XC > 8
                if wide:
                    w("// Stores take the row/plane offset in the per-lane offset and an soffset of 0.  On")
                    w("// gfx950 a store of more than 64 bits (the compiler merges a row's columns into")
                    w("// one) whose data VGPRs the next VALU instruction overwrites can store that")
                    w("// instruction's values in some lanes, also with an SGPR soffset, which the ISA")
                    w("// documents and the compiler exempt from the wait-state rule; with soffset 0")
                    w("// the compiler inserts the wait states.  off + soff cannot wrap: the runtime")
                    w("// keeps the out-of-range base plus the largest row/plane offset below 2^32.")
                    w("// The column's constant `imm` is added last, so it becomes the instruction's")
                    w("// immediate offset, and one add per row serves all its columns (which the")
                    w("// compiler then merges into one wide store).")
                    vo, so = "(off + soff) + imm", "0"
                else:
                    w("// Stores of at most 64 bits: row/plane offset in the SGPR soffset.")
                    vo, so = "off + imm", "soff"
                w(f"__device__ __forceinline__ void tcsg_st(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, "
                  f"uint32_t imm, {T} v) {{")
                w(f"  __builtin_amdgcn_raw_buffer_store_b{bits}(__builtin_bit_cast({ity}, v), r, {vo}, {so}, {aux});")
                w("}")
                if d.remainder and d.xc > 1:
                    # one store instruction for consecutive columns of an output row
                    for n in ((2, 4) if self.f32 else (2,)):
                        if n * bits > 128 or n > d.xc:
                            continue
                        args = ", ".join(f"{T} v{c}" for c in range(n))
                        vals = ", ".join(f"v{c}" for c in range(n))
                        w(f"__device__ __forceinline__ void tcsg_st{n}(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, "
                          f"uint32_t imm, {args}) {{")
                        w(f"  const __attribute__((ext_vector_type({n}))) {T} v = {{{vals}}};")
                        w(f"  __builtin_amdgcn_raw_buffer_store_b{n * bits}("
                          f"__builtin_bit_cast(__attribute__((ext_vector_type({n * bits // 32}))) int, v), r, {vo}, {so}, {aux});")
                        w("}")
        w("")

    def emulation_runner(self):
        """Host loop that runs every wave of a launch through the emulated kernel."""
        w = self.w
        w("tcsg_emu_dim3u blockIdx, blockDim, gridDim;")
        w("thread_local tcsg_emu_dim3l threadIdx;")
        w("thread_local unsigned long tcsg_emu_phase = 0;")
        w('extern "C" { unsigned long tcsg_emu_bad_access = 0; }')
        w('extern "C" int tcsg_emulate(const tcsg_kargs* a, uint32_t gx, uint32_t gy, uint32_t gz,')
        w("                            uint32_t bx, uint32_t by, uint32_t bz) {")
        w("  const uint32_t nt = bx * by * bz;")
        w("  if (nt == 0 || nt % 64) return -1;  // whole waves only")
        w("  gridDim = {gx, gy, gz};")
        w("  blockDim = {bx, by, bz};")
        w("  for (uint32_t k = 0; k < gz; ++k) for (uint32_t j = 0; j < gy; ++j) for (uint32_t i = 0; i < gx; ++i) {")
        w("    blockIdx = {i, j, k};")
        if self.d.variant == STREAM and self.d.staged:
            w("    // the waves of the workgroup run as host threads that meet at each __syncthreads")
            w("    tcsg_emu_stage.reset();")
            w("    tcsg_emu_bar().start(nt / 64u);")
            w("    tcsg_emu_run_waves(nt / 64u, [=](unsigned wv) {")
            w("      for (int l = 0; l < 64; ++l) {")
            w("        const uint32_t lid = wv * 64u + l;")
            w("        threadIdx.x.v[l] = lid % bx; threadIdx.y.v[l] = (lid / bx) % by; threadIdx.z.v[l] = lid / (bx * by);")
            w("      }")
            w("      tcsg_emu_phase = 0;")
            w("      tcsg_update(*a);")
            w("    });")
            w("  }")
        else:
            w("    for (uint32_t w0 = 0; w0 < nt; w0 += 64) {")
            w("      for (int l = 0; l < 64; ++l) {")
            w("        const uint32_t lid = w0 + l;")
            w("        threadIdx.x.v[l] = lid % bx; threadIdx.y.v[l] = (lid / bx) % by; threadIdx.z.v[l] = lid / (bx * by);")
            w("      }")
            w("      tcsg_update(*a);")
            w("    }")
            w("  }")
        w("  return 0;")
        w("}")
        if self.d.batched:
            w("// One launch of the batched kernel: the workgroups take turns (round robin) to")
            w("// take a tile from the queue and run every wave on it; a workgroup that finds")
            w("// the queue drained finishes as the GPU kernel's does.")
            w('extern "C" int tcsg_emulate_batched(const void* grids, const void* tiles, uint32_t ntiles,')
            w("                                    uint32_t* queue, uint32_t nwg, uint32_t bx, uint32_t by, uint32_t bz) {")
            w("  const uint32_t nt = bx * by * bz;")
            w("  if (nt == 0 || nt % 64 || nwg == 0) return -1;")
            w("  gridDim = {nwg, 1, 1};")
            w("  blockDim = {bx, by, bz};")
            w("  tcsg_batch_args ba{(const tcsg_kargs*)grids, (const tcsg_tile*)tiles, ntiles, 0, queue};")
            w("  const uint32_t nr = nwg >= 8u ? 8u : 1u;")
            w("  std::vector<uint32_t> s(nwg, 0u);")
            w("  std::vector<char> live(nwg, 1);")
            w("  for (uint32_t nlive = nwg; nlive;) {")
            w("    for (uint32_t i = 0; i < nwg; ++i) {")
            w("      if (!live[i]) continue;")
            w("      const uint32_t t = tcsg_fetch(queue, ntiles, nr == 8u ? (i & 7u) : 0u, nr, s[i]);")
            w("      if (t == 0xFFFFFFFFu) { tcsg_finish(queue, nwg); live[i] = 0; --nlive; continue; }")
            w("      blockIdx = {i, 0, 0};")
            w("      for (uint32_t w0 = 0; w0 < nt; w0 += 64) {")
            w("        for (int l = 0; l < 64; ++l) {")
            w("          const uint32_t lid = w0 + l;")
            w("          threadIdx.x.v[l] = lid % bx; threadIdx.y.v[l] = (lid / bx) % by; threadIdx.z.v[l] = lid / (bx * by);")
            w("        }")
            w("        tcsg_update_batched_tile(ba, t);")
            w("      }")
            w("    }")
            w("  }")
            w("  return 0;")
            w("}")

    # ------------------------------------------------------------ body parts
    def block_coords(self):
        w = self.w
        if self.batched_body and self.d.pack:
            w("  // packed tile: run tl.bx of blockDim.x * blockDim.y consecutive threads of the grid's")
            w("  // xy thread space (x fastest, a run continues on the next thread row), thread planes")
            w("  // from tl.bz * blockDim.z on")
            w("  const u32 bzi = tl.bz;")
            w("  const u32 lin = tl.bx * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;")
            w("  // (a grid has packed tiles only if thx * thy > 0)")
            w("  const u32 gx = lin % thx, gy = lin / thx;")
            w("  const u32 gz = bzi * blockDim.z + threadIdx.z;")
            return
        if self.batched_body:
            w("  const u32 bxi = tl.bx, byi = tl.by, bzi = tl.bz;")
        elif self.d.xcd_remap:
            w("  // XCD-aware remap: hardware hands consecutive workgroups to the 8 XCDs")
            w("  // round-robin; give each XCD a contiguous range of tiles instead.")
            w("  const u32 nbx = gridDim.x, nby = gridDim.y, nbz = gridDim.z;")
            w("  const u32 nblk = nbx * nby * nbz;")
            w("  u32 lin = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);")
            w("  {")
            w("    const u32 q = nblk >> 3, rm = nblk & 7u, xcd = lin & 7u, k = lin >> 3;")
            w("    lin = tcsg_sel(xcd < rm, xcd * (q + 1u) + k, rm * (q + 1u) + (xcd - rm) * q + k);")
            w("  }")
            if self.d.zfirst:
                w("  const u32 bzi = lin % nbz, byi = (lin / nbz) % nby, bxi = lin / (nbz * nby);")
            else:
                w("  const u32 bxi = lin % nbx, byi = (lin / nbx) % nby, bzi = lin / (nbx * nby);")
        else:
            w("  const u32 bxi = blockIdx.x, byi = blockIdx.y, bzi = blockIdx.z;")
        w("  const u32 gx = bxi * blockDim.x + threadIdx.x;")
        w("  const u32 gy = byi * blockDim.y + threadIdx.y;")
        w("  const u32 gz = bzi * blockDim.z + threadIdx.z;")

    def thread_counts(self):
        d = self.d
        R, LO = self.R, d.left_offset
        w = self.w
        w("  const u32 xdim = a.xdim, ydim = a.ydim;")
        w(f"  const u32 ex = xdim - a.right_offset;  // end of the x extent in use")
        # with remainder handling the last thread of an axis may own fewer than
        # its coarsening's points (rounded-up counts); otherwise the runtime
        # only passes extents whose interiors are multiples of the coarsening
        up = lambda c: f" + {c - 1}u" if d.remainder and c > 1 else ""
        w(f"  const u32 thx = tcsg_sel(ex >= {2 * R + LO}u, (ex - {2 * R + LO}u{up(self.XC)}) / {self.XC}u, 0u);")
        w(f"  const u32 thy = tcsg_sel(ydim >= {2 * R}u, (ydim - {2 * R}u{up(self.YC)}) / {self.YC}u, 0u);")
        if self.three:
            w("  const u32 zdim = a.zdim;")
            w(f"  const u32 thz = tcsg_sel(zdim >= {2 * R}u, (zdim - {2 * R}u{up(self.ZC)}) / {self.ZC}u, 0u);")
        else:
            w("  const u32 thz = 1u;")

    # ---------------------------------------------------------------- PLAIN
    def body_plain(self):
        d = self.d
        T = self.T
        w = self.w
        R, LO = self.R, d.left_offset
        self.thread_counts()
        self.block_coords()
        # in the persistent batched kernel the body sits in a one-pass loop, so
        # `break` ends this tile where the plain kernel returns
        w(f"  if (!(gx < thx && gy < thy && gz < thz)) {'break' if self.batched_body else 'return'};")
        if d.remainder:
            # the last thread of an axis whose interior is not a multiple of the
            # coarsening starts early enough to end at the interior's end; outputs it
            # shares with its neighbour get the same bits from both (same inputs, same
            # order).  The runtime requires interiors of at least one coarsening.
            w(f"  const size_t x = (size_t)tcsg_min(gx * {self.XC}u, ex - {2 * R + LO + self.XC}u) + {R + LO};")
            w(f"  const size_t y = (size_t)tcsg_min(gy * {self.YC}u, ydim - {2 * R + self.YC}u) + {R};")
            w(f"  const size_t z = (size_t)tcsg_min(gz * {self.ZC}u, zdim - {2 * R + self.ZC}u) + {R};"
              if self.three else "  const size_t z = 0;")
        else:
            w(f"  const size_t x = (size_t)gx * {self.XC} + {R + LO};")
            w(f"  const size_t y = (size_t)gy * {self.YC} + {R};")
            w(f"  const size_t z = (size_t)gz * {self.ZC} + {R};" if self.three else "  const size_t z = 0;")
        w("  const size_t sx = xdim;")
        w("  const size_t sxy = (size_t)xdim * ydim;")
        if self.batched_body:
            # the grid's pointers come from the grid table in memory, where the compiler cannot
            # tell their address space: cast them to global so its loads and stores are global,
            # not flat, instructions
            T = f"tcsg_g{T}"
            w(f"  const {T}* __restrict__ src = (const {T}*)(a.in + z * sxy + y * sx + x);")
            w(f"  {T}* __restrict__ dst = ({T}*)(a.out + z * sxy + y * sx + x);")
        else:
            w(f"  const {T}* __restrict__ src = a.in + z * sxy + y * sx + x;")
            w(f"  {T}* __restrict__ dst = a.out + z * sxy + y * sx + x;")
        self.PT = T
        self.weights()
        if d.order == "offset":
            return self.body_plain_offset_order()
        pts, contrib = self.footprint()
        last = {}
        for idx, p in enumerate(pts):
            for (o, t, m) in contrib[p]:
                last[o] = idx
        rows = set()
        for idx, p in enumerate(pts):
            px, py, pz = p
            rk = (py, pz)
            if rk not in rows:
                rows.add(rk)
                w(f"  const {T}* __restrict__ q_{_n(py)}_{_n(pz)} = src + (ptrdiff_t)({py}) * (ptrdiff_t)sx"
                  f" + (ptrdiff_t)({pz}) * (ptrdiff_t)sxy;")
            v = f"v_{_n(px)}_{_n(py)}_{_n(pz)}"
            w(f"  const {self.T} {v} = q_{_n(py)}_{_n(pz)}[{px}];")
            for (o, t, m) in sorted(contrib[p], key=lambda c: c[0]):
                w("  " + self.acc_stmt(o, t, m, v))
            for o in self.outputs():
                if last[o] == idx:
                    i, j, k = o
                    w(f"  dst[(ptrdiff_t)({k}) * (ptrdiff_t)sxy + (ptrdiff_t)({j}) * (ptrdiff_t)sx + {i}]"
                      f" = {self.final_expr(o)};")

    def body_plain_offset_order(self):
        """PLAIN body, offset-major: for each stencil offset in summation
        order, the term of every output.  Each output still sums its terms in
        the reference's order; the weights are needed one after another (so a
        large mask streams through the scalar registers once), and a footprint
        point is loaded just before its first use and dies after its last."""
        T = self.PT
        w = self.w
        loaded, rows = set(), set()
        outs = self.outputs()
        for t, (dx, dy, dz) in enumerate(self.offs):
            m = spec.mask_index(self.d, dx, dy, dz)
            for o in outs:
                px, py, pz = o[0] + dx, o[1] + dy, o[2] + dz
                v = f"v_{_n(px)}_{_n(py)}_{_n(pz)}"
                if (py, pz) not in rows:
                    rows.add((py, pz))
                    w(f"  const {T}* __restrict__ q_{_n(py)}_{_n(pz)} = src + (ptrdiff_t)({py}) * (ptrdiff_t)sx"
                      f" + (ptrdiff_t)({pz}) * (ptrdiff_t)sxy;")
                if v not in loaded:
                    loaded.add(v)
                    w(f"  const {self.T} {v} = q_{_n(py)}_{_n(pz)}[{px}];")
                w("  " + self.acc_stmt(o, t, m, v))
        for o in outs:
            i, j, k = o
            w(f"  dst[(ptrdiff_t)({k}) * (ptrdiff_t)sxy + (ptrdiff_t)({j}) * (ptrdiff_t)sx + {i}]"
              f" = {self.final_expr(o)};")

    def weights(self):
        if self.d.coeff != VARIABLE:
            return
        T = self.T
        self.w(f"  const {T}* __restrict__ mk = a.mask;")
        for m in range(self.d.n):
            self.w(f"  const TV w{m} = mk[{m}];")

    # --------------------------------------------------------------- STREAM
    def halo_lanes(self) -> int:
        """H = ceil(r / xc): lanes a thread reaches for its x-neighbours."""
        return (self.R + self.XC - 1) // self.XC

    def body_stream(self):
        d = self.d
        T = self.T
        w = self.w
        R, LO = self.R, d.left_offset
        XC = self.XC
        es = 4 if self.f32 else 8
        dpp = d.xext in ("dpp", "dppw")
        wave_aligned = d.xext == "dppw"
        H = self.halo_lanes()
        ZR = R if self.three else 0
        XL = 0 if dpp else R  # columns left of x0 covered by the per-lane base
        self.thread_counts()
        self.block_coords()
        w("  // inrow: the thread's row/plane is inside the interior; such lanes load (and")
        w("  // can feed x-neighbours to other lanes) even past the interior's x end, where")
        w("  // they read the halo.  valid: the thread owns interior points and stores them.")
        # LDS staging with remainder handling shifts whole blocks (below), so every thread row
        # and plane of a launched block is inside the interior
        sremain = d.staged and d.remainder
        if sremain:
            w("  const lbool inrow = true;")
        else:
            w("  const lbool inrow = gy < thy && gz < thz;")
        w("  const lbool valid = inrow && gx < thx;")
        w(f"  const u32 x0 = gx * {XC}u + {R + LO}u;")
        if d.remainder and not sremain and (self.YC > 1 or (self.three and self.ZC > 1)):
            w("  // remainder: the last thread of an axis whose interior is not a multiple of the")
            w("  // coarsening starts early enough to end at the interior's end; the outputs it")
            w("  // shares with its neighbour get the same value from both (same inputs, same")
            w("  // summation order).  The runtime requires interiors of at least one coarsening")
            w("  // in y and z.")
        if sremain:
            BX_, BY_, BZ_ = d.stage
            TYs = BY_ * self.YC
            w("  // remainder handling with LDS staging: the block's tile (not one thread) starts early")
            w("  // enough to end at the interior's end; the rows it shares with the block before it get")
            w("  // the same bits from both.  The runtime requires an interior of at least one tile in y")
            w("  // and z.")
            w(f"  const u32 yb = tcsg_min(byi * {TYs}u, ydim - {2 * R + TYs}u);  // the block's first footprint row")
            w(f"  const u32 y0 = yb + threadIdx.y * {self.YC}u;            // first row of the footprint")
        elif d.remainder and self.YC > 1:
            w(f"  const u32 y0 = tcsg_min(gy * {self.YC}u, ydim - {2 * R + self.YC}u);  // first row of the footprint")
        else:
            w(f"  const u32 y0 = gy * {self.YC}u;            // first row of the footprint")
        if not self.three:
            w("  const u32 z0 = 0u;")
        elif sremain:
            w(f"  const u32 wb = tcsg_min(bzi * {BZ_ * self.ZC}u, zdim - {2 * R + BZ_ * self.ZC}u);"
              f"  // the block's first footprint plane")
            w(f"  const u32 z0 = wb + threadIdx.z * {self.ZC}u;            // first plane of the footprint")
        elif d.remainder and self.ZC > 1:
            w(f"  const u32 z0 = tcsg_min(gz * {self.ZC}u, zdim - {2 * R + self.ZC}u);  // first plane of the footprint")
        else:
            w(f"  const u32 z0 = gz * {self.ZC}u;            // first plane of the footprint")
        w(f"  const u32 sx = xdim * {es}u;            // row pitch in bytes")
        w(f"  const u32 sxy = xdim * ydim * {es}u;     // plane pitch in bytes")
        # the buffer resources start at the block's first footprint plane (3D) or row (2D), so
        # 32-bit offsets reach any buffer size: the runtime only checks that one block's window
        # (its planes or rows plus the footprint) stays below the out-of-range base
        if self.three:
            zcl = f", zdim - {2 * R + self.ZC}u" if d.remainder and self.ZC > 1 else ""
            if not sremain:
                w(f"  const u32 wb = {'tcsg_min(' if zcl else '('}bzi * blockDim.z * {self.ZC}u{zcl});"
                  f"  // the block's first footprint plane")
            w("  const uint64_t wbyte = (uint64_t)tcsg_uni(wb) * tcsg_uni(sxy);")
            w(f"  const u32 real = ((z0 - wb) * ydim + y0) * xdim * {es}u + (x0 - {XL}u) * {es}u;")
        else:
            ycl = f", ydim - {2 * R + self.YC}u" if d.remainder and self.YC > 1 else ""
            w(f"  const u32 wb = {'tcsg_min(' if ycl else '('}byi * blockDim.y * {self.YC}u{ycl});"
              f"  // the block's first footprint row")
            w("  const uint64_t wbyte = (uint64_t)tcsg_uni(wb) * tcsg_uni(sx);")
            w(f"  const u32 real = (y0 - wb) * xdim * {es}u + (x0 - {XL}u) * {es}u;")
        w("  // lanes outside the interior take base offset `oob`, past the window's record count:")
        w("  // their loads read 0 and their stores are dropped by the range check")
        w("  const u32 oob = a.oob;")
        if dpp:
            w("  // Lanes past the interior's x end still load (valid lanes take halo columns from")
            w("  // them through DPP).  Buffer range checks are per dword on the full offset")
            w("  // (per-lane + SGPR + immediate), so anything such a lane reads past the buffer's")
            w("  // end returns 0 and the in-range part of a straddling access is exact.")
            w(f"  const u32 base = tcsg_sel(inrow, real, oob);")
        else:
            w(f"  const u32 base = tcsg_sel(valid, real, oob);")
        w(f"  const u32 sbase = tcsg_sel(valid, real, oob);")
        # x positions stay lane-contiguous for the neighbour exchange, so with
        # remainder handling the last thread in x keeps its place and drops the
        # columns past the interior's end instead: every thread stores an output
        # row's XC columns with wide stores at sfull (out of range on a thread
        # with fewer columns left), and such a thread stores the columns it has
        # one by one at spart<i> (out of range on every other thread)
        xgroups = d.remainder and XC > 1
        if xgroups:
            w(f"  const u32 nvx = ex - {R}u - x0;  // interior columns from x0 on")
            w(f"  const u32 sfull = tcsg_sel(nvx >= {XC}u, sbase + {XL * es}u, oob);")
            for i in range(XC - 1):
                w(f"  const u32 spart{i} = tcsg_sel(nvx > {i}u && nvx < {XC}u, sbase + {(i + XL) * es}u, "
                  f"oob);")
        w("  const uint64_t nin = a.in_bytes - wbyte, nout = a.out_bytes - wbyte;")
        w("  const auto rin = tcsg_rsrc((const char*)a.in + wbyte, (uint32_t)(nin < a.rec_limit ? nin : a.rec_limit));")
        wide_st = es * XC > 8
        if wide_st:
            # stores wider than 64 bits take soffset 0 (see tcsg_st): one resource per output
            # row (its row offset in the base) and the plane offset added to the per-lane
            # offset once per plane, shared by the plane's row stores
            w("  const uint32_t recout = (uint32_t)(nout < a.rec_limit ? nout : a.rec_limit);")
            for j in range(self.YC):
                w(f"  const auto rout{j} = tcsg_rsrc((const char*)a.out + (wbyte + {j + R}u * (uint64_t)tcsg_uni(sx)), "
                  f"recout - {j + R}u * tcsg_uni(sx));")
        else:
            w("  const auto rout = tcsg_rsrc((const char*)a.out + wbyte, (uint32_t)(nout < a.rec_limit ? nout : a.rec_limit));")
        self.weights()
        if dpp:
            w("  const u32 lid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);")
            w("  const u32 lane = lid & 63u;")
            if wave_aligned:
                w("  // Block rows are whole waves (the runtime only launches block x extents")
                w(f"  // that are multiples of 64): the {H} edge lanes of each wave load the")
                w("  // columns beyond the wave, which enter the lane-shift chains as the DPP")
                w("  // `old` value of lanes that have no source lane.")
                w(f"  const lbool edgeL = lane < {H}u, edgeR = lane + {H}u > 63u;")
            else:
                w("  // Lane segments: runs of lanes in one block row and one wave.  tL/tR are the")
                w("  // lane's distances to its segment's ends; the runtime only launches block")
                w(f"  // shapes whose segments are at least H = {H} lanes long.")
                w("  const u32 tx = threadIdx.x;")
                w("  const u32 tL = tcsg_min(tx, lane);")
                w("  const u32 tR = tcsg_min(blockDim.x - 1u - tx, 63u - lane);")
                w(f"  const lbool edgeL = tL < {H}u, edgeR = tR < {H}u;")
        # the left edge lanes' extra columns start H*XC columns left of x0; x0 >= R + LO,
        # so with R + LO >= H*XC one offset (plus immediates) serves all of them and the
        # compiler merges their loads into one wide load per row
        merged_left = R + LO >= H * XC
        if dpp:
            w("  // Extra columns of the segment-edge lanes.  Left ones are addressed only where")
            w("  // the column index stays non-negative, so no per-lane offset wraps below zero.")
            if merged_left:
                w(f"  const u32 eoffL = tcsg_sel(edgeL, base - {H * XC * es}u, oob);")
            else:
                for ip in range(XC):
                    k = H * XC - ip
                    w(f"  const u32 eoffL{ip} = tcsg_sel(edgeL && x0 >= {k}u, base - {k * es}u, oob);")
            w(f"  const u32 eoffR = tcsg_sel(edgeR, base + {H * XC * es}u, oob);")
        # wave-aligned rows: the left edge lanes (lane < H) and the right ones (lane > 63 - H)
        # never coincide, so one load, with each side's offset in its own lanes, fetches both
        # sides' extra columns into one register that both lane-shift chains read
        both_edges = dpp and wave_aligned and merged_left and 2 * H <= 64
        if both_edges:
            w("  const u32 eoffE = tcsg_sel(edgeL, eoffL, eoffR);")
        staged = d.staged
        if staged:
            BX, BY, BZ = d.stage
            TX, TY = BX * XC, BY * self.YC
            NROW = TY + 2 * R
            NR = -(-NROW // BY)
            w(f"  // LDS halo staging (blocks of {BX}x{BY}x{BZ} threads): each z-slice of the block loads")
            w(f"  // every plane's {NROW} rows ({TY} rows and {R} halo rows per side) by {TX} columns once")
            w(f"  // into one of two LDS buffers, thread row ty the tile rows ty + {BY} i, and each thread")
            w("  // reads its footprint rows there.  Plane p goes to buffer p % 2: its rows are loaded")
            w("  // two planes ahead, written to LDS after the plane before it is done, and read after")
            w("  // the barrier that follows; the barrier after plane p's reads frees its buffer.")
            if not self.target == "cpu":
                w(f"  __shared__ TV tcsg_stage[{BZ * 2 * NROW * TX}];")
            else:
                w("  auto& tcsg_stage = tcsg_emu_stage;")
            if not sremain:
                w(f"  const u32 yb = byi * {TY}u;  // the block's first footprint row")
            w(f"  const u32 sread = (threadIdx.z * {2 * NROW}u + threadIdx.y * {self.YC}u) * {TX}u + threadIdx.x * {XC}u;")
            w(f"  const u32 swrite = (threadIdx.z * {2 * NROW}u + threadIdx.y) * {TX}u + threadIdx.x * {XC}u;")
            w("  // any row and column of the tile can feed a valid output: every lane of a valid z chunk loads")
            w(f"  const u32 stbase = tcsg_sel({'inrow' if sremain else 'gz < thz'}, ((z0 - wb) * ydim + yb + "
              f"threadIdx.y) * xdim * {es}u + (x0 - {XL}u) * {es}u, oob);")
        pts, contrib = self.footprint()
        # edge_pairs: two adjacent rows of one plane that both need edge columns share a load;
        # the second row's columns go to the next H lanes on each side (one row below in the
        # per-lane offset) and a DPP row shift moves them into the edge lanes
        pair_role = {}
        pair_edges = d.edge_pairs and both_edges and 2 * H <= 16
        if pair_edges:
            need = sorted({(pz, py) for (px, py, pz) in pts if not 0 <= px < XC})
            rows = defaultdict(list)
            for pz, py in need:
                rows[pz].append(py)
            for pz, pys in rows.items():
                i = 0
                while i + 1 < len(pys):
                    if pys[i + 1] == pys[i] + 1:
                        pair_role[(pys[i], pz)] = ("a", pys[i])
                        pair_role[(pys[i + 1], pz)] = ("b", pys[i])
                        i += 2
                    else:
                        i += 1
            w(f"  const lbool pairL = lane >= {H}u && lane < {2 * H}u;")
            w(f"  const lbool pairR = lane + {2 * H}u > 63u && lane + {H}u <= 63u;")
            w(f"  const u32 eoffP = tcsg_sel(pairL, base - {2 * H * XC * es}u + sx, "
              f"tcsg_sel(pairR, base + {2 * H * XC * es}u + sx, eoffE));")
        last = {}
        for idx, p in enumerate(pts):
            for (o, t, m) in contrib[p]:
                last[o] = idx

        def soff(py, pz):
            """Uniform row/plane byte offset of thread-relative row (py, pz)."""
            return f"({py + R}u * sx + {pz + ZR}u * sxy)"

        def st_rs(py, pz):
            """(buffer resource, uniform offset) of a store to output row (py, pz)."""
            if wide_st:
                return f"rout{py}", f"{pz + ZR}u * sxy"
            return "rout", soff(py, pz)

        def vname(px, py, pz):
            return f"v_{_n(px)}_{_n(py)}_{_n(pz)}"

        def ename(side, j, ip, py, pz):
            """Edge-column value j lane hops along the chain of `side` (j = 0: as loaded,
            or for the second row of a pair, as moved into the edge lanes)."""
            if j == 0 and (py, pz) in pair_role:
                role, pya = pair_role[(py, pz)]
                return f"eP_{ip}_{_n(pya)}_{_n(pz)}" if role == "a" else f"eP{side}_{ip}_{_n(pya)}_{_n(pz)}"
            if j == 0 and both_edges:
                return f"eE0_{ip}_{_n(py)}_{_n(pz)}"
            return f"e{side}{j}_{ip}_{_n(py)}_{_n(pz)}"

        def hop(px):
            """(side, h, ip): lane hop and source column of an x-neighbour."""
            if px < 0:
                h = (-px + XC - 1) // XC
                return "L", h, h * XC + px
            h = px // XC
            return "R", h, px - h * XC

        # streaming groups: planes in 3D, rows in 2D Row, the whole tile in 2D Column
        if self.three:
            skey = lambda p: p[2]
        elif d.orientation == ROW:
            skey = lambda p: p[1]
        else:
            skey = lambda p: 0
        order = []
        members = defaultdict(list)
        for idx, p in enumerate(pts):
            k = skey(p)
            if k not in members:
                order.append(k)
            members[k].append((idx, p))

        def loads_of(k):
            """(name, offset expr, soff expr) of the loads plane/row k needs."""
            out = []
            seen = set()
            for _, (px, py, pz) in members[k]:
                if 0 <= px < XC or not dpp:
                    srcs = [(px, py, pz)]
                else:
                    side, h, ip = hop(px)
                    srcs = [(ip, py, pz), ("E" + side, ip, py, pz)]
                for sp in srcs:
                    if sp in seen:
                        continue
                    seen.add(sp)
                    if sp[0] in ("EL", "ER"):
                        _, ip, py2, pz2 = sp
                        side = sp[0][1]
                        if (py2, pz2) in pair_role:
                            pya = pair_role[(py2, pz2)][1]
                            out.append((f"eP_{ip}_{_n(pya)}_{_n(pz2)}", f"eoffP + {ip * es}u", soff(pya, pz2)))
                            continue
                        if both_edges:
                            off = f"eoffE + {ip * es}u"
                        elif side == "R":
                            off = f"eoffR + {ip * es}u"
                        else:
                            off = f"eoffL + {ip * es}u" if merged_left else f"eoffL{ip}"
                        out.append((ename(side, 0, ip, py2, pz2), off, soff(py2, pz2)))
                    else:
                        qx, qy, qz = sp
                        vcoord[vname(qx, qy, qz)] = sp
                        out.append((vname(qx, qy, qz), f"base + {(qx + XL) * es}u", soff(qy, qz)))
            return out

        vcoord = {}

        declared = set()

        def emit_loads(k, buf=0):
            for nm, off, so in loads_of(k):
                if nm in declared:
                    continue
                declared.add(nm)
                if staged and nm in vcoord:
                    qx, qy, _ = vcoord[nm]
                    w(f"  const TV {nm} = tcsg_lds_ld(tcsg_stage, sread + {(buf * NROW + qy + R) * TX + qx}u);")
                    continue
                w(f"  const TV {nm} = tcsg_ld(rin, {off}, {so});")

        def stage_loads(gi):
            for i in range(NR):
                for c in range(XC):
                    w(f"  const TV sg{gi}_{i}_{c} = tcsg_ld(rin, stbase + {c * es}u, "
                      f"({i * BY}u * sx + {order[gi] + ZR}u * sxy));")

        def stage_writes(gi):
            for i in range(NR):
                for c in range(XC):
                    st = f"tcsg_lds_st(tcsg_stage, swrite + {((gi % 2) * NROW + i * BY) * TX + c}u, sg{gi}_{i}_{c});"
                    # (thread rows are whole waves: the condition is uniform across each wave)
                    w(f"  if (tcsg_uni(threadIdx.y) + {i * BY}u < {NROW}u) {st}" if (i + 1) * BY > NROW else f"  {st}")

        def dpp_value(p):
            """Emit the lane-shift chain that brings x-neighbour p into this lane (main chain
            from the own columns, halo chain from the edge lanes' extra loads)."""
            px, py, pz = p
            nm = vname(px, py, pz)
            if 0 <= px < XC or nm in declared:
                return
            side, h, ip = hop(px)
            main_fn, e_fn = ("tcsg_shr1", "tcsg_shl1") if side == "L" else ("tcsg_shl1", "tcsg_shr1")
            gname = lambda j: ename(side, j, ip, py, pz)
            if (py, pz) in pair_role and pair_role[(py, pz)][0] == "b" and gname(0) not in declared:
                pya = pair_role[(py, pz)][1]
                declared.add(gname(0))
                w(f"  const TV {gname(0)} = tcsg_row{'shl' if side == 'L' else 'shr'}{H}("
                  f"eP_{ip}_{_n(pya)}_{_n(pz)});")
            for j in range(1, H):
                if gname(j) not in declared:
                    declared.add(gname(j))
                    w(f"  const TV {gname(j)} = {e_fn}({gname(j - 1)});")
            if wave_aligned:
                prev = vname(ip, py, pz)
                for hh in range(1, h + 1):
                    sn = f"m{side}{hh}_{ip}_{_n(py)}_{_n(pz)}"
                    if sn not in declared:
                        declared.add(sn)
                        w(f"  const TV {sn} = {main_fn}_or({gname(H - hh)}, {prev});")
                    prev = sn
                w(f"  const TV {nm} = {prev};")
            else:
                prev = vname(ip, py, pz)
                for hh in range(1, h + 1):
                    sn = f"s{side}{hh}_{ip}_{_n(py)}_{_n(pz)}"
                    if sn not in declared:
                        declared.add(sn)
                        w(f"  const TV {sn} = {main_fn}({prev});")
                    prev = sn
                w(f"  const TV {nm} = tcsg_sel(t{side} >= {h}u, {prev}, {gname(H - h)});")
            declared.add(nm)

        def store_output(o, final_of_row):
            """Store output o; with remainder column groups, the row's wide stores once its last
            column is final (final_of_row)."""
            i, j, kk = o
            if not xgroups:
                rs, so = st_rs(j, kk)
                w(f"  tcsg_st({rs}, sbase, {so}, {(i + XL) * es}u, {self.final_expr(o)});")
                return
            w(f"  const TV f_{i}_{j}_{kk} = {self.final_expr(o)};")
            if final_of_row:
                row = [f"f_{c}_{j}_{kk}" for c in range(XC)]
                rs, so = st_rs(j, kk)
                # wide stores of 2 or 4 elements, at most 16 bytes each
                c0 = 0
                while c0 < XC:
                    n = max(m for m in (1, 2, 4) if m * es <= 16 and c0 + m <= XC)
                    part = row[c0:c0 + n]
                    if n == 1:
                        w(f"  tcsg_st({rs}, sfull, {so}, {c0 * es}u, {part[0]});")
                    else:
                        w(f"  tcsg_st{n}({rs}, sfull, {so}, {c0 * es}u, {', '.join(part)});")
                    c0 += n
                for c in range(XC - 1):
                    w(f"  tcsg_st({rs}, spart{c}, {so}, 0u, {row[c]});")

        if d.order == "offset":
            # offset-major statements (2D): for each stencil offset in summation order, the term
            # of every output, so each weight is used by all outputs at once and streams through
            # the scalar registers; each output's own sum keeps the reference's order.  A row's
            # loads are issued where the row is first needed, its x-neighbours lane-shifted
            # where first needed
            touched = set()
            for t, (dx, dy, dz) in enumerate(self.offs):
                m = spec.mask_index(d, dx, dy, dz)
                for o in self.outputs():
                    p = (o[0] + dx, o[1] + dy, o[2] + dz)
                    if skey(p) not in touched:
                        touched.add(skey(p))
                        emit_loads(skey(p))
                    if dpp:
                        dpp_value(p)
                    w("  " + self.acc_stmt(o, t, m, vname(*p)))
            for o in self.outputs():
                store_output(o, o[0] == XC - 1)
            return

        D = d.prefetch
        if staged:
            assert order == sorted(order) and D == 0
            stage_loads(0)
            stage_writes(0)
            w("  __syncthreads();")
            if len(order) > 1:
                stage_loads(1)
        for i in range(min(D, len(order))):
            emit_loads(order[i])
        for gi, k in enumerate(order):
            if gi + D < len(order):
                emit_loads(order[gi + D], (gi + D) % 2)
            emit_loads(k, gi % 2)
            gpts = [p for _, p in members[k]]
            if dpp:
                for p in gpts:
                    dpp_value(p)
            for idx, p in members[k]:
                v = vname(*p)
                for (o, t, m) in sorted(contrib[p], key=lambda c: c[0]):
                    w("  " + self.acc_stmt(o, t, m, v))
                for o in self.outputs():
                    if last[o] != idx:
                        continue
                    i, j, kk = o
                    store_output(o, not xgroups or idx == max(last[(c, j, kk)] for c in range(XC)))
            if staged and gi + 1 < len(order):
                stage_writes(gi + 1)
                w("  __syncthreads();")
                if gi + 2 < len(order):
                    stage_loads(gi + 2)

    # ------------------------------------------------------------------ all
    # ------------------------------------------------------------- TEMPORAL
    def _temporal_head(self, title: str, lds_elems: int):
        """Includes, helpers, the argument struct, the parameter block and the LDS
        access macro of a temporal kernel (in the CPU build every LDS access is
        range-checked against the `lds_elems` elements of the workgroup's LDS)."""
        d = self.d
        T = self.T
        w = self.w
        cpu = self.target == "cpu"
        w("// Generated by threadcoarsenedstencilgenerator_amd.generator -- do not edit.")
        w(f"// {title}" + (" (CPU emulation build)" if cpu else ""))
        fn = "static inline" if cpu else "__device__ __forceinline__"
        if cpu:
            w("#include <cmath>")
            w("#include <cstddef>")
            w("#include <cstdint>")
            w("#include <vector>")
            w("typedef uint32_t u32;")
            w(f"typedef {T} TV;")
            w("static inline double tcsg_fma(double a, double b, double c) { return std::fma(a, b, c); }")
            w("static inline float tcsg_fma(float a, float b, float c) { return std::fmaf(a, b, c); }")
        else:
            w("#include <hip/hip_runtime.h>")
            w("#include <stdint.h>")
            w("// sums are written with explicit fma; no other product may be fused into an add,")
            w("// whatever -ffp-contract the including build uses")
            w("#pragma clang fp contract(off)")
            w("typedef uint32_t u32;")
            w(f"typedef {T} TV;")
            w("__device__ __forceinline__ double tcsg_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }")
            w("__device__ __forceinline__ float tcsg_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }")
        if self.f32:
            w(f"template <class F> {fn} void tcsg_two_sum(F a, F b, F& s, F& e) {{")
            w("  const F x = a + b;")
            w("  const F bp = x - a;")
            w("  const F ap = x - bp;")
            w("  e = (a - ap) + (b - bp);")
            w("  s = x;")
            w("}")
        w("")
        w("struct tcsg_kargs {")
        w(f"  const {T}* in;")
        w(f"  {T}* out;")
        w(f"  const {T}* mask;")
        w("  uint32_t xdim, ydim, zdim, right_offset;")
        w("  uint64_t in_bytes, out_bytes;  // buffer sizes")
        w("  uint32_t rec_limit, oob;       // STREAM: record count cap of a block's buffer window,")
        w("                                 // base offset of lanes outside the interior")
        w("};")
        blob = ", ".join(str(v) for v in d.params_blob())
        w(f'extern "C" {"" if cpu else "__device__ "}const uint32_t tcsg_params[24] = {{{blob}}};')
        if not cpu:
            w(f"__attribute__((unused)) static const uint32_t tcsg_params_host[24] = {{{blob}}};")
        w("")
        if cpu:
            w("// LDS accesses are range-checked here: one outside the workgroup's LDS is counted")
            w("// (and reads a dummy) instead of touching other memory")
            w('extern "C" { unsigned long tcsg_emu_bad_access = 0; }')
            w("static TV* tcsg_lds_lo = nullptr;")
            w("static TV tcsg_lds_dummy;")
            w(f"static inline TV& tcsg_lds(TV* p) {{")
            w(f"  if (p < tcsg_lds_lo || p >= tcsg_lds_lo + {lds_elems}) {{ ++tcsg_emu_bad_access; return tcsg_lds_dummy; }}")
            w("  return *p;")
            w("}")
            w("#define TCSG_LDS(p, k) tcsg_lds(const_cast<TV*>(p) + (k))")
        else:
            w("#define TCSG_LDS(p, k) (p)[k]")

    def source_temporal3d(self) -> str:
        """3D kernel that advances tsteps = K time steps per launch by marching
        through z.  Workgroup (64, BY) owns the outputs of an x-y tile of TX = 64 xc
        by TY = BY yc points over ZC = zc consecutive planes.  Level 0 holds input
        planes of the tile plus H = K r points per side, level s (1..K-1) the step-s
        values of planes over the tile plus (K - s) r points per side; each level is a
        ring of 2r+1 planes in LDS.  Iteration t loads input plane t (the next one is
        fetched into registers while this one is used), then each level s computes
        its plane t - s r from the 2r+1 planes of level s-1 around it; level K writes
        the output plane.  Points outside the interior (the Dirichlet halo) keep
        their values; points past the allocation read 0 and feed no interior point.
        Every interior point sums its terms in the reference's order with the same
        operations as a one-step kernel, so k launches equal k*K steps bit for bit."""
        d = self.d
        T = self.T
        w = self.w
        cpu = self.target == "cpu"
        R, LO, K = self.R, d.left_offset, d.tsteps
        H = K * R
        BY = d.max_threads // 64
        NT = 64 * BY
        TX, TY, ZC = 64 * d.xc, BY * d.yc, d.zc
        D = 2 * R + 1
        NQ = ZC + 2 * H                                   # input planes per workgroup
        dims = [(TX + 2 * (K - s) * R, TY + 2 * (K - s) * R) for s in range(K)]   # level s region
        offs, o = [], 0
        for (ws, hs) in dims:
            offs.append(o)
            o += D * ws * hs
        total = o
        nbname = "Box" if d.neighbourhood == BOX else "Star"
        self._temporal_head(f"3D {nbname} r={R} {'Variable' if d.coeff == VARIABLE else 'Jacobi'} "
                            f"{'Row' if d.orientation == ROW else 'Column'} alignment "
                            f"{['None', 'Read', 'Write'][d.alignment]}/{d.cl_size} {'f32c' if self.f32 else 'f64'} "
                            f"temporal: {K} steps per launch, tile {TX}x{TY}x{ZC}, block 64x{BY}, "
                            f"LDS {total * (4 if self.f32 else 8)} bytes", total)
        fn = "static inline" if cpu else "__device__ __forceinline__"
        W0, H0 = dims[0]
        NLD = (W0 * H0 + NT - 1) // NT
        w("// interior of the padded grid")
        w(f"{fn} bool tcsg_interior(const tcsg_kargs& a, long long gx, long long gy, long long gz) {{")
        w(f"  return gx >= {R + LO} && gx < (long long)a.xdim - (long long)a.right_offset - {R} &&")
        w(f"         gy >= {R} && gy < (long long)a.ydim - {R} && gz >= {R} && gz < (long long)a.zdim - {R};")
        w("}")
        w("// global x, y of level 0's first column and row, and z of input plane 0")
        w(f"{fn} long long tcsg_x0(u32 bx) {{ return (long long)bx * {TX} + {R + LO - H}; }}")
        w(f"{fn} long long tcsg_y0(u32 by) {{ return (long long)by * {TY} + {R - H}; }}")
        w(f"{fn} long long tcsg_z0(u32 bz) {{ return (long long)bz * {ZC} + {R - H}; }}")
        w("")
        w(f"// the {d.n if d.coeff == VARIABLE else 0} weights, read once per workgroup before any plane is fetched")
        w(f"struct tcsg_w {{ TV w[{d.n if d.coeff == VARIABLE else 1}]; }};")
        w(f"{fn} void tcsg_weights(const tcsg_kargs& a, tcsg_w& W) {{")
        if d.coeff == VARIABLE:
            for mi in range(d.n):
                w(f"  W.w[{mi}] = a.mask[{mi}];")
        else:
            w("  W.w[0] = 0;")
        w("}")
        w(f"// one input plane of level 0 ({W0}x{H0} points), {NLD} per thread, in registers")
        w(f"struct tcsg_pf {{ TV v[{NLD}]; }};")
        w(f"{fn} void tcsg_fetch(const tcsg_kargs& a, u32 bx, u32 by, u32 bz, u32 tid, u32 t, tcsg_pf& pf) {{")
        w("  const long long x0 = tcsg_x0(bx), y0 = tcsg_y0(by), gz = tcsg_z0(bz) + (long long)t;")
        w("  const bool zok = gz >= 0 && gz < (long long)a.zdim;")
        for k in range(NLD):
            w(f"  {{ const u32 p = tid + {k * NT}u;")
            w(f"    const long long gx = x0 + (long long)(p % {W0}u), gy = y0 + (long long)(p / {W0}u);")
            cond = f"p < {W0 * H0}u && " if (k + 1) * NT > W0 * H0 else ""
            w(f"    const bool ok = {cond}zok && gx >= 0 && gx < (long long)a.xdim && gy >= 0 && gy < (long long)a.ydim;")
            w(f"    pf.v[{k}] = ok ? a.in[((size_t)gz * a.ydim + (size_t)gy) * a.xdim + (size_t)gx] : ({T})0; }}")
        w("}")
        w(f"{fn} void tcsg_put(TV* lds, u32 tid, u32 t, const tcsg_pf& pf) {{")
        w(f"  TV* pl = lds + {offs[0]}u + (t % {D}u) * {W0 * H0}u;")
        for k in range(NLD):
            st = f"TCSG_LDS(pl, tid + {k * NT}u) = pf.v[{k}];"
            w(f"  if (tid + {k * NT}u < {W0 * H0}u) {st}" if (k + 1) * NT > W0 * H0 else f"  {st}")
        w("}")
        w("")
        box = d.neighbourhood == BOX
        for s_ in range(1, K + 1):
            Wp, Hp = dims[s_ - 1]
            last = s_ == K
            Ws, Hs = (TX, TY) if last else dims[s_]
            w(f"// level {s_}: plane q = t - {s_ * R} over {Ws}x{Hs} points from level {s_ - 1}"
              + (" into the output grid" if last else ""))
            w(f"{fn} void tcsg_level{s_}(const tcsg_kargs& a, const tcsg_w& W, TV* lds, u32 bx, u32 by, u32 bz, "
              f"u32 tid, u32 t) {{")
            w(f"  if (t < {2 * s_ * R}u || t >= {NQ}u) return;   // planes q in [{s_ * R}, {NQ - s_ * R})")
            w(f"  const u32 q = t - {s_ * R}u;")
            w(f"  const long long gx0 = tcsg_x0(bx) + {s_ * R}, gy0 = tcsg_y0(by) + {s_ * R}, "
              f"gz = tcsg_z0(bz) + (long long)q;")
            for dz in range(-R, R + 1):
                w(f"  const TV* src{_n(dz)} = lds + {offs[s_ - 1]}u + ((q + {D + dz}u) % {D}u) * {Wp * Hp}u;")
            if not last:
                w(f"  TV* dst = lds + {offs[s_]}u + (q % {D}u) * {Ws * Hs}u;")
            if d.coeff == VARIABLE:
                for mi in range(d.n):
                    w(f"  const TV w{mi} = W.w[{mi}];")
            npt = (Ws * Hs + NT - 1) // NT

            def points(checked):
                # the thread's points p = tid + k*NT, k < npt, unrolled so that every point's
                # LDS reads can be in flight together
                for k in range(npt):
                    guard = (k + 1) * NT > Ws * Hs
                    w(f"    {{ const u32 p = tid + {k * NT}u;")
                    if guard:
                        w(f"      if (p < {Ws * Hs}u) {{")
                    ind = "        " if guard else "      "
                    w(f"{ind}const u32 i = p % {Ws}u, j = p / {Ws}u;")
                    w(f"{ind}const u32 c = (j + {R}u) * {Wp}u + i + {R}u;   // the point in level {s_ - 1}")
                    if checked:
                        w(f"{ind}const bool in = tcsg_interior(a, gx0 + (long long)i, gy0 + (long long)j, gz);")
                    w(f"{ind}TV res = TCSG_LDS(src0, c);")
                    w(f"{ind}if (in) {{" if checked else f"{ind}{{")
                    for t, (dx, dy, dz) in enumerate(self.offs):
                        mi = spec.mask_index(d, dx, dy, dz)
                        v = f"TCSG_LDS(src{_n(dz)}, c + ({dy * Wp + dx}))"
                        w(f"{ind}  " + self.acc_stmt((0, 0, 0), t, mi, v))
                    w(f"{ind}  res = {self.final_expr((0, 0, 0))};")
                    w(f"{ind}}}")
                    if last:
                        st = "a.out[((size_t)gz * a.ydim + (size_t)(gy0 + j)) * a.xdim + (size_t)(gx0 + i)] = res;"
                        w(f"{ind}if (in) {st}" if checked else f"{ind}{st}")
                    else:
                        w(f"{ind}TCSG_LDS(dst, p) = res;")
                    if guard:
                        w("      }")
                    w("    }")
            # a plane of the level wholly inside the interior (most of them) needs no per-point test
            w(f"  const bool inner = gz >= {R} && gz < (long long)a.zdim - {R} && gx0 >= {R + LO} && "
              f"gx0 + {Ws} <= (long long)a.xdim - (long long)a.right_offset - {R} &&")
            w(f"                     gy0 >= {R} && gy0 + {Hs} <= (long long)a.ydim - {R};")
            w("  if (inner) {")
            points(False)
            w("  } else {")
            points(True)
            w("  }")
            w("}")
            w("")
        # barriers: after the plane's put, and after each level but the last (whose reads of
        # level K-1 end before the next iteration's put and its barrier); with one step per
        # launch the level also reads level 0, which the next put overwrites
        syncs = [True] * (K - 1) + [K == 1]
        if cpu:
            w("// One launch: the workgroups one after another; within a workgroup each phase runs for")
            w("// every thread before the next phase starts (the kernel's barriers).")
            w('extern "C" int tcsg_emulate(const tcsg_kargs* a, uint32_t gx, uint32_t gy, uint32_t gz,')
            w("                            uint32_t bx, uint32_t by, uint32_t bz) {")
            w(f"  if (bx != 64u || by != {BY}u || bz != 1u) return -1;")
            w(f"  std::vector<TV> lds({total});")
            w("  tcsg_lds_lo = lds.data();")
            w(f"  std::vector<tcsg_pf> pf({NT});")
            w("  tcsg_w W;")
            w("  tcsg_weights(*a, W);")
            w("  for (uint32_t wz = 0; wz < gz; ++wz) for (uint32_t wy = 0; wy < gy; ++wy) for (uint32_t wx = 0; wx < gx; ++wx) {")
            w(f"    for (u32 tid = 0; tid < {NT}u; ++tid) tcsg_fetch(*a, wx, wy, wz, tid, 0u, pf[tid]);")
            w(f"    for (u32 t = 0; t < {NQ}u; ++t) {{")
            w(f"      for (u32 tid = 0; tid < {NT}u; ++tid) {{")
            w("        tcsg_put(lds.data(), tid, t, pf[tid]);")
            w(f"        if (t + 1u < {NQ}u) tcsg_fetch(*a, wx, wy, wz, tid, t + 1u, pf[tid]);")
            w("      }")
            for s_ in range(1, K + 1):
                w(f"      for (u32 tid = 0; tid < {NT}u; ++tid) tcsg_level{s_}(*a, W, lds.data(), wx, wy, wz, tid, t);")
            w("    }")
            w("  }")
            w("  return 0;")
            w("}")
        else:
            w(f'extern "C" __global__ void __launch_bounds__({d.max_threads}) tcsg_update(const tcsg_kargs a)')
            w("{")
            w(f"  __shared__ TV lds[{total}];")
            w("  const u32 tid = threadIdx.x + 64u * threadIdx.y;")
            w("  const u32 bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;")
            w("  tcsg_w W;")
            w("  tcsg_weights(a, W);")
            w("  tcsg_pf pf;")
            w("  tcsg_fetch(a, bx, by, bz, tid, 0u, pf);")
            w(f"  for (u32 t = 0; t < {NQ}u; ++t) {{")
            w("    tcsg_put(lds, tid, t, pf);")
            w(f"    if (t + 1u < {NQ}u) tcsg_fetch(a, bx, by, bz, tid, t + 1u, pf);")
            w("    __syncthreads();")
            for s_ in range(1, K + 1):
                w(f"    tcsg_level{s_}(a, W, lds, bx, by, bz, tid, t);")
                if syncs[s_ - 1]:
                    w("    __syncthreads();")
            w("  }")
            w("}")
        return "\n".join(self.L) + "\n"

    def source_temporal(self) -> str:
        """2D kernel that advances tsteps time steps per launch.  Workgroup
        (64, BY) owns an output tile of TX = 64 xc by TY = BY yc points; it
        loads the tile plus H = tsteps*r points per side into LDS, then each
        sub-step s updates the region s*r inside the LDS tile from the previous
        sub-step's values (ping-pong between two LDS buffers), and the last
        sub-step's tile goes to the output.  Points outside the interior (the
        Dirichlet halo) keep their values; points past the allocation are read
        as 0 and are never inputs of an interior point.  Every interior point
        sums its terms in the reference's order with the same operations as a
        one-step kernel, so k launches equal k*tsteps steps bit for bit."""
        d = self.d
        T = self.T
        w = self.w
        cpu = self.target == "cpu"
        R, LO, K = self.R, d.left_offset, d.tsteps
        H = K * R
        BY = d.max_threads // 64
        TX, TY, LW, LH = d.tile
        self._temporal_head(f"2D {'Box' if d.neighbourhood == BOX else 'Star'} r={R} "
                            f"{'Variable' if d.coeff == VARIABLE else 'Jacobi'} "
                            f"{'Row' if d.orientation == ROW else 'Column'} alignment "
                            f"{['None', 'Read', 'Write'][d.alignment]}/{d.cl_size} {'f32c' if self.f32 else 'f64'} "
                            f"temporal: {K} steps per launch, tile {TX}x{TY}, block 64x{BY}", 2 * LW * LH)
        fn = "static inline" if cpu else "__device__ __forceinline__"
        w(f"#define TCSG_LW {LW}")
        w(f"#define TCSG_LH {LH}")
        w(f"#define TCSG_BUF {LW * LH}")
        w("// interior of the padded grid: x in [r + left offset, xdim - right offset - r), y in [r, ydim - r)")
        w(f"{fn} bool tcsg_interior(const tcsg_kargs& a, long long gx, long long gy) {{")
        w(f"  return gx >= {R + LO} && gx < (long long)a.xdim - (long long)a.right_offset - {R} &&")
        w(f"         gy >= {R} && gy < (long long)a.ydim - {R};")
        w("}")
        w("// global x and y of the LDS tile's first column and row")
        w(f"{fn} long long tcsg_x0(u32 bx) {{ return (long long)bx * {TX} + {R + LO - H}; }}")
        w(f"{fn} long long tcsg_y0(u32 by) {{ return (long long)by * {TY} + {R - H}; }}")
        w("")
        w("// the tile and its halo from the input grid into LDS buffer 0: every load of the thread")
        w("// is issued before the first LDS write (rows ty + k*BY, columns tx + 64c)")
        w(f"{fn} void tcsg_load(const tcsg_kargs& a, TV* tb, u32 bx, u32 by, u32 tx, u32 ty) {{")
        w("  const long long x0 = tcsg_x0(bx), y0 = tcsg_y0(by);")
        nr, ncol = (LH + BY - 1) // BY, (LW + 63) // 64
        for c in range(ncol):
            w(f"  const long long gx{c} = x0 + tx + {64 * c}u;")
            w(f"  const bool cok{c} = {f'tx + {64 * c}u < TCSG_LW && ' if 64 * (c + 1) > LW else ''}"
              f"gx{c} >= 0 && gx{c} < (long long)a.xdim;")
        for r in range(nr):
            w(f"  const long long gy{r} = y0 + ty + {r * BY}u;")
            w(f"  const bool rok{r} = {f'ty + {r * BY}u < TCSG_LH && ' if BY * (r + 1) > LH else ''}"
              f"gy{r} >= 0 && gy{r} < (long long)a.ydim;")
            for c in range(ncol):
                w(f"  const TV l{r}_{c} = rok{r} && cok{c} ? a.in[(size_t)gy{r} * a.xdim + (size_t)gx{c}] : ({T})0;")
        for r in range(nr):
            for c in range(ncol):
                cond = []
                if BY * (r + 1) > LH:
                    cond.append(f"ty + {r * BY}u < TCSG_LH")
                if 64 * (c + 1) > LW:
                    cond.append(f"tx + {64 * c}u < TCSG_LW")
                st = f"TCSG_LDS(tb, (ty + {r * BY}u) * TCSG_LW + tx + {64 * c}u) = l{r}_{c};"
                w(f"  if ({' && '.join(cond)}) {st}" if cond else f"  {st}")
        w("}")
        w("")
        w("// sub-step s (1..tsteps): the region s*r inside the tile, from buffer (s-1)&1 into s&1.")
        w("// Each wave walks a strip of consecutive rows down its columns, keeping the rows")
        w("// above and below the current one in registers: per point it reads from LDS only")
        w("// the row entering the window (and, for a Star, the current row's x-neighbours).")
        w(f"{fn} void tcsg_substep(const tcsg_kargs& a, TV* tb, u32 bx, u32 by, u32 tx, u32 ty, u32 s) {{")
        w("  const long long x0 = tcsg_x0(bx), y0 = tcsg_y0(by);")
        w("  const TV* cur = tb + ((s - 1u) & 1u) * TCSG_BUF;")
        w("  TV* nxt = tb + (s & 1u) * TCSG_BUF;")
        w(f"  const u32 m = s * {R}u;")
        w(f"  const u32 rows = (TCSG_LH - 2u * m + {BY - 1}u) / {BY}u;")
        w("  const u32 ja = m + ty * rows;")
        w("  const u32 je = ja + rows < TCSG_LH - m ? ja + rows : TCSG_LH - m;")
        if d.coeff == VARIABLE:
            w(f"  const {T}* mk = a.mask;")
            for mi in range(d.n):
                w(f"  const TV w{mi} = mk[{mi}];")
        box = d.neighbourhood == BOX
        # the window: rows j-r..j+r in 2r+1 register slots (a Box keeps each row's 2r+1 columns,
        # a Star only its y arm's column); row ja-r+t sits in slot t mod (2r+1), and the walk is
        # unrolled by 2r+1 rows so that each slot keeps its registers (no moves between rows)
        W = 2 * R + 1
        win_dx = list(range(-R, R + 1)) if box else [0]
        sn = lambda dx, k: f"u_{_n(dx)}_{k}"
        def walk(checked):
            w("  for (u32 i = m + tx; i < TCSG_LW - m; i += 64u) {")
            w("    const TV* col = cur + i;  // signed offsets from here: the window reaches r columns left")
            w("    const long long gx = x0 + i;")
            w("    if (ja >= je) break;")
            for t in range(2 * R):
                for dx in win_dx:
                    w(f"    TV {sn(dx, t)} = TCSG_LDS(col, (int)(ja * TCSG_LW) + ({(t - R) * LW + dx}));")
            for dx in win_dx:
                w(f"    TV {sn(dx, 2 * R)};")
            w(f"    for (u32 j0 = ja; j0 < je; j0 += {W}u) {{")
            o = (0, 0, 0)
            for q in range(W):
                ind = "      "
                w(f"{ind}if (j0 + {q}u < je) {{")
                w(f"{ind}  const u32 j = j0 + {q}u;")
                slot = lambda dy: (q + dy + R) % W
                for dx in win_dx:
                    w(f"{ind}  {sn(dx, slot(R))} = TCSG_LDS(col, (int)(j * TCSG_LW) + ({R * LW + dx}));")
                val = {}
                for dy in range(-R, R + 1):
                    for dx in win_dx:
                        val[(dx, dy)] = sn(dx, slot(dy))
                if not box:
                    for dx in range(-R, R + 1):
                        if dx:
                            w(f"{ind}  const TV e_{_n(dx)} = TCSG_LDS(col, (int)(j * TCSG_LW) + ({dx}));")
                            val[(dx, 0)] = f"e_{_n(dx)}"
                w(f"{ind}  TV res = {val[(0, 0)]};")
                w(f"{ind}  if ({'tcsg_interior(a, gx, y0 + j)' if checked else 'true'}) {{")
                for t, (dx, dy, dz) in enumerate(self.offs):
                    mi = spec.mask_index(d, dx, dy, dz)
                    w(f"{ind}    " + self.acc_stmt(o, t, mi, val[(dx, dy)]))
                w(f"{ind}    res = {self.final_expr(o)};")
                w(f"{ind}  }}")
                w(f"{ind}  TCSG_LDS(nxt, j * TCSG_LW + i) = res;")
                w(f"{ind}}}")
            w("    }")
            w("  }")


        # a sub-step region wholly inside the interior (most tiles) needs no per-point test
        w(f"  const bool inner = x0 + m >= {R + LO} && x0 + TCSG_LW - m <= (long long)a.xdim - (long long)a.right_offset - {R} &&")
        w(f"                     y0 + m >= {R} && y0 + TCSG_LH - m <= (long long)a.ydim - {R};")
        w("  if (inner) {")
        walk(False)
        w("  } else {")
        walk(True)
        w("  }")
        w("}")
        w("")
        w("// the tile's interior points of the last sub-step to the output grid")
        w(f"{fn} void tcsg_store(const tcsg_kargs& a, const TV* tb, u32 bx, u32 by, u32 tx, u32 ty) {{")
        w("  const long long x0 = tcsg_x0(bx), y0 = tcsg_y0(by);")
        w(f"  const TV* fin = tb + {K & 1}u * TCSG_BUF;")
        w(f"  for (u32 j = {H}u + ty; j < {H + TY}u; j += {BY}u) {{")
        w(f"    for (u32 i = {H}u + tx; i < {H + TX}u; i += 64u) {{")
        w("      if (tcsg_interior(a, x0 + i, y0 + j))")
        w("        a.out[(size_t)(y0 + j) * a.xdim + (size_t)(x0 + i)] = TCSG_LDS(fin, j * TCSG_LW + i);")
        w("    }")
        w("  }")
        w("}")
        w("")
        if cpu:
            w("// One launch: the workgroups one after another; within a workgroup each phase runs for")
            w("// every thread before the next phase starts (the kernel's barriers).")
            w('extern "C" int tcsg_emulate(const tcsg_kargs* a, uint32_t gx, uint32_t gy, uint32_t gz,')
            w("                            uint32_t bx, uint32_t by, uint32_t bz) {")
            w(f"  if (bx != 64u || by != {BY}u || bz != 1u || gz != 1u) return -1;")
            w("  std::vector<TV> tb(2 * TCSG_BUF);")
            w("  tcsg_lds_lo = tb.data();")
            w("  for (uint32_t wy = 0; wy < gy; ++wy) for (uint32_t wx = 0; wx < gx; ++wx) {")
            w(f"    for (u32 t = 0; t < {64 * BY}u; ++t) tcsg_load(*a, tb.data(), wx, wy, t % 64u, t / 64u);")
            w(f"    for (u32 s = 1; s <= {K}u; ++s)")
            w(f"      for (u32 t = 0; t < {64 * BY}u; ++t) tcsg_substep(*a, tb.data(), wx, wy, t % 64u, t / 64u, s);")
            w(f"    for (u32 t = 0; t < {64 * BY}u; ++t) tcsg_store(*a, tb.data(), wx, wy, t % 64u, t / 64u);")
            w("  }")
            w("  return 0;")
            w("}")
        else:
            w(f'extern "C" __global__ void __launch_bounds__({d.max_threads}) tcsg_update(const tcsg_kargs a)')
            w("{")
            w("  __shared__ TV tb[2 * TCSG_BUF];")
            w("  const u32 tx = threadIdx.x, ty = threadIdx.y;")
            w("  tcsg_load(a, tb, blockIdx.x, blockIdx.y, tx, ty);")
            w("  __syncthreads();")
            w(f"  for (u32 s = 1; s <= {K}u; ++s) {{")
            w("    tcsg_substep(a, tb, blockIdx.x, blockIdx.y, tx, ty, s);")
            w("    __syncthreads();")
            w("  }")
            w("  tcsg_store(a, tb, blockIdx.x, blockIdx.y, tx, ty);")
            w("}")
        return "\n".join(self.L) + "\n"

    def source(self, plugin: bool = False) -> str:
        d = self.d
        if d.variant == TEMPORAL:
            return self.source_temporal3d() if d.dimension == 3 else self.source_temporal()
        self.preamble()
        attr = ""
        if d.variant == STREAM and d.waves_per_eu and self.target != "cpu":
            attr = f" __attribute__((amdgpu_waves_per_eu({d.waves_per_eu})))"
        self.w(f'extern "C" __global__ void __launch_bounds__({d.max_threads}){attr} tcsg_update(const tcsg_kargs a)')
        self.w("{")
        if d.variant == PLAIN:
            self.body_plain()
        else:
            self.body_stream()
        self.w("}")
        if d.batched:
            self.batched_kernel(attr)
        if plugin:
            self.reference_entry_points(attr)
        if self.target == "cpu":
            self.emulation_runner()
        return "\n".join(self.L) + "\n"

    def reference_entry_points(self, attr):
        """The symbols the reference's kernel file provides to the launch unit
        that its generator prints (hip/generateLaunchUpdateGrid.py:55-58):
        updateGrid with the reference's arguments (hip/generate2DStencil_hip.py:70-76,
        hip/generate3DStencil_hip.py:73-79), init with its initial values
        (hip/generate2DStencil_hip.py:322-351, hip/generate3DStencil_hip.py:387-431),
        the __constant__ weights d_mask and the macro range (:221-226).  That
        launch unit checks nothing, so updateGrid runs this kernel's body only
        where it applies (block shape, 32-bit buffer addressing) and the
        reference's code shape, which sums in the same order, elsewhere."""
        d = self.d
        w = self.w
        R, three = d.range, self.three
        aligned = d.alignment != ALIGN_NONE
        w("")
        w("// ---- the reference kernel file's symbols, for its own launchUpdateGrid.cu ----")
        w(f"#define TCSG_PLUGIN_DIM {d.dimension}")
        w(f"#define TCSG_PLUGIN_RIGHT_OFFSET {int(aligned)}")
        w(f"#define TCSG_PLUGIN_VARIABLE {int(d.coeff == VARIABLE)}")
        w(f"#define TCSG_PLUGIN_STENCIL_SIZE {d.n}")
        w("#ifndef range")
        w(f"#define range {R}")
        w("#endif")
        if d.coeff == VARIABLE:
            w(f"__constant__ double d_mask[{d.n}];")

        def device_body(name, desc):
            e = _Emitter(desc, self.target)
            if desc.variant == PLAIN:
                e.body_plain()
            else:
                e.body_stream()
            w(f"__device__ __forceinline__ void {name}(const tcsg_kargs a)")
            w("{")
            self.L.extend(e.L)
            w("}")

        device_body("tcsg_step_fast", d)
        fallback = d.variant == STREAM
        if fallback:
            device_body("tcsg_step_plain", d.replace(variant=PLAIN, xext="load", xcd_remap=False, nt_store=False,
                                                     nt_load=False, zfirst=False, prefetch=0, waves_per_eu=0,
                                                     batched=False, order="point", edge_pairs=False))
        args = "const unsigned long xdim, const unsigned long ydim"
        if three:
            args += ", const unsigned long zdim, double * const __restrict__ array0"
        else:
            args += ", const double * const __restrict__ array0"
        args += ", double * const __restrict__ array1"
        if aligned:
            args += ", const unsigned int rightOffset"
        w(f"__global__ void __launch_bounds__({d.max_threads}){attr} updateGrid({args})")
        w("{")
        w("  tcsg_kargs a;")
        w("  a.in = array0;")
        w("  a.out = array1;")
        w(f"  a.mask = {'d_mask' if d.coeff == VARIABLE else 'nullptr'};")
        w("  a.xdim = (uint32_t)xdim;")
        w("  a.ydim = (uint32_t)ydim;")
        w(f"  a.zdim = {'(uint32_t)zdim' if three else '1u'};")
        w(f"  a.right_offset = {'rightOffset' if aligned else '0u'};")
        w(f"  const uint64_t bytes = (uint64_t)xdim * ydim{' * zdim' if three else ''} * 8u;")
        w("  a.in_bytes = a.out_bytes = bytes;")
        w("  a.rec_limit = a.oob = 0xFFFFFFFFu;")
        if fallback:
            H = self.halo_lanes()
            w("  // this kernel's body addresses one block's window of planes (rows in 2D) with 32-bit")
            w("  // offsets (the limits the runtime's launches compute: record cap rec = window + slack,")
            w("  // base offset oob = rec + slack of lanes outside the interior, no wrap below 2^32)")
            w("  // and, for its x-neighbour exchange, needs the block shapes the runtime's")
            w("  // tcsg_check_block accepts; the reference's code shape, summing in the same order,")
            w("  // takes the launches that do not fit")
            w("  const uint64_t sx = (uint64_t)xdim * 8u, sxy = sx * ydim;")
            w(f"  const uint64_t cols = {(2 * H * d.xc + d.xc + 2 * R) * 8}ull;")
            w(f"  const uint64_t span = {d.zc + 2 * R}ull * {'sxy' if three else '0ull'} + {d.yc + 2 * R}ull * sx + cols;")
            if three:
                w(f"  const uint64_t window = ((uint64_t)blockDim.z * {d.zc}u + {2 * R}u) * sxy;")
            else:
                w(f"  const uint64_t window = ((uint64_t)blockDim.y * {d.yc}u + {2 * R}u) * sx;")
            w("  const uint64_t slack = sx + cols;")
            w("  bool fast = window + 2u * slack + span < (1ull << 32);")
            w("  if (fast) {")
            w("    a.rec_limit = (uint32_t)(window + slack);")
            w("    a.oob = (uint32_t)(window + 2u * slack);")
            w("  }")
            if d.xext == "dppw":
                w("  fast = fast && blockDim.x % 64u == 0u;")
            elif d.xext == "dpp":
                w("  {")
                w("    // every run of lanes in one block row and one wave spans the lanes the exchange reaches")
                w("    const uint32_t nt = blockDim.x * blockDim.y * blockDim.z;")
                w("    const uint32_t lid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);")
                w("    const uint32_t lane = lid & 63u, tx = threadIdx.x;")
                w("    const uint32_t tl = tx < lane ? tx : lane;")
                w("    uint32_t tr = blockDim.x - 1u - tx < 63u - lane ? blockDim.x - 1u - tx : 63u - lane;")
                w("    tr = tr < nt - 1u - lid ? tr : nt - 1u - lid;")
                w(f"    fast = __syncthreads_and(fast && tl + tr + 1u >= {H}u) != 0;")
                w("  }")
            w("  if (fast) tcsg_step_fast(a); else tcsg_step_plain(a);")
        else:
            w("  tcsg_step_fast(a);")
        w("}")
        lo = d.left_offset if d.alignment == ALIGN_WRITE else 0
        iargs = "double * array, const unsigned long xdim, const unsigned long ydim"
        if three:
            iargs += ", const unsigned long zdim"
        if aligned:
            iargs += ", const unsigned int rightOffset"
        w(f"__global__ void init({iargs})")
        w("{")
        w(f"  const size_t x = blockIdx.x * blockDim.x + threadIdx.x + {lo};")
        w("  const size_t y = blockIdx.y * blockDim.y + threadIdx.y;")
        if three:
            w("  const size_t z = blockIdx.z * blockDim.z + threadIdx.z;")
        w(f"  if (x < xdim{' - rightOffset' if aligned else ''} && y < ydim{' && z < zdim' if three else ''}) {{")
        w(f"    const size_t idx = x + y * xdim{' + z * xdim * ydim' if three else ''};")
        w(f"    const bool halo = x - {lo} < range || y < range{' || z < range' if three else ''};")
        w(f"    array[idx] = halo ? {'(idx % 2 == 0 ? 0.75 : 0.25)' if three else '1.0'} : 0.0;")
        w("  }")
        w("}")

    def batched_kernel(self, attr):
        """Persistent kernel: the workgroups take tiles from a work queue over a
        tile table whose entries name a grid (its kernel arguments) and a block
        of that grid; each tile is the work of one block of tcsg_update."""
        w = self.w
        d = self.d
        cpu = self.target == "cpu"
        w("")
        if not cpu:
            w(f"typedef __attribute__((address_space(1))) {self.T} tcsg_g{self.T};")
        w("struct tcsg_tile { uint32_t grid, bx, by, bz; };")
        w("struct tcsg_batch_args { const tcsg_kargs* grids; const tcsg_tile* tiles; uint32_t ntiles, pad;")
        w("                         uint32_t* queue; };")
        w("// Work queue: the tile table is split into one contiguous range per XCD (the")
        w("// hardware hands workgroup i to XCD i % 8; neighbouring entries are neighbouring")
        w("// blocks of one grid, sharing halo lines in that XCD's L2), or a single range with")
        w("// fewer than 8 workgroups.  queue[x] counts the tiles taken from range x;")
        w("// queue[8] the workgroups that found every range drained.  A workgroup drains its")
        w("// own range, then takes from the following ones.  The last workgroup to finish")
        w("// clears the queue, so each launch (and each replay of a captured step loop)")
        w("// starts from zero; launches of one batch must not overlap.")
        if cpu:
            HD = "static inline"
            w("static inline uint32_t tcsg_atomic_add(uint32_t* p, uint32_t v) { const uint32_t o = *p; *p = o + v; return o; }")
            w("static inline void tcsg_atomic_store(uint32_t* p, uint32_t v) { *p = v; }")
            w("static inline void tcsg_fence() {}")
        else:
            HD = "__device__ __forceinline__"
            w("__device__ __forceinline__ uint32_t tcsg_atomic_add(uint32_t* p, uint32_t v) { return atomicAdd(p, v); }")
            w("__device__ __forceinline__ void tcsg_atomic_store(uint32_t* p, uint32_t v) { (void)atomicExch(p, v); }")
            w("__device__ __forceinline__ void tcsg_fence() { __threadfence(); }")
        w("// next tile for a workgroup on XCD xcd of nr ranges, or ~0u when all are drained;")
        w("// s: ranges (counted from xcd) this workgroup has found drained")
        w(f"{HD} uint32_t tcsg_fetch(uint32_t* q, uint32_t ntiles, uint32_t xcd, uint32_t nr, uint32_t& s) {{")
        w("  for (; s < nr; ++s) {")
        w("    const uint32_t x = (xcd + s) % nr;")
        w("    const uint32_t beg = (uint32_t)(((uint64_t)ntiles * x) / nr);")
        w("    const uint32_t end = (uint32_t)(((uint64_t)ntiles * (x + 1u)) / nr);")
        w("    if (beg == end) continue;")
        w("    const uint32_t k = tcsg_atomic_add(&q[x], 1u);")
        w("    if (k < end - beg) return beg + k;")
        w("  }")
        w("  return 0xFFFFFFFFu;")
        w("}")
        if cpu:
            w("template <class S> static inline S tcsg_uniform(const S& v) { return v; }")
        else:
            w("// A value every lane of the wave holds alike, moved to scalar registers word by word.")
            w("// The tile entry and the grid's arguments are read from memory the kernel also writes")
            w("// (the queue, other grids), so the compiler keeps them in vector registers; the")
            w("// buffer resources built from them must be scalar, and a vector resource would make")
            w("// it wrap every buffer access in a loop over the distinct values of the wave's lanes.")
            w("template <class S> __device__ __forceinline__ S tcsg_uniform(const S& v) {")
            w("  static_assert(sizeof(S) % 4 == 0, \"whole words\");")
            w("  uint32_t wd[sizeof(S) / 4];")
            w("  __builtin_memcpy(wd, &v, sizeof(S));")
            w("#pragma unroll")
            w("  for (unsigned i = 0; i < sizeof(S) / 4; ++i) wd[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)wd[i]);")
            w("  S u;")
            w("  __builtin_memcpy(&u, wd, sizeof(S));")
            w("  return u;")
            w("}")
        w(f"{HD} void tcsg_finish(uint32_t* q, uint32_t nwg) {{")
        w("  tcsg_fence();")
        w("  // the last workgroup here: every other one is done with the queue")
        w("  if (tcsg_atomic_add(&q[8], 1u) == nwg - 1u) {")
        w("    for (uint32_t x = 0; x < 8u; ++x) tcsg_atomic_store(&q[x], 0u);")
        w("    tcsg_atomic_store(&q[8], 0u);")
        w("  }")
        w("}")
        if cpu:
            # the emulation runs one wave at a time: the runner takes the tiles from the
            # queue (tcsg_fetch / tcsg_finish above) and runs every wave of the workgroup
            # on each, where the GPU kernel hands the tile to its waves through LDS
            w("void tcsg_update_batched_tile(const tcsg_batch_args ba, const uint32_t t)")
            w("{")
        else:
            w(f'extern "C" __global__ void __launch_bounds__({d.max_threads}){attr} '
              f"tcsg_update_batched(const tcsg_batch_args ba)")
            w("{")
            w("  __shared__ uint32_t tcsg_slot[2];  // the tile, from the leader to every wave")
            w("  const uint32_t nwg = gridDim.x;")
            w("  const uint32_t nr = nwg >= 8u ? 8u : 1u;")
            w("  const uint32_t xcd = nr == 8u ? (blockIdx.x & 7u) : 0u;")
            w("  const bool leader = threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0;")
            w("  uint32_t s = 0;")
            w("  for (uint32_t it = 0;; ++it) {")
            w("  // alternate slots: the leader writes the next one only after every wave has")
            w("  // passed this barrier, i.e. has read the previous one")
            w("  if (leader) tcsg_slot[it & 1u] = tcsg_fetch(ba.queue, ba.ntiles, xcd, nr, s);")
            w("  __syncthreads();")
            if d.variant == PLAIN:
                # (the PLAIN body builds no buffer resources; taking its tile as wave-uniform
                # measured 15-21% slower on the 64-grid batch)
                w("  const uint32_t t = tcsg_slot[it & 1u];")
            else:
                w("  // the same tile for every lane, taken as wave-uniform: the buffer resources built")
                w("  // from the grid's arguments are then scalar values")
                w("  const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)tcsg_slot[it & 1u]);")
            w("  if (t == 0xFFFFFFFFu) break;")
        if d.variant == PLAIN:
            w("  const tcsg_tile tl = ba.tiles[t];")
            w("  const tcsg_kargs a = ba.grids[tl.grid];")
        else:
            w("  const tcsg_tile tl = tcsg_uniform(ba.tiles[t]);")
            w("  const tcsg_kargs a = tcsg_uniform(ba.grids[tl.grid]);")
        self.batched_body = True
        if d.variant == PLAIN:
            w("  do {")
            self.body_plain()
            w("  } while (0);")
        else:
            self.body_stream()
        self.batched_body = False
        if not cpu:
            w("  }")
            w("  if (leader) tcsg_finish(ba.queue, nwg);")
        w("}")


# Straight-line kernels grow with outputs-per-thread x stencil size; past this
# many accumulate statements hipcc takes minutes and the code no longer fits
# the instruction cache, so the generator refuses the parameter set.
MAX_TERMS_PER_THREAD = 40000


def terms_per_thread(desc: StencilDesc) -> int:
    return desc.xc * desc.yc * desc.zc * desc.n


def emit_source(desc: StencilDesc, target: str = "gfx950", plugin: bool = False) -> str:
    """Source text of the kernel for ``desc``: HIP for gfx950, or (target
    "cpu", STREAM variant only) the same kernel body over the wave-level CPU
    emulation in cpu_emu.h, used to test the emitted code on the host.
    plugin (FP64 kernels for gfx950): also define the reference kernel file's
    updateGrid, init, d_mask and range, so the source can stand in for the
    reference's kernels.cu under the launch unit its generator prints."""
    if plugin and (target != "gfx950" or desc.dtype != F64):
        raise ValueError("the reference's kernel-file symbols exist for FP64 kernels built for the GPU")
    if terms_per_thread(desc) > MAX_TERMS_PER_THREAD:
        raise ValueError(f"coarsening {desc.coarsening} x stencil size {desc.n} = {terms_per_thread(desc)} "
                         f"accumulate statements per thread exceeds {MAX_TERMS_PER_THREAD}")
    if target == "cpu" and desc.variant == PLAIN:
        raise ValueError("CPU emulation covers the STREAM and TEMPORAL variants")
    if plugin and desc.variant == TEMPORAL:
        raise ValueError("the reference's updateGrid advances one step per launch; a temporal kernel has no such symbol")
    if plugin and desc.staged:
        raise ValueError("the reference's updateGrid takes any block shape; an LDS-staging kernel runs one")
    if target not in ("gfx950", "cpu"):
        raise ValueError(f"unknown target {target}")
    return _Emitter(desc, target).source(plugin)
