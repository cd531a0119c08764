// libtcsg.so -- runtime behind include/tcsg.h.
//
// Loads generated gfx950 code objects, owns their mask buffers, computes the
// padded geometry, launches steps and times them.  The device kernels in this
// file are the ones every stencil shares: grid initialisation, FP64->FP32
// conversion and the FP64 error reduction used by the FP32-storage variant.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <vector>

#include "tcsg.h"

namespace {

thread_local int g_last_hip = 0;

#define TCSG_HIP(call)                                  \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) {                         \
            g_last_hip = (int)e_;                       \
            return TCSG_ERR_HIP;                        \
        }                                               \
    } while (0)

constexpr uint32_t kMagic = 0x54435347u;
constexpr uint32_t kAbi = 1u;

// Kernel argument block; layout matches the emitted `struct tcsg_kargs`.
struct KArgs {
    const void *in;
    void *out;
    const void *mask;
    uint32_t xdim, ydim, zdim, right_offset;
    uint64_t in_bytes, out_bytes;  // buffer sizes
    uint32_t rec_limit, oob;       // STREAM: record count cap of a block's buffer window,
                                   // base offset of lanes outside the interior
};
static_assert(sizeof(KArgs) == 64, "kernel argument layout");

uint32_t stencil_size(const tcsg_desc &d)
{
    uint32_t w = 2 * d.range + 1;
    if (d.neighbourhood == TCSG_BOX) return d.dimension == 2 ? w * w : w * w * w;
    return 2 * d.dimension * d.range + 1;
}

size_t elem_size(const tcsg_desc &d) { return d.dtype == TCSG_F32C ? 4 : 8; }

bool desc_valid(const tcsg_desc &d)
{
    if (d.dimension != 2 && d.dimension != 3) return false;
    if (d.neighbourhood > 1 || d.orientation > 1 || d.coeff > 1 || d.dtype > 1 || d.variant > 2) return false;
    if (d.range < 1 || d.xc < 1 || d.yc < 1 || d.zc < 1) return false;
    if (d.dimension == 2 && d.zc != 1) return false;
    if (d.alignment > 2) return false;
    if (d.alignment != TCSG_ALIGN_NONE && d.cl_size == 0) return false;
    return true;
}

uint32_t left_offset(const tcsg_desc &d)
{
    if (d.alignment == TCSG_ALIGN_WRITE && d.range % d.cl_size) return d.cl_size - d.range % d.cl_size;
    return 0;
}

// ------------------------------------------------------------------ devices
__device__ __forceinline__ uint64_t sm64(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t seed, uint64_t idx)
{
    return (double)(sm64(seed, idx) >> 11) * 0x1.0p-53;
}

// Initial values over the whole padded grid.  Reference mode follows the
// generated init kernels (hip/generate2DStencil_hip.py:322-351,
// hip/generate3DStencil_hip.py:387-431); padding, which they leave unwritten,
// is set to 0 here.
template <typename T>
__global__ void init_kernel(T *a, uint32_t xdim, uint32_t ydim, uint32_t zdim, uint32_t lo, uint32_t ro,
                            uint32_t r, uint32_t dim, int mode, uint64_t seed)
{
    const uint64_t n = (uint64_t)xdim * ydim * zdim;
    const uint64_t nx = xdim - lo - ro - 2 * r, ny = ydim - 2 * r;
    const uint32_t zr = dim == 3 ? r : 0;
    for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; idx < n;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(idx % xdim);
        const uint64_t jk = idx / xdim;
        const uint32_t j = (uint32_t)(jk % ydim);
        const uint32_t k = (uint32_t)(jk / ydim);
        double v = 0.0;
        if (i >= lo && i < xdim - ro) {
            if (mode == TCSG_INIT_REFERENCE) {
                const bool halo = (i - lo < r) || (j < r) || (dim == 3 && k < r);
                if (halo) v = dim == 2 ? 1.0 : ((idx % 2 == 0) ? 0.75 : 0.25);
            } else {
                const bool inside = i >= lo + r && i < xdim - ro - r && j >= r && j < ydim - r &&
                                    k >= zr && k < zdim - zr;
                if (inside) {
                    const uint64_t li = (i - lo - r) + nx * ((j - r) + ny * (uint64_t)(k - zr));
                    if (mode == TCSG_INIT_U01) v = u01(seed, li);
                    else if (mode == TCSG_INIT_UPM1) v = (double)(sm64(seed, li) >> 11) * 0x1.0p-52 - 1.0;
                    else {
                        const double a1 = 1.0 - u01(seed, 2 * li), a2 = u01(seed, 2 * li + 1);
                        v = sqrt(-2.0 * log(a1)) * cos(6.283185307179586 * a2);
                    }
                }
            }
        }
        a[idx] = (T)v;
    }
}

__global__ void convert_kernel(const double *__restrict__ s, float *__restrict__ d, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}

__device__ __forceinline__ void atomic_max_nonneg(double *p, double v)
{
    // non-negative doubles order like their bit patterns
    atomicMax((unsigned long long *)p, (unsigned long long)__double_as_longlong(v));
}

// Layout of one grid buffer for the error reduction: row pitch, rows per
// plane and the offset of the first interior point.
struct GridIndex {
    uint64_t pitch, rows, first;
    __host__ __device__ uint64_t at(uint64_t i, uint64_t j, uint64_t k) const
    {
        return first + i + pitch * (j + rows * k);
    }
};

// FP64 error reduction of an FP32 state against an FP64 state over the
// interior (nx x ny x nz points); each state in its own padded layout.
__global__ void error_kernel(const float *__restrict__ a, GridIndex ga, const double *__restrict__ b,
                             GridIndex gb, uint64_t nx, uint64_t ny, uint64_t nz, double floor_v, double *out)
{
    const uint64_t n = nx * ny * nz;
    double mx = 0.0, mb = 0.0, s2 = 0.0, b2 = 0.0, mrel = 0.0;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = t % nx, jk = t / nx, j = jk % ny, k = jk / ny;
        const double bv = b[gb.at(i, j, k)];
        const double d = fabs((double)a[ga.at(i, j, k)] - bv);
        mx = fmax(mx, d);
        mb = fmax(mb, fabs(bv));
        s2 += d * d;
        b2 += bv * bv;
        mrel = fmax(mrel, d / fmax(fabs(bv), floor_v));
    }
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_down(mx, off));
        mb = fmax(mb, __shfl_down(mb, off));
        s2 += __shfl_down(s2, off);
        b2 += __shfl_down(b2, off);
        mrel = fmax(mrel, __shfl_down(mrel, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(&out[0], mx);
        atomic_max_nonneg(&out[1], mb);
        atomicAdd(&out[2], s2);
        atomicAdd(&out[3], b2);
        atomic_max_nonneg(&out[4], mrel);
    }
}

// remainder: keep every point of an axis whose interior is not a multiple of
// the coarsening (kernels generated for it start the last thread's footprint
// early enough to end at the interior's end; STREAM kernels in x drop its
// columns past the interior instead); otherwise the axis is truncated to a
// multiple, as the reference does.
int geometry_compute(const tcsg_desc &d, tcsg_dim3 size, tcsg_geometry &g, bool remainder = false)
{
    g = tcsg_geometry{};
    const uint32_t r = d.range;
    uint32_t sz[3] = {size.x, size.y, d.dimension == 3 ? size.z : 1u};
    const uint32_t c[3] = {d.xc, d.yc, d.zc};
    uint32_t t[3] = {1, 1, 1}, e[3] = {1, 1, 1};
    for (uint32_t a = 0; a < d.dimension; ++a) {
        if (sz[a] <= 2 * r) return TCSG_ERR_ARG;
        t[a] = remainder ? sz[a] : sz[a] - (sz[a] - 2 * r) % c[a];
        e[a] = t[a] - 2 * r;
        // remainder kernels shift the last thread of an axis back by less than
        // one coarsening (y and z; x too in PLAIN kernels), which needs an
        // interior of at least one
        // (temporal kernels test every point against the interior instead)
        if (remainder && (a > 0 || d.variant == TCSG_PLAIN) && d.variant != TCSG_TEMPORAL && e[a] < c[a])
            return TCSG_ERR_ARG;
    }
    uint64_t xmem = t[0], lo = 0, ro = 0;
    if (d.alignment == TCSG_ALIGN_READ) {
        if (xmem % d.cl_size) { ro = d.cl_size - xmem % d.cl_size; xmem += ro; }
    } else if (d.alignment == TCSG_ALIGN_WRITE) {
        lo = left_offset(d);
        xmem += lo;
        if (xmem % d.cl_size) ro = d.cl_size - xmem % d.cl_size;
        xmem += ro;
    }
    if (xmem > 0xFFFFFFFFull) return TCSG_ERR_TOO_LARGE;
    g.size = {(uint32_t)xmem, t[1], d.dimension == 3 ? t[2] : 1u};
    g.elements = {e[0], e[1], d.dimension == 3 ? e[2] : 1u};
    g.threads = {(e[0] + c[0] - 1) / c[0], (e[1] + c[1] - 1) / c[1],
                 d.dimension == 3 ? (e[2] + c[2] - 1) / c[2] : 1u};
    g.left_offset = (uint32_t)lo;
    g.right_offset = (uint32_t)ro;
    g.dimension = d.dimension;
    g.range = r;
    g.n_points = (uint64_t)g.size.x * g.size.y * g.size.z;
    return TCSG_OK;
}

} // namespace

// Lets the other translation units of the library (the tuner's network)
// record a failed HIP call for tcsg_last_hip_error.
__attribute__((visibility("hidden"))) void tcsg_note_hip_error(int e) { g_last_hip = e; }

struct tcsg_kernel {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    tcsg_desc desc{};
    uint32_t params[24]{};
    void *mask = nullptr;
    bool mask_set = false;
    int legal_max = 0;
    int xext_mode = 0;    // STREAM x-neighbour exchange (params[16]): 0 loads, 1 DPP, 2 wave-aligned DPP
    bool remainder = false;  // kernel generated with remainder handling (params[18] bit 2)
    uint32_t tsteps = 1;     // time steps one launch advances (temporal kernels: params[18] bits 8-15)
    tcsg_dim3 stage{0, 0, 0};  // STREAM with LDS staging: the block shape it stages for (params[18]
                               // bit 3; bx/64, by, bz in bits 16-20, 21-25, 26-30), else 0
    bool pack = false;  // batched kernel over packed tiles (params[23] bit 1): a tile is a run of
                        // block.x * block.y consecutive threads of a grid's xy thread space (x
                        // fastest) times block.z thread planes
};

extern "C" {

const char *tcsg_status_string(int s)
{
    switch (s) {
    case TCSG_OK: return "ok";
    case TCSG_ERR_ARG: return "invalid argument";
    case TCSG_ERR_HIP: return "HIP runtime error";
    case TCSG_ERR_MISMATCH: return "code object parameters differ from the descriptor";
    case TCSG_ERR_BLOCK_SHAPE: return "block shape not launchable for this kernel";
    case TCSG_ERR_TOO_LARGE: return "grid too large for this kernel's addressing";
    case TCSG_ERR_NO_MASK: return "Variable kernel launched before its mask was set";
    case TCSG_ERR_FILE: return "code object missing or unreadable";
    default: return "unknown status";
    }
}

int tcsg_last_hip_error(void) { return g_last_hip; }

int tcsg_geometry_compute(const tcsg_desc *desc, tcsg_dim3 size, tcsg_geometry *out)
{
    if (!desc || !out || !desc_valid(*desc)) return TCSG_ERR_ARG;
    return geometry_compute(*desc, size, *out);
}

tcsg_dim3 tcsg_grid_blocks(const tcsg_geometry *g, tcsg_dim3 b)
{
    tcsg_dim3 r{0, 0, 0};
    if (!g || !b.x || !b.y || !b.z) return r;
    r.x = (g->threads.x + b.x - 1) / b.x;
    r.y = (g->threads.y + b.y - 1) / b.y;
    r.z = (g->threads.z + b.z - 1) / b.z;
    return r;
}

int tcsg_kernel_load(tcsg_kernel **out, const char *path, const tcsg_desc *expect)
{
    if (!out || !path) return TCSG_ERR_ARG;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) return TCSG_ERR_FILE;
    std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (bytes.empty()) return TCSG_ERR_FILE;
    tcsg_kernel *k = new tcsg_kernel();
    hipError_t e = hipModuleLoadData(&k->module, bytes.data());
    if (e != hipSuccess) { g_last_hip = (int)e; delete k; return TCSG_ERR_HIP; }
    auto fail = [&](int code) { (void)hipModuleUnload(k->module); delete k; return code; };
    e = hipModuleGetFunction(&k->fn, k->module, "tcsg_update");
    if (e != hipSuccess) { g_last_hip = (int)e; return fail(TCSG_ERR_HIP); }
    hipDeviceptr_t p = nullptr;
    size_t bytes_p = 0;
    e = hipModuleGetGlobal(&p, &bytes_p, k->module, "tcsg_params");
    if (e != hipSuccess || bytes_p != sizeof(k->params)) { g_last_hip = (int)e; return fail(TCSG_ERR_HIP); }
    e = hipMemcpyDtoH(k->params, p, sizeof(k->params));
    if (e != hipSuccess) { g_last_hip = (int)e; return fail(TCSG_ERR_HIP); }
    const uint32_t *q = k->params;
    if (q[0] != kMagic || q[1] != kAbi) return fail(TCSG_ERR_MISMATCH);
    tcsg_desc d{q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11], q[12], q[13]};
    if (!desc_valid(d) || q[14] != stencil_size(d)) return fail(TCSG_ERR_MISMATCH);
    if (expect && std::memcmp(expect, &d, sizeof(d)) != 0) return fail(TCSG_ERR_MISMATCH);
    k->desc = d;
    int v = 0;
    e = hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, k->fn);
    if (e != hipSuccess) { g_last_hip = (int)e; return fail(TCSG_ERR_HIP); }
    k->legal_max = std::min<int>(v, (int)q[15]);
    // The attribute bounds the block by the launch bound only; a kernel whose
    // registers do not fit that many waves on one CU cannot be launched with
    // such a block, so lower the bound until the occupancy calculator places
    // at least one workgroup.
    k->legal_max -= k->legal_max % 64;
    while (k->legal_max >= 64) {
        int nb = 0;
        e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k->fn, k->legal_max, 0);
        if (e != hipSuccess) { g_last_hip = (int)e; return fail(TCSG_ERR_HIP); }
        if (nb > 0) break;
        k->legal_max -= 64;
    }
    if (k->legal_max < 64) return fail(TCSG_ERR_BLOCK_SHAPE);
    k->xext_mode = d.variant == TCSG_STREAM ? (int)q[16] : 0;
    if (k->xext_mode > 2) return fail(TCSG_ERR_MISMATCH);
    k->remainder = (q[18] & 4u) != 0;
    k->pack = (q[23] & 2u) != 0;
    if (k->pack && (d.dimension != 3 || k->xext_mode != 0 || (q[18] & 8u) || !(q[23] & 1u)))
        return fail(TCSG_ERR_MISMATCH);
    k->tsteps = d.variant == TCSG_TEMPORAL ? (q[18] >> 8) & 0xFFu : 1u;
    if (k->tsteps < 1 || (d.variant == TCSG_TEMPORAL && q[15] % 64)) return fail(TCSG_ERR_MISMATCH);
    if (q[18] & 8u) {
        k->stage = tcsg_dim3{((q[18] >> 16) & 31u) * 64u, (q[18] >> 21) & 31u, (q[18] >> 26) & 31u};
        if (d.variant != TCSG_STREAM || d.dimension != 3 || k->xext_mode != 2 || !k->stage.x ||
            !k->stage.y || !k->stage.z)
            return fail(TCSG_ERR_MISMATCH);
    }
    *out = k;
    return TCSG_OK;
}

int tcsg_kernel_free(tcsg_kernel *k)
{
    if (!k) return TCSG_OK;
    if (k->mask) (void)hipFree(k->mask);
    if (k->module) (void)hipModuleUnload(k->module);
    delete k;
    return TCSG_OK;
}

int tcsg_kernel_desc(const tcsg_kernel *k, tcsg_desc *out)
{
    if (!k || !out) return TCSG_ERR_ARG;
    *out = k->desc;
    return TCSG_OK;
}

int tcsg_set_mask(tcsg_kernel *k, const double *mask, uint32_t n)
{
    if (!k || !mask) return TCSG_ERR_ARG;
    if (n != stencil_size(k->desc)) return TCSG_ERR_ARG;
    const size_t es = elem_size(k->desc);
    if (!k->mask) TCSG_HIP(hipMalloc(&k->mask, n * es));
    if (es == 8) {
        TCSG_HIP(hipMemcpy(k->mask, mask, n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(n);
        for (uint32_t i = 0; i < n; ++i) f[i] = (float)mask[i];
        TCSG_HIP(hipMemcpy(k->mask, f.data(), n * 4, hipMemcpyHostToDevice));
    }
    k->mask_set = true;
    return TCSG_OK;
}

int tcsg_max_block_size(tcsg_kernel *k, int *occ, int *legal)
{
    if (!k) return TCSG_ERR_ARG;
    if (occ) {
        int min_grid = 0, block = 0;
        TCSG_HIP(hipModuleOccupancyMaxPotentialBlockSize(&min_grid, &block, k->fn, 0, 0));
        *occ = block;
    }
    if (legal) *legal = k->legal_max;
    return TCSG_OK;
}

int tcsg_kernel_resources(tcsg_kernel *k, uint32_t threads, int *vgprs, int *scratch_bytes, int *workgroups_per_cu)
{
    if (!k) return TCSG_ERR_ARG;
    if (vgprs) TCSG_HIP(hipFuncGetAttribute(vgprs, HIP_FUNC_ATTRIBUTE_NUM_REGS, k->fn));
    if (scratch_bytes) TCSG_HIP(hipFuncGetAttribute(scratch_bytes, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k->fn));
    if (workgroups_per_cu) {
        *workgroups_per_cu = 0;
        if (threads == 0 || threads > (uint32_t)k->legal_max) return TCSG_OK;
        int n = 0;
        TCSG_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, k->fn, (int)threads, 0));
        *workgroups_per_cu = n;
    }
    return TCSG_OK;
}

int tcsg_check_block(const tcsg_kernel *k, tcsg_dim3 b)
{
    if (!k) return TCSG_ERR_ARG;
    if (!b.x || !b.y || !b.z || b.x > 1024 || b.y > 1024 || b.z > 1024) return TCSG_ERR_BLOCK_SHAPE;
    const uint64_t t = (uint64_t)b.x * b.y * b.z;
    if (t > (uint64_t)k->legal_max) return TCSG_ERR_BLOCK_SHAPE;
    if (k->xext_mode == 2 && b.x % 64 != 0) return TCSG_ERR_BLOCK_SHAPE;  // block rows must be whole waves
    // a temporal kernel's LDS tile is sized for its compiled block (64, launch bound / 64, 1)
    if (k->desc.variant == TCSG_TEMPORAL && (b.x != 64 || b.y != k->params[15] / 64 || b.z != 1))
        return TCSG_ERR_BLOCK_SHAPE;
    // an LDS-staging kernel's tile is sized and indexed for the block it was generated for
    if (k->stage.x && (b.x != k->stage.x || b.y != k->stage.y || b.z != k->stage.z)) return TCSG_ERR_BLOCK_SHAPE;
    if (k->xext_mode == 1) {
        // every run of lanes in one block row and one wave must span the
        // H = ceil(r/xc) lanes the x-neighbour exchange reaches
        const uint32_t H = (k->desc.range + k->desc.xc - 1) / k->desc.xc;
        for (uint64_t lid = 0; lid < t; ++lid) {
            const uint32_t tx = (uint32_t)(lid % b.x), lane = (uint32_t)(lid & 63);
            const uint32_t tl = std::min(tx, lane);
            uint32_t tr = std::min(b.x - 1 - tx, 63u - lane);
            // a partial last wave ends at the block's last thread
            tr = (uint32_t)std::min<uint64_t>(tr, t - 1 - lid);
            if (tl + tr + 1 < H) return TCSG_ERR_BLOCK_SHAPE;
        }
    }
    return TCSG_OK;
}

// STREAM kernels address their buffers with 32-bit offsets from a base at the first
// footprint plane (3D) or row (2D) of the launching block.  For block extent `bext` along
// that axis:
//   window W = (bext*zc + 2r) plane pitches (2D: (bext*yc + 2r) row pitches): the block's
//              planes or rows;
//   slack  S = one row pitch + (2H*xc + xc + 2r) elements: how far the edge lanes' loads
//              reach past the window's last row;
//   span     = (zc + 2r) plane pitches + (yc + 2r) row pitches + (2H*xc + xc + 2r) elements:
//              the most a lane adds to its base offset.
// Valid lanes' accesses lie below rec = W + S, the record-count cap; lanes outside the
// interior get the base offset oob = rec + S, so their accesses (at least oob - 2H*xc
// elements) miss the records, and oob + span < 2^32 keeps their offsets from wrapping.
struct StreamLimits {
    uint32_t rec, oob;
};

static int stream_limits(const tcsg_desc &d, const tcsg_geometry *g, uint32_t bext, uint32_t brows, StreamLimits *out)
{
    const uint64_t es = elem_size(d);
    const uint64_t sx = (uint64_t)g->size.x * es, sxy = sx * g->size.y;
    const uint64_t H = (d.range + d.xc - 1) / d.xc;
    const uint64_t cols = (2 * H * d.xc + d.xc + 2ull * d.range) * es;
    // rows a lane's loads reach from its base: its footprint's (brows = 1), or with LDS staging
    // those of its block's brows thread rows
    const uint64_t rows = (uint64_t)brows * d.yc + 2ull * d.range;
    const uint64_t span = (d.zc + 2ull * d.range) * (d.dimension == 3 ? sxy : 0) + rows * sx + cols;
    const uint64_t window = d.dimension == 3 ? ((uint64_t)bext * d.zc + 2ull * d.range) * sxy
                                             : ((uint64_t)bext * d.yc + 2ull * d.range) * sx;
    const uint64_t slack = sx + cols;
    if (window + 2 * slack + span >= (1ull << 32)) return TCSG_ERR_TOO_LARGE;
    if (out) *out = StreamLimits{(uint32_t)(window + slack), (uint32_t)(window + 2 * slack)};
    return TCSG_OK;
}

static int check_geom(const tcsg_desc &d, const tcsg_geometry *g, bool remainder = false)
{
    if (g->dimension != d.dimension || g->range != d.range) return TCSG_ERR_ARG;
    tcsg_geometry ref;
    // the geometry must be one this kernel's coarsening and alignment produce
    tcsg_dim3 user{g->size.x - g->left_offset - g->right_offset, g->size.y, g->size.z};
    if (geometry_compute(d, user, ref, remainder) != TCSG_OK) return TCSG_ERR_ARG;
    auto eq3 = [](tcsg_dim3 a, tcsg_dim3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    if (!eq3(ref.size, g->size) || !eq3(ref.elements, g->elements) || !eq3(ref.threads, g->threads) ||
        ref.left_offset != g->left_offset || ref.right_offset != g->right_offset || ref.n_points != g->n_points)
        return TCSG_ERR_ARG;
    // (the smallest block window: launches check theirs)
    if (d.variant == TCSG_STREAM) return stream_limits(d, g, 1, 1, nullptr);
    return TCSG_OK;
}

// Block extent along the axis a STREAM kernel's buffer window covers.
static uint32_t window_extent(const tcsg_desc &d, tcsg_dim3 block) { return d.dimension == 3 ? block.z : block.y; }
// thread rows whose footprint rows one lane's loads cover (more than its own with LDS staging)
static uint32_t load_rows(const tcsg_kernel *k, tcsg_dim3 block) { return k->stage.x ? block.y : 1u; }

static void fill_kargs(const tcsg_desc &d, const tcsg_geometry *g, tcsg_dim3 block, uint32_t brows, const void *in,
                       void *out, const void *mask, KArgs &a)
{
    a = KArgs{};
    a.in = in;
    a.out = out;
    a.mask = mask;
    a.xdim = g->size.x;
    a.ydim = g->size.y;
    a.zdim = g->size.z;
    a.right_offset = g->right_offset;
    a.in_bytes = a.out_bytes = g->n_points * elem_size(d);
    a.rec_limit = 0xFFFFFFFFu;
    a.oob = 0xFFFFFFFFu;
    StreamLimits lim;
    if (d.variant == TCSG_STREAM && stream_limits(d, g, window_extent(d, block), brows, &lim) == TCSG_OK) {
        a.rec_limit = lim.rec;
        a.oob = lim.oob;
    }
}

int tcsg_geometry_check(const tcsg_desc *desc, const tcsg_geometry *g)
{
    if (!desc || !g || !desc_valid(*desc)) return TCSG_ERR_ARG;
    return check_geom(*desc, g);
}

int tcsg_kernel_geometry(const tcsg_kernel *k, tcsg_dim3 size, tcsg_geometry *out)
{
    if (!k || !out) return TCSG_ERR_ARG;
    return geometry_compute(k->desc, size, *out, k->remainder);
}

int tcsg_kernel_steps_per_launch(const tcsg_kernel *k, int *out)
{
    if (!k || !out) return TCSG_ERR_ARG;
    *out = (int)k->tsteps;
    return TCSG_OK;
}

int tcsg_kernel_remainder(const tcsg_kernel *k, int *out)
{
    if (!k || !out) return TCSG_ERR_ARG;
    *out = k->remainder ? 1 : 0;
    return TCSG_OK;
}

int tcsg_launch_init(tcsg_kernel *k, const tcsg_geometry *g, int mode, uint64_t seed, void *a, void *stream)
{
    if (!k || !g || !a || mode < 0 || mode > 3) return TCSG_ERR_ARG;
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((g->n_points + 255) / 256, 65536);
    if (k->desc.dtype == TCSG_F32C)
        hipLaunchKernelGGL(init_kernel<float>, dim3(blocks), dim3(256), 0, s, (float *)a, g->size.x, g->size.y,
                           g->size.z, g->left_offset, g->right_offset, g->range, g->dimension, mode, seed);
    else
        hipLaunchKernelGGL(init_kernel<double>, dim3(blocks), dim3(256), 0, s, (double *)a, g->size.x,
                           g->size.y, g->size.z, g->left_offset, g->right_offset, g->range, g->dimension, mode,
                           seed);
    TCSG_HIP(hipGetLastError());
    return TCSG_OK;
}

static int launch_one(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks, const void *in,
                      void *out, hipStream_t s)
{
    KArgs a;
    fill_kargs(k->desc, g, block, load_rows(k, block), in, out, k->mask, a);
    size_t sz = sizeof(a);
    void *extra[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &a, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
    const hipError_t e = hipModuleLaunchKernel(k->fn, blocks.x, blocks.y, blocks.z, block.x, block.y, block.z, 0, s,
                                               nullptr, extra);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a rejected launch must not surface in the caller's next HIP call
        g_last_hip = (int)e;
        return TCSG_ERR_HIP;
    }
    return TCSG_OK;
}

static int precheck(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks)
{
    if (!k || !g) return TCSG_ERR_ARG;
    int rc = check_geom(k->desc, g, k->remainder);
    if (rc) return rc;
    rc = tcsg_check_block(k, block);
    if (rc) return rc;
    if (k->desc.variant == TCSG_STREAM) {
        rc = stream_limits(k->desc, g, window_extent(k->desc, block), load_rows(k, block), nullptr);
        if (rc) return rc;
    }
    // an LDS-staging kernel with remainder handling starts its last block of rows (planes)
    // early enough to end at the interior's end: the interior must hold one block's tile
    if (k->stage.x && k->remainder &&
        ((uint64_t)g->elements.y < (uint64_t)k->stage.y * k->desc.yc ||
         (k->desc.dimension == 3 && (uint64_t)g->elements.z < (uint64_t)k->stage.z * k->desc.zc)))
        return TCSG_ERR_ARG;
    if (!blocks.x || !blocks.y || !blocks.z) return TCSG_ERR_ARG;
    if ((uint64_t)blocks.x * block.x < g->threads.x || (uint64_t)blocks.y * block.y < g->threads.y ||
        (uint64_t)blocks.z * block.z < g->threads.z)
        return TCSG_ERR_ARG;  // grid would not cover the interior
    // the launch API's grid limits: 2^31-1 blocks in x, 65535 in y and z
    if (blocks.x > 0x7FFFFFFFu || blocks.y > 65535u || blocks.z > 65535u) return TCSG_ERR_BLOCK_SHAPE;
    if (k->desc.coeff == TCSG_VARIABLE && !k->mask_set) return TCSG_ERR_NO_MASK;
    return TCSG_OK;
}

int tcsg_update_grid(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks, const void *in,
                     void *out, void *stream)
{
    int rc = precheck(k, g, block, blocks);
    if (rc) return rc;
    if (!in || !out || in == out) return TCSG_ERR_ARG;
    return launch_one(k, g, block, blocks, in, out, (hipStream_t)stream);
}

int tcsg_run(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks, void *d0, void *d1,
             int steps, void *stream, int *final_index)
{
    int rc = precheck(k, g, block, blocks);
    if (rc) return rc;
    if (!d0 || !d1 || d0 == d1 || steps < 0) return TCSG_ERR_ARG;
    void *b[2] = {d0, d1};
    for (int s = 0; s < steps; ++s) {
        rc = launch_one(k, g, block, blocks, b[s & 1], b[(s + 1) & 1], (hipStream_t)stream);
        if (rc) return rc;
    }
    if (final_index) *final_index = steps & 1;
    return TCSG_OK;
}

int tcsg_benchmark_update_grid(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks,
                               double max_time, void *d0, void *d1, double *durations, size_t cap, size_t *n)
{
    int rc = precheck(k, g, block, blocks);
    if (rc) return rc;
    if (!d0 || !d1 || d0 == d1 || max_time < 0) return TCSG_ERR_ARG;
    using clk = std::chrono::steady_clock;
    std::vector<clk::time_point> ts;
    ts.reserve(1024);
    ts.push_back(clk::now());
    double runtime = 0.0;
    void *a0 = d0, *a1 = d1;
    for (unsigned long long it = 1; runtime <= max_time; it *= 2) {
        for (unsigned long long i = 0; i < it; ++i) {
            rc = launch_one(k, g, block, blocks, a0, a1, nullptr);
            if (rc) return rc;
            TCSG_HIP(hipDeviceSynchronize());
            std::swap(a0, a1);
            ts.push_back(clk::now());
        }
        runtime = std::chrono::duration<double>(ts.back() - ts.front()).count();
    }
    // per-step durations: a launch of a temporal kernel advances tsteps steps, each given an
    // equal share of the launch's time
    const size_t launches = ts.size() - 1, K = k->tsteps, cnt = launches * K;
    for (size_t i = 0; i < cnt && i < cap; ++i)
        durations[i] = std::chrono::duration<double>(ts[i / K + 1] - ts[i / K]).count() / (double)K;
    if (n) *n = cnt;
    return TCSG_OK;
}

int tcsg_time_steps(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks, void *d0,
                    void *d1, int warmup, int steps, void *stream, float *ms)
{
    int rc = precheck(k, g, block, blocks);
    if (rc) return rc;
    if (!d0 || !d1 || d0 == d1 || warmup < 0 || steps < 1 || !ms) return TCSG_ERR_ARG;
    const hipStream_t s = (hipStream_t)stream;
    void *b[2] = {d0, d1};
    int step = 0;
    for (int w = 0; w < warmup; ++w, ++step) {
        rc = launch_one(k, g, block, blocks, b[step & 1], b[(step + 1) & 1], s);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    TCSG_HIP(hipEventCreate(&e0));
    TCSG_HIP(hipEventCreate(&e1));
    TCSG_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < steps; ++i, ++step) {
        rc = launch_one(k, g, block, blocks, b[step & 1], b[(step + 1) & 1], s);
        if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
    }
    TCSG_HIP(hipEventRecord(e1, s));
    TCSG_HIP(hipEventSynchronize(e1));
    TCSG_HIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return TCSG_OK;
}

} // extern "C"

namespace {
struct Tile {
    uint32_t grid, bx, by, bz;
};
struct BatchArgs {
    const KArgs *grids;
    const Tile *tiles;
    uint32_t ntiles, pad;
    uint32_t *queue;  // the kernel's work queue: 8 per-XCD range counters + finished workgroups
};
static_assert(sizeof(BatchArgs) == 32, "batch argument layout");
constexpr size_t kQueueWords = 16;
} // namespace

struct tcsg_batch {
    tcsg_kernel *k = nullptr;
    hipFunction_t fn = nullptr;
    std::vector<tcsg_geometry> geoms;
    std::vector<tcsg_dim3> blocks;
    std::vector<void *> b0, b1;
    tcsg_dim3 block{};
    uint32_t ntiles = 0, nwg = 0;
    uint64_t points = 0;
    KArgs *d_grids[2] = {nullptr, nullptr};
    Tile *d_tiles = nullptr;
    uint32_t *d_queue = nullptr;  // zero between launches (the kernel's last workgroup clears it)
    hipStream_t own = nullptr;
    std::vector<std::pair<int, hipGraphExec_t>> graphs;
};

extern "C" {

int tcsg_batch_free(tcsg_batch *b)
{
    if (!b) return TCSG_OK;
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto *p : b->d_grids)
        if (p) (void)hipFree(p);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_queue) (void)hipFree(b->d_queue);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
    return TCSG_OK;
}

int tcsg_batch_create(tcsg_batch **out, tcsg_kernel *k, uint32_t n, const tcsg_geometry *geoms, void *const *bufs0,
                      void *const *bufs1, tcsg_dim3 block, uint32_t workgroups)
{
    if (!out || !k || !n || !geoms || !bufs0 || !bufs1) return TCSG_ERR_ARG;
    *out = nullptr;
    int rc = tcsg_check_block(k, block);
    if (rc) return rc;
    if (k->desc.coeff == TCSG_VARIABLE && !k->mask_set) return TCSG_ERR_NO_MASK;
    auto *b = new tcsg_batch();
    b->k = k;
    b->block = block;
    if (hipModuleGetFunction(&b->fn, k->module, "tcsg_update_batched") != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return TCSG_ERR_MISMATCH;  // code object generated without the batched kernel
    }
    std::vector<KArgs> args[2];
    std::vector<Tile> tiles;
    for (uint32_t i = 0; i < n; ++i) {
        rc = check_geom(k->desc, &geoms[i], k->remainder);
        if (!rc && k->desc.variant == TCSG_STREAM)
            rc = stream_limits(k->desc, &geoms[i], window_extent(k->desc, block), load_rows(k, block), nullptr);
        if (!rc && (!bufs0[i] || !bufs1[i] || bufs0[i] == bufs1[i])) rc = TCSG_ERR_ARG;
        if (rc) { delete b; return rc; }
        const tcsg_dim3 nb = tcsg_grid_blocks(&geoms[i], block);
        b->geoms.push_back(geoms[i]);
        b->blocks.push_back(nb);
        b->b0.push_back(bufs0[i]);
        b->b1.push_back(bufs1[i]);
        b->points += (uint64_t)geoms[i].elements.x * geoms[i].elements.y * geoms[i].elements.z;
        for (int p = 0; p < 2; ++p) {
            KArgs a;
            fill_kargs(k->desc, &geoms[i], block, load_rows(k, block), p ? bufs1[i] : bufs0[i],
                       p ? bufs0[i] : bufs1[i], k->mask, a);
            args[p].push_back(a);
        }
        if (k->pack) {
            // runs of block.x * block.y threads over the grid's xy thread space, per block.z planes
            const uint64_t nxy = (uint64_t)geoms[i].threads.x * geoms[i].threads.y;
            const uint64_t per = (uint64_t)block.x * block.y;
            // the kernel numbers a grid's xy threads in 32 bits
            if (nxy + per > 0xFFFFFFFFull) { delete b; return TCSG_ERR_TOO_LARGE; }
            const uint32_t nr = (uint32_t)((nxy + per - 1) / per);
            for (uint32_t z = 0; z < nb.z; ++z)
                for (uint32_t x = 0; x < nr; ++x) tiles.push_back({i, x, 0, z});
        } else {
            for (uint32_t z = 0; z < nb.z; ++z)
                for (uint32_t y = 0; y < nb.y; ++y)
                    for (uint32_t x = 0; x < nb.x; ++x) tiles.push_back({i, x, y, z});
        }
    }
    b->ntiles = (uint32_t)tiles.size();
    uint32_t nwg = workgroups;
    if (!nwg) {
        // persistent grid sized to fill the device: as many workgroups per CU as
        // the batched kernel's registers allow at this block size
        int dev = 0, cus = 256, per_cu = 0;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const int threads = (int)(block.x * block.y * block.z);
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->fn, threads, 0) != hipSuccess) {
            (void)hipGetLastError();
            per_cu = 0;
        }
        nwg = (uint32_t)cus * (uint32_t)std::max(1, per_cu);
    }
    b->nwg = std::min(nwg, b->ntiles);
    hipError_t e = hipSuccess;
    for (int p = 0; p < 2 && e == hipSuccess; ++p) {
        e = hipMalloc(&b->d_grids[p], n * sizeof(KArgs));
        if (e == hipSuccess) e = hipMemcpy(b->d_grids[p], args[p].data(), n * sizeof(KArgs), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc(&b->d_tiles, tiles.size() * sizeof(Tile));
    if (e == hipSuccess) e = hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&b->d_queue, kQueueWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b->d_queue, 0, kQueueWords * sizeof(uint32_t));  // before any capture
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_last_hip = (int)e;
        tcsg_batch_free(b);
        return TCSG_ERR_HIP;
    }
    *out = b;
    return TCSG_OK;
}

static int batch_launch(tcsg_batch *b, int parity, hipStream_t s)
{
    BatchArgs a{b->d_grids[parity], b->d_tiles, b->ntiles, 0, b->d_queue};
    size_t sz = sizeof(a);
    void *extra[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &a, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
    TCSG_HIP(hipModuleLaunchKernel(b->fn, b->nwg, 1, 1, b->block.x, b->block.y, b->block.z, 0, s, nullptr, extra));
    return TCSG_OK;
}

int tcsg_batch_run(tcsg_batch *b, int steps, int use_graph, void *stream, int *final_index)
{
    if (!b || steps < 0) return TCSG_ERR_ARG;
    const hipStream_t s = stream ? (hipStream_t)stream : b->own;
    if (!use_graph) {
        for (int t = 0; t < steps; ++t) {
            const int rc = batch_launch(b, t & 1, s);
            if (rc) return rc;
        }
    } else if (steps > 0) {
        hipGraphExec_t exec = nullptr;
        for (auto &g : b->graphs)
            if (g.first == steps) exec = g.second;
        if (!exec) {
            // capture on the library's stream, then replay on the caller's
            TCSG_HIP(hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal));
            int rc = TCSG_OK;
            for (int t = 0; t < steps && !rc; ++t) rc = batch_launch(b, t & 1, b->own);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(b->own, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            TCSG_HIP(e);
            const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            TCSG_HIP(e2);
            b->graphs.emplace_back(steps, exec);
        }
        TCSG_HIP(hipGraphLaunch(exec, s));
    }
    if (!stream) TCSG_HIP(hipStreamSynchronize(b->own));
    if (final_index) *final_index = steps & 1;
    return TCSG_OK;
}

int tcsg_batch_run_separate(tcsg_batch *b, int steps, void *stream, int *final_index)
{
    if (!b || steps < 0) return TCSG_ERR_ARG;
    const hipStream_t s = stream ? (hipStream_t)stream : b->own;
    for (int t = 0; t < steps; ++t)
        for (size_t i = 0; i < b->geoms.size(); ++i) {
            void *in = (t & 1) ? b->b1[i] : b->b0[i];
            void *out = (t & 1) ? b->b0[i] : b->b1[i];
            const int rc = launch_one(b->k, &b->geoms[i], b->block, b->blocks[i], in, out, s);
            if (rc) return rc;
        }
    if (!stream) TCSG_HIP(hipStreamSynchronize(b->own));
    if (final_index) *final_index = steps & 1;
    return TCSG_OK;
}

int tcsg_batch_info(const tcsg_batch *b, uint64_t *points, uint32_t *tiles, uint32_t *workgroups)
{
    if (!b) return TCSG_ERR_ARG;
    if (points) *points = b->points;
    if (tiles) *tiles = b->ntiles;
    if (workgroups) *workgroups = b->nwg;
    return TCSG_OK;
}

static GridIndex grid_index(const tcsg_geometry &g)
{
    const uint64_t r = g.range, zr = g.dimension == 3 ? g.range : 0;
    return GridIndex{g.size.x, g.size.y, (g.left_offset + r) + (uint64_t)g.size.x * (r + (uint64_t)g.size.y * zr)};
}

// Reduction scratch of tcsg_error_f32_vs_f64: 5 doubles per device, allocated by the first
// call on that device and kept until the process ends (calls that overlap on one device race).
static std::mutex g_err_mu;
static double *g_err_scratch[256] = {};

static hipError_t error_scratch(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 256) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_err_mu);
    if (!g_err_scratch[dev]) {
        e = hipMalloc(&g_err_scratch[dev], 5 * sizeof(double));
        if (e != hipSuccess) { g_err_scratch[dev] = nullptr; return e; }
    }
    *out = g_err_scratch[dev];
    return hipSuccess;
}

// Clears the 5 partial results, reduces, copies them to the host (null stream, synchronous).
static int error_reduce(const tcsg_geometry *ga, const float *a, const tcsg_geometry *gb, const double *b,
                        double floor_v, double *d, double *out)
{
    hipError_t e = hipMemsetAsync(d, 0, 5 * sizeof(double), nullptr);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(error_kernel, dim3(1024), dim3(256), 0, nullptr, a, grid_index(*ga), b, grid_index(*gb),
                           (uint64_t)ga->elements.x, (uint64_t)ga->elements.y, (uint64_t)ga->elements.z, floor_v, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d, 5 * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { g_last_hip = (int)e; return TCSG_ERR_HIP; }
    return TCSG_OK;
}

static bool same_interior(const tcsg_geometry *ga, const tcsg_geometry *gb)
{
    return ga->dimension == gb->dimension && ga->range == gb->range && ga->elements.x == gb->elements.x &&
           ga->elements.y == gb->elements.y && ga->elements.z == gb->elements.z;
}

int tcsg_error_f32_vs_f64(const tcsg_geometry *ga, const float *a, const tcsg_geometry *gb, const double *b,
                          double floor_v, double *out)
{
    if (!ga || !gb || !a || !b || !out) return TCSG_ERR_ARG;
    if (!same_interior(ga, gb)) return TCSG_ERR_ARG;  // the two states must cover the same interior
    double *d = nullptr;
    TCSG_HIP(error_scratch(&d));
    return error_reduce(ga, a, gb, b, floor_v, d, out);
}

// ---- the error reduction without allocation (no counterpart in the reference) ----
// Partial results of the two-pass reduction: one row of 5 per workgroup of the first
// pass, in a file-scope device array (kErrBlocks rows, 640 KiB on each GPU the library
// runs on, held until the process ends).  Each first-pass workgroup reduces a contiguous
// range of at most kErrPointsPerBlock interior points, so interiors up to
// kErrBlocks * kErrPointsPerBlock = 2^30 points fit; larger ones are refused.
constexpr uint32_t kErrBlocks = 16384;
constexpr uint32_t kErrThreads = 256;
constexpr uint64_t kErrPointsPerBlock = 65536;
static __device__ __align__(256) double scratch[kErrBlocks * 5];

// first pass: workgroup w reduces points [w*chunk, min((w+1)*chunk, n)) in a fixed order
// (each thread its strided points in turn, then a fixed tree over the threads) into row w
__global__ void __launch_bounds__(kErrThreads) error_partial_kernel(const float *__restrict__ a, GridIndex ga,
                                                                    const double *__restrict__ b, GridIndex gb,
                                                                    uint64_t nx, uint64_t ny, uint64_t nz,
                                                                    uint64_t chunk, double floor_v, double *part)
{
    const uint64_t n = nx * ny * nz;
    const uint64_t p0 = blockIdx.x * chunk, p1 = p0 + chunk < n ? p0 + chunk : n;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint64_t t = p0 + threadIdx.x; t < p1; t += kErrThreads) {
        const uint64_t i = t % nx, jk = t / nx, j = jk % ny, k = jk / ny;
        const double bv = b[gb.at(i, j, k)];
        const double d = fabs((double)a[ga.at(i, j, k)] - bv);
        v[0] = fmax(v[0], d);
        v[1] = fmax(v[1], fabs(bv));
        v[2] += d * d;
        v[3] += bv * bv;
        v[4] = fmax(v[4], d / fmax(fabs(bv), floor_v));
    }
    __shared__ double red[5][kErrThreads];
    for (int c = 0; c < 5; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (uint32_t h = kErrThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            for (int c = 0; c < 5; ++c) {
                const double o = red[c][threadIdx.x + h];
                red[c][threadIdx.x] = (c == 2 || c == 3) ? red[c][threadIdx.x] + o : fmax(red[c][threadIdx.x], o);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) part[5 * (uint64_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// second pass: one workgroup folds the rows [0, rows) in a fixed order into out[5]
__global__ void __launch_bounds__(kErrThreads) error_final_kernel(const double *__restrict__ part, uint32_t rows,
                                                                  double *__restrict__ out)
{
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t w = threadIdx.x; w < rows; w += kErrThreads) {
        for (int c = 0; c < 5; ++c) {
            const double o = part[5 * (uint64_t)w + c];
            v[c] = (c == 2 || c == 3) ? v[c] + o : fmax(v[c], o);
        }
    }
    __shared__ double red[5][kErrThreads];
    for (int c = 0; c < 5; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (uint32_t h = kErrThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            for (int c = 0; c < 5; ++c) {
                const double o = red[c][threadIdx.x + h];
                red[c][threadIdx.x] = (c == 2 || c == 3) ? red[c][threadIdx.x] + o : fmax(red[c][threadIdx.x], o);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) out[threadIdx.x] = red[threadIdx.x][0];
}

// Device address of `scratch` on each device, looked up on the first call on that device
// (with it current); host code uses these addresses, never the array's name.
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[256] = {};

static hipError_t scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 256) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (double *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

int tcsg_error_f32_vs_f64_noalloc(const tcsg_geometry *ga, const float *a, const tcsg_geometry *gb,
                                  const double *b, double floor_v, double *out, void *stream)
{
    if (!ga || !gb || !a || !b || !out) return TCSG_ERR_ARG;
    if (!same_interior(ga, gb)) return TCSG_ERR_ARG;  // the two states must cover the same interior
    const uint64_t n = (uint64_t)ga->elements.x * ga->elements.y * ga->elements.z;
    if (n == 0) return TCSG_ERR_ARG;
    if (n > (uint64_t)kErrBlocks * kErrPointsPerBlock) return TCSG_ERR_TOO_LARGE;  // beyond the scratch rows
    double *part = nullptr;
    TCSG_HIP(scratch_address(&part));
    const uint64_t chunk = std::max<uint64_t>(kErrThreads, (n + kErrBlocks - 1) / kErrBlocks);
    const uint32_t rows = (uint32_t)((n + chunk - 1) / chunk);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(error_partial_kernel, dim3(rows), dim3(kErrThreads), 0, s, a, grid_index(*ga), b,
                       grid_index(*gb), (uint64_t)ga->elements.x, (uint64_t)ga->elements.y,
                       (uint64_t)ga->elements.z, chunk, floor_v, part);
    TCSG_HIP(hipGetLastError());
    hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(kErrThreads), 0, s, (const double *)part, rows, out);
    TCSG_HIP(hipGetLastError());
    return TCSG_OK;
}

int tcsg_convert_f64_to_f32(const double *src, float *dst, uint64_t n, void *stream)
{
    if (!src || !dst) return TCSG_ERR_ARG;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(convert_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, (hipStream_t)stream, src, dst, n);
    TCSG_HIP(hipGetLastError());
    return TCSG_OK;
}

} // extern "C"
