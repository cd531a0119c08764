// HBM stream probes: plain copy, read-only and write-only kernels whose
// bandwidth is the practical ceiling the stencil kernels are measured
// against (a one-step stencil reads one grid and writes the other, so the
// copy rate bounds it).  16-byte accesses, each lane handling 1-16
// elements per iteration with all loads issued before the stores, optional
// nontemporal hints, 256-1024 threads per workgroup, grid-stride over the
// buffer.
#include <hip/hip_runtime.h>

#include "tcsg.h"

void tcsg_note_hip_error(int e);  // tcsg_runtime.hip

namespace {

typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

template <bool NT>
__device__ __forceinline__ u32x4 ld(const u32x4 *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool NT>
__device__ __forceinline__ void st(u32x4 *p, u32x4 v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// mode 0: dst = src; 1: read src (a lane stores only if its xor of
// everything it read equals `never`, which the host picks so it does not);
// 2: dst = constant.  UNROLL elements of 16 bytes per lane per iteration
// (UNROLL * 16 bytes in flight per lane), BLOCK threads per workgroup.
template <int MODE, bool NT, int UNROLL, int BLOCK>
__global__ void __launch_bounds__(BLOCK) probe_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                      uint64_t n, uint32_t never)
{
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK * UNROLL;
    u32x4 acc = {0u, 0u, 0u, 0u};
    for (uint64_t base = (uint64_t)blockIdx.x * BLOCK * UNROLL + threadIdx.x; base < n; base += stride) {
        u32x4 v[UNROLL];
        if constexpr (MODE != 2) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const uint64_t i = base + (uint64_t)u * BLOCK;
                if (i < n) v[u] = ld<NT>(src + i);
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const uint64_t i = base + (uint64_t)u * BLOCK;
            if (i >= n) continue;
            if constexpr (MODE == 0) st<NT>(dst + i, v[u]);
            else if constexpr (MODE == 1) acc ^= v[u];
            else st<NT>(dst + i, u32x4{(uint32_t)i, 1u, 2u, 3u});
        }
    }
    if constexpr (MODE == 1) {
        if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == never) dst[blockIdx.x * BLOCK + threadIdx.x] = acc;
    }
}

template <int MODE, int UNROLL, int BLOCK>
hipError_t launch3(int nt, uint32_t wg, hipStream_t s, const u32x4 *src, u32x4 *dst, uint64_t n)
{
    if (nt) hipLaunchKernelGGL((probe_kernel<MODE, true, UNROLL, BLOCK>), dim3(wg), dim3(BLOCK), 0, s, src, dst, n,
                               0x9E3779B9u);
    else hipLaunchKernelGGL((probe_kernel<MODE, false, UNROLL, BLOCK>), dim3(wg), dim3(BLOCK), 0, s, src, dst, n,
                            0x9E3779B9u);
    return hipGetLastError();
}

template <int MODE, int UNROLL>
hipError_t launch2(int nt, uint32_t block, uint32_t wg, hipStream_t s, const u32x4 *src, u32x4 *dst, uint64_t n)
{
    if (block == 256) return launch3<MODE, UNROLL, 256>(nt, wg, s, src, dst, n);
    if (block == 512) return launch3<MODE, UNROLL, 512>(nt, wg, s, src, dst, n);
    return launch3<MODE, UNROLL, 1024>(nt, wg, s, src, dst, n);
}

template <int MODE>
hipError_t launch(int nt, uint32_t unroll, uint32_t block, uint32_t wg, hipStream_t s, const u32x4 *src, u32x4 *dst,
                  uint64_t n)
{
    switch (unroll) {
    case 1: return launch2<MODE, 1>(nt, block, wg, s, src, dst, n);
    case 2: return launch2<MODE, 2>(nt, block, wg, s, src, dst, n);
    case 4: return launch2<MODE, 4>(nt, block, wg, s, src, dst, n);
    case 8: return launch2<MODE, 8>(nt, block, wg, s, src, dst, n);
    default: return launch2<MODE, 16>(nt, block, wg, s, src, dst, n);
    }
}

}  // namespace

extern "C" int tcsg_hbm_probe_ex(int mode, int nontemporal, uint32_t unroll, uint32_t block, const void *src,
                                 void *dst, uint64_t bytes, uint32_t workgroups, int reps, void *stream,
                                 float *ms_per_rep)
{
    if (mode < 0 || mode > 2 || reps < 1 || !ms_per_rep || bytes < 16 || bytes % 16) return TCSG_ERR_ARG;
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8 && unroll != 16) return TCSG_ERR_ARG;
    if (block != 256 && block != 512 && block != 1024) return TCSG_ERR_ARG;
    if ((mode != 2 && !src) || !dst) return TCSG_ERR_ARG;
    const uint64_t n = bytes / 16;
    // read mode stores at most one element per launched lane
    uint32_t wg = workgroups;
    if (!wg) wg = (uint32_t)((n + (uint64_t)block * unroll - 1) / ((uint64_t)block * unroll));
    if (mode == 1 && (uint64_t)wg * block > n) return TCSG_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto fail = [&](hipError_t e) {
        tcsg_note_hip_error((int)e);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return TCSG_ERR_HIP;
    };
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return fail(e);
    const u32x4 *sp = (const u32x4 *)src;
    u32x4 *dp = (u32x4 *)dst;
    auto one = [&]() {
        if (mode == 0) return launch<0>(nontemporal, unroll, block, wg, s, sp, dp, n);
        if (mode == 1) return launch<1>(nontemporal, unroll, block, wg, s, sp, dp, n);
        return launch<2>(nontemporal, unroll, block, wg, s, sp, dp, n);
    };
    if ((e = one()) != hipSuccess) return fail(e);  // warm-up
    if ((e = hipEventRecord(e0, s)) != hipSuccess) return fail(e);
    for (int r = 0; r < reps; ++r)
        if ((e = one()) != hipSuccess) return fail(e);
    if ((e = hipEventRecord(e1, s)) != hipSuccess) return fail(e);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return fail(e);
    float ms = 0.0f;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return fail(e);
    *ms_per_rep = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return TCSG_OK;
}

extern "C" int tcsg_hbm_probe(int mode, int nontemporal, const void *src, void *dst, uint64_t bytes,
                              uint32_t workgroups, int reps, void *stream, float *ms_per_rep)
{
    return tcsg_hbm_probe_ex(mode, nontemporal, 4, 256, src, dst, bytes, workgroups, reps, stream, ms_per_rep);
}
