// HBM stream probes: plain copy, read-only and write-only kernels whose
// bandwidth is the practical ceiling the stencil kernels are measured
// against (a one-step stencil reads one grid and writes the other, so the
// copy rate bounds it).  16-byte accesses, each lane handling UNROLL
// elements per iteration with all loads issued before the stores, optional
// nontemporal hints, grid-stride over the buffer.
#include <hip/hip_runtime.h>

#include "tcsg.h"

void tcsg_note_hip_error(int e);  // tcsg_runtime.hip

namespace {

constexpr int kUnroll = 4;
constexpr uint32_t kBlock = 256;

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
// 2: dst = constant.
template <int MODE, bool NT>
__global__ void __launch_bounds__(kBlock) probe_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                       uint64_t n, uint32_t never)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock * kUnroll;
    u32x4 acc = {0u, 0u, 0u, 0u};
    for (uint64_t base = (uint64_t)blockIdx.x * kBlock * kUnroll + threadIdx.x; base < n; base += stride) {
        u32x4 v[kUnroll];
        if constexpr (MODE != 2) {
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint64_t i = base + (uint64_t)u * kBlock;
                if (i < n) v[u] = ld<NT>(src + i);
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint64_t i = base + (uint64_t)u * kBlock;
            if (i >= n) continue;
            if constexpr (MODE == 0) st<NT>(dst + i, v[u]);
            else if constexpr (MODE == 1) acc ^= v[u];
            else st<NT>(dst + i, u32x4{(uint32_t)i, 1u, 2u, 3u});
        }
    }
    if constexpr (MODE == 1) {
        if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == never) dst[blockIdx.x * kBlock + threadIdx.x] = acc;
    }
}

template <int MODE>
hipError_t launch(int nt, uint32_t wg, hipStream_t s, const u32x4 *src, u32x4 *dst, uint64_t n)
{
    if (nt) hipLaunchKernelGGL((probe_kernel<MODE, true>), dim3(wg), dim3(kBlock), 0, s, src, dst, n, 0x9E3779B9u);
    else hipLaunchKernelGGL((probe_kernel<MODE, false>), dim3(wg), dim3(kBlock), 0, s, src, dst, n, 0x9E3779B9u);
    return hipGetLastError();
}

}  // namespace

extern "C" int tcsg_hbm_probe(int mode, int nontemporal, const void *src, void *dst, uint64_t bytes,
                              uint32_t workgroups, int reps, void *stream, float *ms_per_rep)
{
    if (mode < 0 || mode > 2 || reps < 1 || !ms_per_rep || bytes < 16 || bytes % 16) return TCSG_ERR_ARG;
    if ((mode != 2 && !src) || !dst) return TCSG_ERR_ARG;
    const uint64_t n = bytes / 16;
    // read mode stores at most one element per launched lane
    uint32_t wg = workgroups;
    if (!wg) wg = (uint32_t)((n + (uint64_t)kBlock * kUnroll - 1) / ((uint64_t)kBlock * kUnroll));
    if (mode == 1 && (uint64_t)wg * kBlock > n) return TCSG_ERR_ARG;
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
        if (mode == 0) return launch<0>(nontemporal, wg, s, sp, dp, n);
        if (mode == 1) return launch<1>(nontemporal, wg, s, sp, dp, n);
        return launch<2>(nontemporal, wg, s, sp, dp, n);
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
