// Probe of the raw-buffer range check that generated STREAM kernels rely on
// (buffer resource with stride 0): every lane loads from a buffer whose record
// count ends inside the probed range, while the allocation behind it is larger,
// so every address stays inside the allocation whatever the check does.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef int tcsg_v4i __attribute__((ext_vector_type(4)));

// dword load at per-lane offset vbase + 4*lane plus SGPR offset soff
extern "C" __global__ void tcsg_probe_b32(const uint32_t *p, uint32_t records, uint32_t vbase, uint32_t soff,
                                          uint32_t *out)
{
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)records, 0x00020000);
    out[threadIdx.x] = __builtin_amdgcn_raw_buffer_load_b32(r, vbase + 4u * threadIdx.x, soff, 0);
}

// 16-byte load at per-lane offset vbase + 16*lane plus SGPR offset soff
extern "C" __global__ void tcsg_probe_b128(const uint32_t *p, uint32_t records, uint32_t vbase, uint32_t soff,
                                           uint32_t *out)
{
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)records, 0x00020000);
    const tcsg_v4i v = __builtin_amdgcn_raw_buffer_load_b128(r, vbase + 16u * threadIdx.x, soff, 0);
    out[4 * threadIdx.x + 0] = v.x;
    out[4 * threadIdx.x + 1] = v.y;
    out[4 * threadIdx.x + 2] = v.z;
    out[4 * threadIdx.x + 3] = v.w;
}
