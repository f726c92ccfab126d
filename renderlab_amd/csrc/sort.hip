// Ray reordering for incoherent batches: a 30-bit key per ray (direction
// octant, then a 9-bit-per-axis Morton code of the origin inside the scene's
// bounds) sorted with rocPRIM's radix sort. The traversal then
// takes rays in key order, so a wave's lanes start from nearby origins in the
// same octant and stay in step longer; every ray's result is still written
// to its own index.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.h"

namespace rq {
namespace sortk {

__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x1ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void key_kernel(const rq_ray* rays, uint32_t n, const uint32_t* count_ptr, float lx, float ly, float lz,
                           float sx, float sy, float sz, uint32_t* keys, uint32_t* vals) {
    const uint32_t count = count_ptr ? min(*count_ptr, n) : n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t key = 0xffffffffu;
        if (i < count) {
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            float4 a = rp[0], b = rp[1];
            auto q = [](float v, float l, float s) {
                float f = (v - l) * s;
                f = f != f ? 0.0f : fminf(fmaxf(f, 0.0f), 511.0f);
                return (uint32_t)f;
            };
            uint32_t m = spread3(q(a.x, lx, sx)) | (spread3(q(a.y, ly, sy)) << 1) | (spread3(q(a.z, lz, sz)) << 2);
            uint32_t oct = (b.x < 0.0f ? 1u : 0u) | (b.y < 0.0f ? 2u : 0u) | (b.z < 0.0f ? 4u : 0u);
            key = (oct << 27) | m;
        }
        keys[i] = key;
        vals[i] = i;
    }
}

}  // namespace sortk

// Fills `order` (n entries) with the ray indices in key order. `scratch` is
// grown as needed by the caller: query with scratch == nullptr.
hipError_t ray_order(const rq_ray* rays, uint32_t n, const uint32_t* count_ptr, const float bounds[6], uint32_t* keys_in,
                     uint32_t* keys_out, uint32_t* vals_in, uint32_t* order, void* temp, size_t& temp_bytes,
                     hipStream_t s) {
    if (!temp) {
        return rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in, keys_out, vals_in, order, (size_t)n, 0u, 30u, s);
    }
    float sx = 511.0f / fmaxf(bounds[3] - bounds[0], 1e-20f);
    float sy = 511.0f / fmaxf(bounds[4] - bounds[1], 1e-20f);
    float sz = 511.0f / fmaxf(bounds[5] - bounds[2], 1e-20f);
    unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(sortk::key_kernel, dim3(blocks), dim3(256), 0, s, rays, n, count_ptr, bounds[0], bounds[1],
                       bounds[2], sx, sy, sz, keys_in, vals_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, vals_in, order, (size_t)n, 0u, 30u, s);
}

}  // namespace rq
