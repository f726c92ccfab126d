// Device-side pieces shared by the GPU refit and the GPU build: the host
// builder's min/max and the quantization of one BVH4 node from its children
// (crates/asset/src/near_field.rs:402-430 child_bounds, :481-525 quantized_node).
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace rq {
namespace bvhdev {

// Keep the first argument unless the second is strictly smaller / larger, as
// the host builder does: the device must round and order like the build.
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ float bf(uint32_t w) { return __uint_as_float(w); }
__device__ __forceinline__ uint32_t fb(float f) { return __float_as_uint(f); }
// f32 total order as a signed integer (atomic min/max of floats).
__device__ __forceinline__ int32_t okey(float f) {
    const int32_t k = __float_as_int(f);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}
__device__ __forceinline__ float okey_inv(int32_t k) { return __int_as_float(k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1)); }

// Recomputes node n of a mesh in `words` (arrays at word offsets nodes, leaves,
// positions, indices): each child's box -- a leaf's corners, or the union of
// both decodes of a child node, which must be final already -- then the
// origin, steps and outward 8-bit codes quantized_node gives them.
// A leaf's box: its triangles' corners folded in index order (the host's order,
// so even the sign of a zero bound matches).
__device__ inline void leaf_box(const uint32_t* words, uint32_t leaves, uint32_t positions, uint32_t indices,
                                uint32_t l, float lo[3], float hi[3]) {
    for (int x = 0; x < 3; ++x) {
        lo[x] = __builtin_inff();
        hi[x] = -__builtin_inff();
    }
    const uint32_t off = words[leaves + 2ull * l], cnt = words[leaves + 2ull * l + 1];
    for (uint32_t i = off; i < off + cnt; ++i) {
        const uint32_t* p = words + positions + 3ull * words[indices + i];
        for (int x = 0; x < 3; ++x) {
            lo[x] = mn(lo[x], bf(p[x]));
            hi[x] = mx(hi[x], bf(p[x]));
        }
    }
}

// `boxes`: the mesh's leaf boxes (6 floats a leaf), or null to fold the triangles here.
__device__ inline void refit_node(uint32_t* words, uint32_t nodes, uint32_t leaves, uint32_t positions,
                                  uint32_t indices, uint32_t n, const float* boxes = nullptr) {
    uint32_t* node = words + nodes + 16ull * n;
    uint32_t children[4];
    float blo[4][3], bhi[4][3];
    uint32_t count = 0;
    for (int c = 0; c < 4; ++c) {
        const uint32_t child = node[6 + c];
        if (child == EMPTY) continue;
        float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        if (child & LEAF_BIT) {
            const uint32_t l = child & ~LEAF_BIT;
            if (boxes) {
                for (int x = 0; x < 3; ++x) {
                    lo[x] = boxes[6ull * l + x];
                    hi[x] = boxes[6ull * l + 3 + x];
                }
            } else {
                leaf_box(words, leaves, positions, indices, l, lo, hi);
            }
        } else {
            const uint32_t* cn = words + nodes + 16ull * child;
            for (int slot = 0; slot < 4; ++slot) {
                if (cn[6 + slot] == EMPTY) continue;
                for (int fused = 0; fused < 2; ++fused)
                    for (int x = 0; x < 3; ++x) {
                        const float o = bf(cn[x]), s = bf(cn[3 + x]);
                        const float ql = (float)((cn[10 + x] >> (8 * slot)) & 0xffu);
                        const float qh = (float)((cn[13 + x] >> (8 * slot)) & 0xffu);
                        const float l = fused ? __builtin_fmaf(s, ql, o) : o + s * ql;
                        const float h = fused ? __builtin_fmaf(s, qh, o) : o + s * qh;
                        lo[x] = mn(lo[x], l);
                        hi[x] = mx(hi[x], h);
                    }
            }
        }
        children[count] = child;
        for (int x = 0; x < 3; ++x) {
            blo[count][x] = lo[x];
            bhi[count][x] = hi[x];
        }
        ++count;
    }
    // quantized_node
    float all_lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float all_hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint32_t c = 0; c < count; ++c)
        for (int x = 0; x < 3; ++x) {
            all_lo[x] = mn(all_lo[x], blo[c][x]);
            all_hi[x] = mx(all_hi[x], bhi[c][x]);
        }
    uint32_t out[16];
    float step[3];
    const float grow = 1.0f + 1.0f / 1048576.0f;
    for (int x = 0; x < 3; ++x) {
        const float extent = all_hi[x] - all_lo[x];
        float s = mx(extent / 255.0f, mx(1.17549435e-38f, fabsf(all_lo[x]) * 1e-7f));
        // grow until code 255 covers the top in both decodes (bounded: s grows
        // by 2^-20 relative per pass and starts within a few ulps of the answer)
        for (int guard = 0; guard < (1 << 22) && (all_lo[x] + s * 255.0f < all_hi[x] ||
                                                   __builtin_fmaf(s, 255.0f, all_lo[x]) < all_hi[x]);
             ++guard)
            s *= grow;
        step[x] = s;
        out[x] = fb(all_lo[x]);
        out[3 + x] = fb(s);
    }
    for (int c = 0; c < 4; ++c) out[6 + c] = c < (int)count ? children[c] : EMPTY;
    for (int x = 0; x < 3; ++x) {
        uint32_t lw = 0, hw = 0;
        const float o = all_lo[x], s = step[x];
        for (uint32_t c = 0; c < count; ++c) {
            const float lv = blo[c][x], hv = bhi[c][x];
            float fl = floorf((lv - o) / s);
            fl = mn(mx(fl, 0.0f), 255.0f);
            uint32_t low = (uint32_t)fl;
            while (low > 0 && (o + s * (float)low > lv || __builtin_fmaf(s, (float)low, o) > lv)) low -= 1;
            float fh = ceilf((hv - o) / s);
            fh = mn(mx(fh, 0.0f), 255.0f);
            uint32_t high = (uint32_t)fh;
            while (high < 255 && (o + s * (float)high < hv || __builtin_fmaf(s, (float)high, o) < hv)) high += 1;
            lw |= low << (8 * c);
            hw |= high << (8 * c);
        }
        out[10 + x] = lw;
        out[13 + x] = hw;
    }
    for (int w = 0; w < 16; ++w) node[w] = out[w];
}

}  // namespace bvhdev
}  // namespace rq
