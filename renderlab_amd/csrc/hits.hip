// The reference's buffer form of a hit and its hit decoders, over a traced
// batch's rq_hit records:
//
// - the 16-byte record vec4<u32>(placement + 1, primitive, bits(u), bits(v)),
//   zero on a miss (shaders/primary_rays.wgsl:19; reflection rays' RsHit.record,
//   shaders/reflections_software_bvh4.wgsl:42, 90);
// - gi_hit_identity (shaders/gi_rays_software.wgsl:297-309): instance word 24,
//   or for a merged world (word 24 zero) the identity + 1 of the last
//   (first position, identity + 1) pair of the table at word 25 (word 26 pairs)
//   whose first position is at most tri.x;
// - gi_hit_primitive (:310-313): tri.x when the instance carries a primitive
//   table (word 28 nonzero), else GI_NO_PRIMITIVE (shaders/gi_cells.wgsl:186);
// - gi_hit_placement (:314-316): instance word 29.
//
// One __host__ __device__ decode serves the host entry point and the kernel, so
// both give the same words. Reflection scenes (rl_geometry) carry the placement
// slot and primitive in the hit itself (rq_hit.slot, tri[0]) and no identity.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace rq {
namespace hits {

constexpr uint32_t NO_PRIMITIVE = 0xffffffffu;  // shaders/gi_cells.wgsl:186

// The merged-world table's entry for position p: the identity of the last pair
// whose first position is <= p. Pairs are in position order (checked when words
// are adopted), so the last such pair is found by bisection.
__host__ __device__ inline uint32_t map_identity(const uint32_t* g, uint64_t map, uint32_t pairs, uint32_t p) {
    uint32_t lo = 0, hi = pairs;  // answer: the count of pairs with first <= p
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (g[map + 2ull * mid] <= p)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo ? g[map + 2ull * (lo - 1) + 1] : 0u;
}

// out: [0..3] the record, [4] identity. `reflection`: rl_geometry hits.
__host__ __device__ inline void decode(const uint32_t* g, uint64_t n, bool reflection, const rq_hit& h, uint32_t out[5]) {
    out[0] = out[1] = out[2] = out[3] = out[4] = 0;
    if (!(h.flags & RQ_HIT_HIT)) return;
    uint32_t ub, vb;
    memcpy(&ub, &h.u, 4);
    memcpy(&vb, &h.v, 4);
    if (reflection) {
        out[0] = h.slot + 1u;
        out[1] = h.tri[0];
        out[2] = ub;
        out[3] = vb;
        return;
    }
    const uint32_t inst = h.inst;
    const uint64_t r = HEADER_WORDS + (uint64_t)inst * INSTANCE_WORDS;
    if (n < HEADER_WORDS || inst >= g[4] || r + INSTANCE_WORDS > n) return;  // no such record
    const bool opaque = inst < g[0];
    out[0] = g[r + 29] + 1u;
    out[1] = opaque && g[r + 28] != 0u ? h.tri[0] : NO_PRIMITIVE;
    out[2] = ub;
    out[3] = vb;
    if (!opaque) return;
    const uint32_t own = g[r + 24];
    out[4] = own != 0u ? own : map_identity(g, g[r + 25], g[r + 26], h.tri[0]);
}

__global__ void __launch_bounds__(256) records_kernel(const uint32_t* __restrict__ g, uint64_t n, int reflection,
                                                      const rq_hit* __restrict__ hits, uint32_t count,
                                                      uint4* __restrict__ records, uint32_t* __restrict__ identities) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t o[5];
    decode(g, n, reflection != 0, hits[i], o);
    if (records) records[i] = make_uint4(o[0], o[1], o[2], o[3]);
    if (identities) identities[i] = o[4];
}

rq_status fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RQ_ERR_OUT_OF_MEMORY : RQ_ERR_DEVICE;
}

}  // namespace hits
}  // namespace rq

using namespace rq;

extern "C" rq_status rq_gi_hit_decode(const rq_scene* scene, const rq_hit* hit, uint32_t record[4],
                                      uint32_t* identity) {
    if (!scene || !hit || !record) {
        set_error("rq_gi_hit_decode: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    const bool rs = scene->kind == RQ_SCENE_REFLECTION;
    if (!rs) {
        const rq_status st = sync_host_words(scene);
        if (st != RQ_OK) return st;
    }
    uint32_t o[5];
    hits::decode(scene->words.data(), scene->words.size(), rs, *hit, o);
    std::copy(o, o + 4, record);
    if (identity) *identity = o[4];
    return RQ_OK;
}

extern "C" rq_status rq_hit_records(const rq_scene* scene, const rq_hit* hits, uint32_t count, rq_hit_record* records,
                                    uint32_t* identities, void* stream) {
    if (!scene || (count && (!hits || (!records && !identities)))) {
        set_error("rq_hit_records: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    const bool rs = scene->kind == RQ_SCENE_REFLECTION;
    const int layout = rs ? RQ_LAYOUT_RS_BVH4Q8 : RQ_LAYOUT_REF_BVH4Q8;
    if (!rs && !scene->dev[layout].ready) {
        set_error("rq_hit_records: needs the REF layout uploaded");
        return RQ_ERR_INVALID_INPUT;
    }
    if (!count) return RQ_OK;
    const uint32_t* g = rs ? nullptr : reinterpret_cast<const uint32_t*>(scene->dev[layout].base);
    hipLaunchKernelGGL(hits::records_kernel, dim3((count + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), g, rs ? 0ull : (uint64_t)scene_words(scene), rs ? 1 : 0,
                       hits, count, reinterpret_cast<uint4*>(records), identities);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RQ_OK : hits::fail(e, "rq_hit_records launch");
}
