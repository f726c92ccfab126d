// Device side: the traversal kernels for the three layouts, ray generation,
// FP64 verification, and the C entry points that launch them.
//
// Traversal semantics are those of gi_traverse / gi_instance
// (shaders/gi_rays_software.wgsl:98-279): a top level of 4-wide float nodes
// (or the flat instance list when header word 1 is 0), per-instance
// world-box test against gi_limit, world-to-object rows, a 48-entry stack
// per level, child boxes sorted by entry distance and pushed far to near,
// Moller-Trumbore in leaf order, first hit returned in any-hit modes.
//
// MI355X mapping (gfx950, wave64):
//  - one lane per ray inside a persistent grid sized to the device's
//    occupancy; a wave refills its idle lanes from a global atomic ray cursor
//    (one atomic per refill, lane ranks by mbcnt), so finished rays stop
//    costing SIMD slots;
//  - the two per-level stacks share one per-lane stack: a ring of S entries in
//    LDS laid out [entry][lane] (bank-conflict free), spilling the oldest
//    entries to a per-lane global area when deeper than S;
//  - min/max are v_minimum3/v_maximum3 (IEEE 754-2019), division and square
//    root are correctly rounded, no contraction (-ffp-contract=off), so the
//    kernels compute the oracle's arithmetic bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "internal.h"

namespace rq {
namespace dev {

constexpr uint32_t NONE = 0xffffffffu;
constexpr float FMAXV = 3.4028235e38f;
// Per-lane stack words: a top-level and a BLAS bound of 48 each for the
// 4-wide layouts (the reference's; depth <= 15 pushes at most 1 + 3*15), of
// 112 each for CW8 (7 per level of at most 16; an 8-wide tree is never deeper
// than the 4-wide tree it is collapsed from, and the relayout rejects one
// deeper than 16 levels).
constexpr uint32_t SPILL_WORDS = TRACE_SPILL_WORDS;
constexpr uint32_t STACK_CAP = 48;
constexpr uint32_t CW8_LEVELS = 16;  // the CW8 relayout rejects deeper trees
constexpr uint32_t CW8_CAP = 7 * CW8_LEVELS;
constexpr uint32_t LEAF_MASK = (1u << 27) - 1u;

enum : int { REF = 0, PLAIN = 1, CW8 = 2, RS = 3, RSW = 4 };  // RSW: rl_geometry over CW8 nodes

__device__ __forceinline__ float fmn(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_elementwise_maximum(a, b); }
__device__ __forceinline__ float bf(uint32_t w) { return __uint_as_float(w); }
__device__ __forceinline__ uint32_t fb(float f) { return __float_as_uint(f); }

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float safe_inv1(float d) {
    float tiny = d >= 0.0f ? 1e-30f : -1e-30f;
    return 1.0f / (fabsf(d) < 1e-30f ? tiny : d);
}
__device__ __forceinline__ V3 safe_inverse(V3 d) { return mk(safe_inv1(d.x), safe_inv1(d.y), safe_inv1(d.z)); }
// The same clamp with the hardware reciprocal (1 ulp): for box tests that
// carry their own error slack (the CW8 layout), never for hit arithmetic.
// |d| is raised to 1e-30 keeping d's sign (a zero's too: the near and far
// planes are chosen by the reciprocal's sign, so either sign culls correctly).
__device__ __forceinline__ float fast_inv1(float d) {
    return __builtin_amdgcn_rcpf(copysignf(fmaxf(fabsf(d), 1e-30f), d));
}
__device__ __forceinline__ V3 fast_inverse(V3 d) { return mk(fast_inv1(d.x), fast_inv1(d.y), fast_inv1(d.z)); }

// A box's exit distance scaled by 1 + 4 * 2^-23 >= 1 + 2 gamma_3 before the
// near <= far test makes it conservative against the roundings of (b - o) * inv
// (Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013): the watertight mode's
// culling, so a ray leaving through a box face exactly at a vertex keeps the box.
constexpr float ROBUST_FAR = 1.00000048f;

// gi_slab_near (:35-40); ROBUST: the exit distance scaled by ROBUST_FAR
template <bool ROBUST = false>
__device__ __forceinline__ float slab_near(V3 o, V3 inv, V3 lo, V3 hi, float tmin, float tmax) {
    float ax = (lo.x - o.x) * inv.x, bx = (hi.x - o.x) * inv.x;
    float ay = (lo.y - o.y) * inv.y, by = (hi.y - o.y) * inv.y;
    float az = (lo.z - o.z) * inv.z, bz = (hi.z - o.z) * inv.z;
    float near = fmx(fmx(fmx(fmn(ax, bx), fmn(ay, by)), fmn(az, bz)), tmin);
    float exit = fmn(fmn(fmx(ax, bx), fmx(ay, by)), fmx(az, bz));
    if (ROBUST) exit *= ROBUST_FAR;
    float far = fmn(exit, tmax);
    return near <= far ? near : FMAXV;
}

// gi_triangle (:42-58) on vertex a and the edges e1 = b - a, e2 = c - a;
// returns t < 0 on a miss.
__device__ __forceinline__ float tri_test(V3 o, V3 d, V3 a, V3 e1, V3 e2, float tmin, float tmax, float& u_out,
                                          float& v_out) {
    V3 p = cross3(d, e2);
    float det = dot3(e1, p);
    V3 s = sub3(o, a);
    const float su = dot3(s, p);  // before the det test, so that a's load is issued with e1's and e2's
    if (fabsf(det) < 1e-30f) return -1.0f;
    float inv = 1.0f / det;
    float u = su * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    V3 q = cross3(s, e1);
    float v = dot3(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    float t = dot3(e2, q) * inv;
    if (t > tmin && t < tmax) {
        u_out = u;
        v_out = v;
        return t;
    }
    return -1.0f;
}

// Kahan's difference of products a*b - c*d: the rounding error of c*d taken back
// with one fma. The result lies within 2 ulps of the exact value (Jeannerod,
// Louvet and Muller, "Further analysis of Kahan's algorithm for the accurate
// computation of 2x2 determinants", Math. Comp. 82, 2013), so its sign is the
// exact sign and it is zero exactly when a*b == c*d (no overflow or underflow).
__device__ __forceinline__ float diff_of_products(float a, float b, float c, float d) {
    const float w = c * d;
    const float e = __builtin_fmaf(-c, d, w);
    const float f = __builtin_fmaf(a, b, -w);
    return f + e;
}
__device__ __forceinline__ float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013), both faces, all in f32: the vertices are
// translated to the ray origin and sheared so the ray runs along its dominant
// axis; the three 2D edge functions then decide the hit. Computed with the
// exactly-signed difference of products (in place of the paper's f64 retry),
// an edge shared by two triangles gives opposite signs or zero in both, so a
// ray through the edge or a shared vertex cannot miss both. Barycentrics are
// the weights of b and c, as gi_triangle's (u, v); t < 0 on a miss.
__device__ __forceinline__ float wt_tri_test(V3 o, V3 d, V3 a, V3 b, V3 c, float tmin, float tmax, float& u_out,
                                             float& v_out) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const int kz = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dz = comp(d, kz);
    if (dz < 0.0f) {  // keep the winding
        const int k = kx;
        kx = ky;
        ky = k;
    }
    const float sx = comp(d, kx) / dz, sy = comp(d, ky) / dz, sz = 1.0f / dz;
    const V3 A = sub3(a, o), B = sub3(b, o), C = sub3(c, o);
    const float az_ = comp(A, kz), bz_ = comp(B, kz), cz_ = comp(C, kz);
    const float Ax = comp(A, kx) - sx * az_, Ay = comp(A, ky) - sy * az_;
    const float Bx = comp(B, kx) - sx * bz_, By = comp(B, ky) - sy * bz_;
    const float Cx = comp(C, kx) - sx * cz_, Cy = comp(C, ky) - sy * cz_;
    const float U = diff_of_products(Cx, By, Cy, Bx);
    const float V = diff_of_products(Ax, Cy, Ay, Cx);
    const float W = diff_of_products(Bx, Ay, By, Ax);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return -1.0f;
    const float det = U + V + W;
    if (det == 0.0f) return -1.0f;
    const float T = U * (sz * az_) + V * (sz * bz_) + W * (sz * cz_);
    const float inv = 1.0f / det;
    const float t = T * inv;
    if (!(t > tmin && t < tmax)) return -1.0f;
    u_out = V * inv;
    v_out = W * inv;
    return t;
}

// rs_trace's triangle test (shaders/reflections_software_bvh4.wgsl:124-135): the
// same Moller-Trumbore arithmetic with RS_EDGE slack on the barycentric tests
// and no interval test (the caller applies tmin, the best t and the tie rule).
// False on a rejection.
constexpr float RS_EDGE = 4e-6f;
__device__ __forceinline__ bool rs_tri_test(V3 o, V3 d, V3 a, V3 e1, V3 e2, float& t_out, float& u_out,
                                            float& v_out) {
    V3 p = cross3(d, e2);
    float det = dot3(e1, p);
    if (fabsf(det) < 1e-30f) return false;
    float inv = 1.0f / det;
    V3 s = sub3(o, a);
    float u = dot3(s, p) * inv;
    if (u < -RS_EDGE || u > 1.0f + RS_EDGE) return false;
    V3 q = cross3(s, e1);
    float v = dot3(d, q) * inv;
    if (v < -RS_EDGE || u + v > 1.0f + RS_EDGE) return false;
    t_out = dot3(e2, q) * inv;
    u_out = u;
    v_out = v;
    return true;
}

// shaders/scene.wgsl:86-90 applied to mat3x3(r0,r1,r2) * cross(e1, e2)
__device__ __forceinline__ V3 hit_normal(V3 r0, V3 r1, V3 r2, V3 e1, V3 e2) {
    V3 x = cross3(e1, e2);
    V3 n = mk(r0.x * x.x + r1.x * x.y + r2.x * x.z, r0.y * x.x + r1.y * x.y + r2.y * x.z,
              r0.z * x.x + r1.z * x.y + r2.z * x.z);
    float scale = fmx(fmx(fabsf(n.x), fabsf(n.y)), fabsf(n.z));
    float s = fmx(scale, 1e-30f);
    V3 sc = mk(n.x / s, n.y / s, n.z / s);
    float k = 1.0f / sqrtf(fmx(dot3(sc, sc), 1e-30f));
    return mk(sc.x * k, sc.y * k, sc.z * k);
}

__device__ __forceinline__ bool ray_invalid(float4 a, float4 b) {
    bool bad = !isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z) || !isfinite(b.x) || !isfinite(b.y) ||
               !isfinite(b.z);
    bad = bad || !isfinite(a.w) || a.w < 0.0f || b.w != b.w || b.w > 1e30f;
    return bad;
}

struct SceneView {
    const uint32_t* words;  // REF: the packed words
    const uint32_t* inst;   // PLAIN/CW8 instance records (32 words)
    const uint32_t* top;    // top-level wide nodes
    const uint32_t* nodes;  // PLAIN 32 words, CW8 24 words a node
    const float* tris;      // TRI_FLOATS a triangle: a, b - a, c - a
    const uint32_t* tri_ids;  // TRI_ID_WORDS a triangle: ia, ib, ic, slot, primitive, 0
    const float* verts;     // TRI_FLOATS a triangle: a, b, c (the watertight test)
    const uint32_t* roots;  // CW8: a copy of each instance's BLAS root node (24 words)
    uint32_t n_inst, top_on, top_root, pad;
};

struct BatchDev {
    SceneView sv;
    const rq_ray* rays;
    rq_hit* hits;
    uint32_t* vis;
    uint32_t mode, count;
};

struct TraceArgs {
    SceneView sv;
    const rq_ray* rays;
    rq_hit* hits;
    uint32_t* vis;
    uint32_t* counter;
    uint32_t* spill;
    unsigned long long* bytes;  // COUNT instantiations: global bytes moved
    const uint32_t* count_ptr;  // when set, the ray count is min(*count_ptr, count)
    const BatchDev* batches;    // BATCHED: the batch table
    const uint32_t* prefix;     // BATCHED: first global ray of each batch
    const uint32_t* order;      // when set, the i-th ray taken is ray order[i]
    // AM 2 without BATCHED: up to 4 ranges of one scene; range j holds the
    // global rays [rstart[j], rstart[j + 1]) (unused entries: the total).
    const rq_ray* rrays[4];
    rq_hit* rhits[4];
    uint32_t* rvis[4];
    uint32_t rstart[5];
    uint32_t rany;              // bit j: range j is an any-hit range
    uint32_t n_batches;
    uint32_t count, refill_min;
    uint32_t speculate;  // CW8: lanes with triangles pending take node steps with the wave
    uint32_t drain;      // once the rays run out, every live lane takes its step each iteration
};

// Marks the 20 words as used here: the loads that produce them are all issued
// before this point, so they take one memory round trip together, instead of the
// compiler sinking each group into the branch of the test that first reads it
// (three round trips one after another at an instance entry).
__device__ __forceinline__ void hold_words(const uint32_t (&w)[20]) {
    asm volatile("" ::"v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5]), "v"(w[6]), "v"(w[7]),
                 "v"(w[8]), "v"(w[9]), "v"(w[10]), "v"(w[11]), "v"(w[12]), "v"(w[13]), "v"(w[14]), "v"(w[15]),
                 "v"(w[16]), "v"(w[17]), "v"(w[18]), "v"(w[19]));
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <int L>
struct Loader;
template <>
struct Loader<REF> {
    static __device__ __forceinline__ uint32_t w(const uint32_t* p, uint64_t i) { return p[i]; }
};

// 8-key sorting network, descending on key (Batcher's odd-even merge sort, 19 comparators).
__device__ __forceinline__ void cas_desc(uint32_t& ka, uint32_t& va, uint32_t& kb, uint32_t& vb) {
    bool sw = ka < kb;
    uint32_t k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    uint32_t v0 = sw ? vb : va, v1 = sw ? va : vb;
    ka = k0; kb = k1; va = v0; vb = v1;
}

// COUNT: the same traversal, also tallying the bytes every global load and
// store moves (rays, records, nodes, triangles, spills, hit records); used
// once, untimed, to give the roofline its algorithmic bytes per ray.
// AM: 0 closest hit, 1 any hit, 2 per batch or range (mixed modes).
// BATCHED: rays come from a table of batches over several scenes; a lane
// takes its batch's scene, mode and output buffers when it fetches a ray.
// AM 2 without BATCHED: up to 4 ranges of one scene described in the
// arguments; a lane takes its range's mode and outputs the same way.
//
// Loop structure: every iteration each lane takes exactly one small step --
// visit one node, enter one instance, or test the next two triangles of its
// current leaf (the reference's pair step) -- so a wave never executes a
// whole leaf loop for one lane while the others wait on nodes, and the lanes
// whose rays finished are refilled at the top of every iteration. Each ray
// still visits exactly the sequence a one-item-at-a-time traversal visits.
// WT: the watertight leaf test (wt_tri_test) in place of Moller-Trumbore, on the
// triangles' own vertices (REF: through the indices; PLAIN/CW8: their vertex
// records).
template <int LAYOUT, int BLOCK, int S, int AM, bool COUNT = false, bool BATCHED = false, bool WT = false>
__global__ __launch_bounds__(BLOCK, (COUNT || BATCHED) ? 1 : 5) void trace_kernel(TraceArgs a) {
    static_assert((S & (S - 1)) == 0, "stack ring is a power of two");
    // RSF: reflection words (records, triangles, rs_trace's rules); W8: 8-wide nodes.
    constexpr bool RSF = LAYOUT == RS || LAYOUT == RSW;
    constexpr bool W8 = LAYOUT == CW8 || LAYOUT == RSW;
    __shared__ uint32_t lds[S * BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    uint32_t* st = lds + tid;
    uint32_t* spill = a.spill + ((uint64_t)blockIdx.x * BLOCK + tid) * SPILL_WORDS;
    constexpr uint32_t TOP_CAP = W8 ? CW8_CAP : STACK_CAP;
    constexpr uint32_t BLAS_CAP = W8 ? CW8_CAP : STACK_CAP;

    // Lane state.
    uint32_t ray = NONE;
    V3 wo{0, 0, 0}, wd{0, 0, 0}, winv{0, 0, 0}, oo{0, 0, 0}, od{0, 0, 0}, oinv{0, 0, 0};
    float tmin = 0.0f, bt = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t binst = NONE, btri = 0, cur = 0;
    bool bhit = false, invalid = false, in_blas = false;
    uint32_t sp = 0, lo = 0, floor_ = 0, flat_next = 0;
    uint32_t m_nodes = 0, m_leaves = 0, m_pos = 0, m_idx = 0;
    uint32_t tri_next = 0, tri_left = 0;  // the current leaf's untested triangles
    uint32_t nxt = NONE;                  // the next item to visit (the top of the stack, in a register)
    bool exhausted = false;  // the wave found no rays left to claim
    unsigned long long nb = 0;
    // COUNT: [1] wave iterations, [2] node, [3] triangle-pair, [4] instance
    // steps (lane sums), [5] refill events (per wave)
    unsigned long long st_iter = 0, st_node = 0, st_tri = 0, st_inst = 0, st_refill = 0;
    // [6] live lanes, [7] [8] [9] lanes wanting a node / triangle-pair / instance
    // step, summed over wave iterations (wave-uniform)
    unsigned long long st_live = 0, st_want_node = 0, st_want_tri = 0, st_want_inst = 0;
    // [10] [11] wave iterations running node / triangle-pair steps, [12] node
    // iterations in which some lane entered an instance, [13] speculative node
    // steps (lane sums), [14] drain iterations, [15] top-level node steps (lane sums)
    unsigned long long st_node_it = 0, st_tri_it = 0, st_inst_it = 0, st_spec = 0, st_drain = 0, st_top = 0;
    bool entered = false;  // COUNT: the lane entered an instance this iteration
    const uint32_t count = a.count_ptr ? min(*a.count_ptr, a.count) : a.count;

    // ---- the ray cursor. The rays [0, count) are dealt out in chunks of 64
    // (chunk c = rays [64c, 64c + 64), or, in a launch of several ranges, each
    // range's 64-ray chunks in turn); chunk c belongs to partition c mod
    // CURSOR_PARTS, and each partition has its own counter of claimed chunks. Claims
    // thus spread over CURSOR_PARTS addresses instead of queueing on one (an
    // address takes one atomic at a time: a single cursor made every refill wait
    // its turn), while all partitions advance through the rays together, so the
    // grid still works on one narrow window of the ray order. Ray sets come in
    // screen order, so neighbouring rays visit the same nodes, and a narrow
    // window keeps those in L2 (contiguous partitions, one window each, measured
    // 20% slower). Wave w of the grid starts on chunk w without an atomic, then
    // claims one chunk at a time from its home partition, w mod CURSOR_PARTS, and
    // hands the rays to its idle lanes over the next refills. A wave whose
    // partition runs dry moves on to one that still holds chunks and is exhausted
    // when none does.
    // The cursor state is wave-uniform and kept in scalar registers (readfirstlane).
    constexpr uint32_t PARTS = CURSOR_PARTS;
    static_assert(PARTS <= 32, "a 32-bit mask of empty partitions; one lane per partition when probing");
    const uint32_t gwave = blockIdx.x * (uint32_t)(BLOCK / 64) + __builtin_amdgcn_readfirstlane(tid) / 64u;
    const uint64_t waves = (uint64_t)gridDim.x * (BLOCK / 64);
    // A launch of several ranges (AM 2) numbers each range's chunks on from the
    // previous range's, so no chunk holds rays of two ranges: the rays of one
    // refill then share their range, whose buffers are picked once per refill.
    // cend[j]: the chunks of ranges 0..j.
    uint32_t cend[4] = {0u, 0u, 0u, 0u};
    uint64_t chunks;
    if (AM == 2 && !BATCHED) {
        uint32_t c = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c += (a.rstart[j + 1] - a.rstart[j] + 63u) / 64u;
            cend[j] = c;
        }
        chunks = c;
    } else {
        chunks = ((uint64_t)count + 63u) / 64u;
    }
    uint32_t part = gwave % PARTS;
    uint32_t dead = 0u;  // partitions this wave found empty
    // partition p's chunks below `waves` are first slices: its counter counts from there
    auto first_round = [&](uint32_t p) -> uint64_t { return p < waves ? (waves - 1u - p) / PARTS + 1u : 0ull; };
    // the wave's claimed rays not yet handed out: [res_next, res_end), all of range res_j
    uint32_t res_next = 0u, res_end = 0u, res_j = 0u;
    // the rays of chunk c < chunks (64-bit bounds compared by hand: there is no 64-bit
    // min overload that stays integer)
    auto take_chunk = [&](uint64_t c) {
        if (AM == 2 && !BATCHED) {
            res_j = (c >= cend[0] ? 1u : 0u) + (c >= cend[1] ? 1u : 0u) + (c >= cend[2] ? 1u : 0u);
            const uint32_t first = res_j == 0 ? 0u : res_j == 1 ? cend[0] : res_j == 2 ? cend[1] : cend[2];
            const uint32_t lo = res_j == 0 ? a.rstart[0] : res_j == 1 ? a.rstart[1] : res_j == 2 ? a.rstart[2] : a.rstart[3];
            const uint32_t hi = res_j == 0 ? a.rstart[1] : res_j == 1 ? a.rstart[2] : res_j == 2 ? a.rstart[3] : a.rstart[4];
            res_next = lo + 64u * ((uint32_t)c - first);
            res_end = hi - res_next > 64u ? res_next + 64u : hi;
        } else {
            res_next = (uint32_t)(64ull * c);
            res_end = 64ull * c + 64ull < count ? res_next + 64u : count;
        }
    };
    if (gwave < chunks) take_chunk(gwave);
    // Claims the wave's next chunk from its current partition. A partition whose
    // claim fails has no chunks left: the wave moves on to the next partition it
    // has not found empty (and tries it at its next refill), and is exhausted once
    // it has found them all empty.
    auto claim = [&]() {
        uint32_t old = 0;
        if (lane == 0) old = atomicAdd(a.counter + part * CURSOR_STRIDE, 1u);
        old = __builtin_amdgcn_readfirstlane(old);  // every lane is active here: lane 0's value
        const uint64_t c = (first_round(part) + old) * PARTS + part;
        if (c < chunks) {
            take_chunk(c);
            return;
        }
        dead |= 1u << part;
        const uint32_t open = ~dead & (PARTS == 32 ? ~0u : (1u << PARTS) - 1u);
        if (open == 0u) {
            exhausted = true;
            return;
        }
        const uint32_t after = part < 31u ? open & (~0u << (part + 1u)) : 0u;
        part = (uint32_t)__builtin_ctz(after ? after : open);
    };
    SceneView v = a.sv;
    bool any_hit = AM == 1;
    rq_hit* out_hits = a.hits;
    uint32_t* out_vis = a.vis;
    uint32_t out_idx = 0;

    // A node step pushes at most n entries: spill the oldest ring entries
    // first so that n more fit, then push without per-entry checks.
    auto room = [&](uint32_t n) {
        while (sp - lo + n > (uint32_t)S) {
            spill[lo] = st[(lo & (S - 1)) * BLOCK];
            lo++;
            if (COUNT) nb += 4;
        }
    };
    // Branch-free push after room(): the slot above the top is written either
    // way and kept only when `c` holds.
    auto push_if = [&](uint32_t x, bool c) {
        st[(sp & (S - 1)) * BLOCK] = x;
        sp += c ? 1u : 0u;
    };
    // Unconditional write at stack position `at` (room() made it free).
    auto put = [&](uint32_t x, uint32_t at) { st[(at & (S - 1)) * BLOCK] = x; };
    auto pop = [&]() -> uint32_t {
        sp--;
        if (sp < lo) {
            lo = sp;
            if (COUNT) nb += 4;
            return spill[sp];
        }
        return st[(sp & (S - 1)) * BLOCK];
    };
    // gi_limit (:64-70); rs_limit keeps equal-distance candidates of the same
    // instance eligible too (lower primitive wins), hence i <= best there.
    auto limit = [&](uint32_t i) -> float {
        return (bhit && (RSF ? i <= binst : i < binst)) ? bf(fb(bt) + 1u) : bt;
    };
    // ---- hit writeback (the record of a finished ray)
    auto writeback = [&]() {
        uint32_t tri0 = 0, tri1 = 0, tri2 = 0, slot = 0;
        V3 n = mk(0.0f, 0.0f, 0.0f);
        if (RSF && bhit) {
            // rs_trace's record (slot + 1, primitive, u, v): the slot, the primitive
            tri0 = btri;
            slot = v.words[RS_HEADER_WORDS + (uint64_t)binst * RS_RECORD_WORDS + 21u];
            if (COUNT) nb += 4;
        } else if (bhit) {
            V3 r0, r1, r2, e1, e2;
            if (LAYOUT == REF) {
                const uint32_t* r = v.words + 16ull + (uint64_t)binst * 32u;
                r0 = mk(bf(r[0]), bf(r[1]), bf(r[2]));
                r1 = mk(bf(r[4]), bf(r[5]), bf(r[6]));
                r2 = mk(bf(r[8]), bf(r[9]), bf(r[10]));
                uint32_t pos = r[20], idx = r[21], prim_base = r[28];
                uint64_t x = (uint64_t)idx + 3ull * btri;
                tri0 = v.words[x];
                tri1 = v.words[x + 1];
                tri2 = v.words[x + 2];
                const uint32_t* P = v.words + pos;
                const V3 va = mk(bf(P[3ull * tri0]), bf(P[3ull * tri0 + 1]), bf(P[3ull * tri0 + 2]));
                e1 = sub3(mk(bf(P[3ull * tri1]), bf(P[3ull * tri1 + 1]), bf(P[3ull * tri1 + 2])), va);
                e2 = sub3(mk(bf(P[3ull * tri2]), bf(P[3ull * tri2 + 1]), bf(P[3ull * tri2 + 2])), va);
                slot = btri;
                if (!any_hit && prim_base != 0u) {
                    tri0 = v.words[prim_base + (uint64_t)btri];
                    tri1 = 0;
                    tri2 = 0;
                }
            } else {
                const uint32_t* r = v.inst + (uint64_t)binst * 32u;
                r0 = mk(bf(r[0]), bf(r[1]), bf(r[2]));
                r1 = mk(bf(r[4]), bf(r[5]), bf(r[6]));
                r2 = mk(bf(r[8]), bf(r[9]), bf(r[10]));
                const uint32_t flags = r[19];
                const float3* tp = reinterpret_cast<const float3*>(v.tris + (uint64_t)TRI_FLOATS * btri);
                const float3 t1 = tp[1], t2 = tp[2];
                e1 = mk(t1.x, t1.y, t1.z);
                e2 = mk(t2.x, t2.y, t2.z);
                const uint2* ids = reinterpret_cast<const uint2*>(v.tri_ids + (uint64_t)TRI_ID_WORDS * btri);
                const uint2 i0 = ids[0], i1 = ids[1], i2 = ids[2];
                // all of it in one memory round trip (the flags word would otherwise load
                // inside the primitive-table branch, after the rest)
                hold_words({fb(r0.x), fb(r0.y), fb(r0.z), fb(r1.x), fb(r1.y), fb(r1.z), fb(r2.x), fb(r2.y), fb(r2.z),
                            flags, fb(e1.x), fb(e1.y), fb(e1.z), fb(e2.x), fb(e2.y), fb(e2.z), i0.x, i0.y, i1.x, i1.y});
                tri0 = i0.x;
                tri1 = i0.y;
                tri2 = i1.x;
                slot = i1.y;
                if (!any_hit && (flags & 2u)) {
                    tri0 = i2.x;
                    tri1 = 0;
                    tri2 = 0;
                }
            }
            n = hit_normal(r0, r1, r2, e1, e2);
            if (COUNT) nb += LAYOUT == REF ? (36 + 24 + 12 + 36 + (!any_hit ? 4 : 0)) : (36 + 24 + 24);
        }
        if (out_hits) {
            uint4* hp = reinterpret_cast<uint4*>(out_hits + out_idx);
            uint32_t flags = invalid ? 2u : (bhit ? 1u : 0u);
            hp[0] = make_uint4(fb(bt), bhit ? binst : NONE, tri0, tri1);
            hp[1] = make_uint4(tri2, fb(bhit ? bu : 0.0f), fb(bhit ? bv : 0.0f), fb(n.x));
            hp[2] = make_uint4(fb(n.y), fb(n.z), flags, slot);
        }
        if (COUNT) nb += (out_hits ? 48 : 0) + (out_vis && !bhit ? 4 : 0);
        if (out_vis && !bhit) atomicOr(out_vis + (out_idx >> 5), 1u << (out_idx & 31u));
    };

    bool done = false;  // the ray is finished; its record is written at the next refill
    for (;;) {
        // ---- write back finished rays and refill idle lanes (wave-uniform decision):
        // only once enough lanes are free, so this code runs once per many rays.
        const uint64_t idle = __ballot(ray == NONE || done);
        const bool refill_now = idle != 0ull &&
                                ((uint32_t)__popcll(idle) >= a.refill_min || idle == ~0ull || exhausted);
        if (refill_now && done) {
            writeback();
            ray = NONE;
            done = false;
        }
        if (COUNT && refill_now && !exhausted) st_refill++;
        if (!exhausted && refill_now) {
            // hand the wave's claimed rays to its idle lanes in lane order (claiming a
            // chunk first if none is left); lanes left over wait for the next refill
            if (res_next == res_end) claim();
            const uint32_t k = min((uint32_t)__popcll(idle), res_end - res_next);
            const uint32_t rank = lane_rank(idle);
            const uint32_t mine = res_next + rank;
            res_next += k;
            if (ray == NONE && rank < k) {
                {
                    uint32_t r = mine;
                    if (!BATCHED && a.order) {
                        r = a.order[r];
                        if (COUNT) nb += 4;
                    }
                    ray = r;
                    const rq_ray* src = a.rays + r;
                    out_idx = r;
                    if (AM == 2 && !BATCHED) {
                        const uint32_t j = res_j;  // every ray of this refill is of range res_j
                        const rq_ray* rb = j == 0 ? a.rrays[0] : j == 1 ? a.rrays[1] : j == 2 ? a.rrays[2] : a.rrays[3];
                        out_hits = j == 0 ? a.rhits[0] : j == 1 ? a.rhits[1] : j == 2 ? a.rhits[2] : a.rhits[3];
                        out_vis = j == 0 ? a.rvis[0] : j == 1 ? a.rvis[1] : j == 2 ? a.rvis[2] : a.rvis[3];
                        out_idx = r - (j == 0 ? a.rstart[0] : j == 1 ? a.rstart[1] : j == 2 ? a.rstart[2] : a.rstart[3]);
                        any_hit = ((a.rany >> j) & 1u) != 0u;
                        src = rb + out_idx;
                    }
                    if (BATCHED) {
                        // batch b holds global rays [prefix[b], prefix[b+1])
                        uint32_t lo_b = 0, hi_b = a.n_batches;
                        while (hi_b - lo_b > 1) {
                            uint32_t mid = (lo_b + hi_b) >> 1;
                            if (a.prefix[mid] <= r) lo_b = mid; else hi_b = mid;
                        }
                        const BatchDev& bd = a.batches[lo_b];
                        v = bd.sv;
                        any_hit = bd.mode != 0;
                        out_hits = bd.hits;
                        out_vis = bd.vis;
                        out_idx = r - a.prefix[lo_b];
                        src = bd.rays + out_idx;
                        if (COUNT) nb += 8 + sizeof(BatchDev);
                    }
                    const float4* rp = reinterpret_cast<const float4*>(src);
                    float4 r0 = rp[0], r1 = rp[1];
                    if (COUNT) nb += 32;
                    wo = mk(r0.x, r0.y, r0.z);
                    tmin = r0.w;
                    wd = mk(r1.x, r1.y, r1.z);
                    bt = r1.w;
                    invalid = ray_invalid(r0, r1);
                    winv = LAYOUT == CW8 ? fast_inverse(wd) : safe_inverse(wd);
                    binst = NONE;
                    bhit = false;
                    btri = 0;
                    bu = 0.0f;
                    bv = 0.0f;
                    sp = 0;
                    lo = 0;
                    in_blas = false;
                    flat_next = 0;
                    tri_left = 0;
                    nxt = (!invalid && v.top_on) ? v.top_root : NONE;
                    if (RSF && !invalid) {
                        // rs_trace's scene-bounds pre-cull (:175-187): a ray that misses
                        // the padded union of the instances is a miss; otherwise the best
                        // t starts 2 ulps past the exit so boundary triangles stay eligible.
                        bool out = v.n_inst == 0u;
                        if (!out) {
                            const uint4* bp = reinterpret_cast<const uint4*>(v.words + v.words[4]);
                            const uint4 q0 = bp[0], q1 = bp[1];
                            if (COUNT) nb += 32;
                            const float ax = (bf(q0.x) - wo.x) * winv.x, bx = (bf(q1.x) - wo.x) * winv.x;
                            const float ay = (bf(q0.y) - wo.y) * winv.y, by = (bf(q1.y) - wo.y) * winv.y;
                            const float az = (bf(q0.z) - wo.z) * winv.z, bz = (bf(q1.z) - wo.z) * winv.z;
                            const float near = fmx(fmx(fmx(fmn(ax, bx), fmn(ay, by)), fmn(az, bz)), tmin);
                            const float far = fmn(fmn(fmn(fmx(ax, bx), fmx(ay, by)), fmx(az, bz)), bt);
                            if (near > far) out = true;
                            else bt = fmn(bt, bf(fb(far) + 2u));
                        }
                        if (out) {
                            nxt = NONE;
                            flat_next = v.n_inst;
                        }
                    }
                }
            }
        }
        if (exhausted && __ballot(ray != NONE) == 0ull) break;
        if (COUNT) {
            st_iter++;
            st_live += (unsigned long long)__popcll(__ballot(ray != NONE && !done));
        }
        if (ray == NONE || done) continue;

        bool fin = invalid;
        // ---- the next item of a lane that has none: pop it, or leave the BLAS
        if (!fin && tri_left == 0) {
            if (nxt == NONE) {
                if (in_blas && sp == floor_) in_blas = false;
                if (!in_blas && !v.top_on) {
                    if (flat_next < v.n_inst) nxt = LEAF_BIT | flat_next++;
                } else if (sp > 0) {
                    nxt = pop();
                }
                if (nxt == NONE) fin = true;
            }
            if (!fin && in_blas && (nxt & LEAF_BIT)) {
                // a leaf: its triangles are tested by the next steps
                if (LAYOUT == REF) {
                    uint64_t l = m_leaves + (uint64_t)(nxt & ~LEAF_BIT) * 2u;
                    tri_next = v.words[l] / 3u;
                    tri_left = v.words[l + 1] / 3u;
                    if (COUNT) nb += 8;
                } else if (RSF) {
                    tri_next = (nxt & ~LEAF_BIT) >> 4;  // LEAF | first triangle << 4 | count
                    tri_left = nxt & 15u;
                } else {
                    tri_next = nxt & LEAF_MASK;
                    tri_left = ((nxt >> 27) & 0xfu) + 1u;
                }
                nxt = NONE;
            }
        }
        // ---- the wave runs one kind of step at a time (1 triangle pair, 2 instance
        // entry, 3 node), chosen by how many lanes want each. The others wait a step;
        // each lane's own sequence of steps is unchanged.
        // The 8-wide layouts enter an instance and visit its BLAS root in one node step
        // (kind 3 covers both), so instance entries share the node steps' iterations;
        // each lane's sequence of visits is unchanged. (PLAIN measured slower merged.)
        // Their node step costs about twice a pair step, so they take node steps only
        // while twice as many lanes want one as want a pair (the measured optimum on
        // the 4M-triangle frame); PLAIN takes whichever most lanes want.
        constexpr bool MERGE = W8;
        constexpr uint32_t TRI_WEIGHT = W8 ? 2u : 1u;
        const int kind = fin ? 0 : (tri_left > 0 ? 1 : ((nxt & LEAF_BIT) && !MERGE ? 2 : 3));
        const uint32_t n_tri = (uint32_t)__popcll(__ballot(kind == 1));
        const uint32_t n_inst = MERGE ? 0u : (uint32_t)__popcll(__ballot(kind == 2));
        const uint32_t n_node = (uint32_t)__popcll(__ballot(kind == 3));
        const int pick = (n_node >= n_tri * TRI_WEIGHT && n_node >= n_inst) ? 3 : (n_tri >= n_inst ? 1 : 2);
        // CW8 speculation (after Aila and Laine, "Understanding the Efficiency of Ray
        // Traversal on GPUs", HPG 2009): while the wave runs node steps, a lane whose
        // leaf still has triangles to test takes a node step too -- its pending node,
        // or the node on top of its stack -- and tests the triangles later. Each lane
        // visits the same nodes and triangles it may hit; only the order changes, so
        // closest hits differ at most between equal-distance candidates (the CW8
        // contract) and visibility not at all.
        const bool dual = a.drain != 0u && exhausted;
        bool spec = false;
        uint32_t spec_item = NONE;
        if (W8 && a.speculate && !dual && pick == 3 && kind == 1 && in_blas) {
            if (nxt != NONE) {
                if (!(nxt & LEAF_BIT)) {
                    spec = true;
                    spec_item = nxt;
                }
            } else if (sp > floor_ && sp > lo) {
                const uint32_t top = st[((sp - 1) & (S - 1)) * BLOCK];
                if (!(top & LEAF_BIT)) {
                    spec = true;
                    spec_item = top;
                    sp--;
                }
            }
        }
        if (COUNT) {
            // once per wave iteration: by the lowest lane still here (summed over lanes below)
            if (lane == (uint32_t)__ffsll((unsigned long long)__ballot(true)) - 1u) {
                st_want_node += n_node;
                st_want_tri += n_tri;
                st_want_inst += n_inst;
            }
            st_tri += kind == 1 && (pick == 1 || dual);  // node and instance steps: counted where they run
        }
        if (kind == 1 && (pick == 1 || dual)) {
            // ---- triangle step (:126-168): the next two triangles in leaf order
            const uint32_t k = tri_left >= 2 ? 2u : 1u;
            // Both triangles' records (the second is the first again for a lone
            // triangle, so every load stays inside the leaf) are loaded before the
            // first test: one memory round trip for the pair, not one each.
            const uint32_t tri2 = tri_next + (k == 2 ? 1u : 0u);
            uint4 rsw[2][3];
            float3 tw[2][3];
            if (RSF) {
                const uint4* tp0 = reinterpret_cast<const uint4*>(v.words + m_idx + (uint64_t)RS_TRI_WORDS * tri_next);
                const uint4* tp1 = reinterpret_cast<const uint4*>(v.words + m_idx + (uint64_t)RS_TRI_WORDS * tri2);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    rsw[0][q] = tp0[q];
                    rsw[1][q] = tp1[q];
                }
                hold_words({rsw[0][0].x, rsw[0][0].y, rsw[0][0].z, rsw[0][0].w, rsw[0][1].x, rsw[0][1].y, rsw[0][1].z,
                            rsw[0][1].w, rsw[0][2].x, rsw[0][2].y, rsw[1][0].x, rsw[1][0].y, rsw[1][0].z, rsw[1][0].w,
                            rsw[1][1].x, rsw[1][1].y, rsw[1][1].z, rsw[1][1].w, rsw[1][2].x, rsw[1][2].y});
            } else if (LAYOUT != REF) {
                const float* base = WT ? v.verts : v.tris;
                const float3* tp0 = reinterpret_cast<const float3*>(base + (uint64_t)TRI_FLOATS * tri_next);
                const float3* tp1 = reinterpret_cast<const float3*>(base + (uint64_t)TRI_FLOATS * tri2);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    tw[0][q] = tp0[q];
                    tw[1][q] = tp1[q];
                }
                hold_words({fb(tw[0][0].x), fb(tw[0][0].y), fb(tw[0][0].z), fb(tw[0][1].x), fb(tw[0][1].y),
                            fb(tw[0][1].z), fb(tw[0][2].x), fb(tw[0][2].y), fb(tw[0][2].z), fb(tw[1][0].x),
                            fb(tw[1][0].y), fb(tw[1][0].z), fb(tw[1][1].x), fb(tw[1][1].y), fb(tw[1][1].z),
                            fb(tw[1][2].x), fb(tw[1][2].y), fb(tw[1][2].z), 0u, 0u});
            }
#pragma unroll
            for (uint32_t j = 0; j < 2; ++j) {
                if (j >= k || fin) break;
                if (RSF) {
                    // rs_trace leaf (:118-141): words 0-2 corner a, 3-5 e1, 6-8 e2, 9 primitive
                    const uint4 w0 = rsw[j][0], w1 = rsw[j][1], w2 = rsw[j][2];
                    if (COUNT) nb += 48;
                    float t, u, vv;
                    if (rs_tri_test(oo, od, mk(bf(w0.x), bf(w0.y), bf(w0.z)), mk(bf(w0.w), bf(w1.x), bf(w1.y)),
                                    mk(bf(w1.z), bf(w1.w), bf(w2.x)), t, u, vv)) {
                        const bool tied = bhit && t == bt && (cur < binst || (cur == binst && w2.y < btri));
                        if (t > tmin && (t < bt || tied)) {
                            bt = t; bu = u; bv = vv; binst = cur; bhit = true;
                            btri = w2.y;  // the source primitive: the tie key and the record
                            if (any_hit) fin = true;
                        }
                    }
                    continue;
                }
                V3 va, e1, e2;
                const uint32_t slot_id = tri_next + j;
                float u, vv, t;
                if (LAYOUT == REF) {
                    uint64_t x = (uint64_t)m_idx + 3ull * (tri_next + j);
                    uint32_t ia = v.words[x], ib = v.words[x + 1], ic = v.words[x + 2];
                    const uint32_t* P = v.words + m_pos;
                    va = mk(bf(P[3ull * ia]), bf(P[3ull * ia + 1]), bf(P[3ull * ia + 2]));
                    const V3 vb = mk(bf(P[3ull * ib]), bf(P[3ull * ib + 1]), bf(P[3ull * ib + 2]));
                    const V3 vc = mk(bf(P[3ull * ic]), bf(P[3ull * ic + 1]), bf(P[3ull * ic + 2]));
                    if (WT) {
                        t = wt_tri_test(oo, od, va, vb, vc, tmin, limit(cur), u, vv);
                    } else {
                        e1 = sub3(vb, va);
                        e2 = sub3(vc, va);
                        t = tri_test(oo, od, va, e1, e2, tmin, limit(cur), u, vv);
                    }
                } else if (WT) {
                    const float3 t0 = tw[j][0], t1 = tw[j][1], t2 = tw[j][2];
                    t = wt_tri_test(oo, od, mk(t0.x, t0.y, t0.z), mk(t1.x, t1.y, t1.z), mk(t2.x, t2.y, t2.z), tmin,
                                    limit(cur), u, vv);
                } else {
                    const float3 t0 = tw[j][0], t1 = tw[j][1], t2 = tw[j][2];
                    va = mk(t0.x, t0.y, t0.z);
                    e1 = mk(t1.x, t1.y, t1.z);
                    e2 = mk(t2.x, t2.y, t2.z);
                    t = tri_test(oo, od, va, e1, e2, tmin, limit(cur), u, vv);
                }
                if (COUNT) nb += LAYOUT == REF ? 48 : 4 * TRI_FLOATS;
                if (t >= 0.0f) {
                    bt = t; bu = u; bv = vv; binst = cur; bhit = true;
                    btri = slot_id;
                    if (any_hit) fin = true;
                }
            }
            tri_next += k;
            tri_left -= k;
        } else if ((kind >= 2 && (kind == pick || dual)) || spec) {
            uint32_t item = spec ? spec_item : nxt;
            nxt = NONE;  // consumed (a speculating lane's pending node, or its stack top already popped)
            // an instance (a top-level or flat-loop leaf; speculation takes nodes only) or a node
            bool node_step = !(item & LEAF_BIT);
            // 8-wide layouts: the node this step tests is loaded first -- for an instance entry the
            // copy of its BLAS root kept beside the records -- so its loads are in flight
            // together with the instance record's (one memory round trip for the step).
            uint4 cq0, cq1, cq2, cq3, cq4, cq5;
            if (W8) {
                const uint32_t* np = node_step ? (in_blas ? v.nodes : v.top) + (uint64_t)item * 24u
                                               : v.roots + (uint64_t)(item & ~LEAF_BIT) * 24u;
                const uint4* w4 = reinterpret_cast<const uint4*>(np);
                cq0 = w4[0]; cq1 = w4[1]; cq2 = w4[2]; cq3 = w4[3]; cq4 = w4[4]; cq5 = w4[5];
            }
            if (!node_step) {
                if (COUNT) {
                    st_inst++;
                    entered = true;
                }
                // gi_instance prologue (:100-113)
                const uint32_t i = item & ~LEAF_BIT;
                uint32_t rw[20];
                if (RSF) {
                    // rs_instance (:79-99): record words 0-11 rows, 12-17 world box,
                    // 18 node base, 19 triangle base, 20 root, 21 slot, 22 casts no shadow
                    const uint4* r4 = reinterpret_cast<const uint4*>(v.words + RS_HEADER_WORDS + (uint64_t)i * RS_RECORD_WORDS);
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        uint4 q = r4[k];
                        rw[4 * k] = q.x; rw[4 * k + 1] = q.y; rw[4 * k + 2] = q.z; rw[4 * k + 3] = q.w;
                    }
                    const uint4 tail = r4[5];
                    m_nodes = rw[18];
                    m_idx = rw[19];
                    rw[18] = tail.x;  // root
                    rw[19] = tail.z != 0u ? 1u : 0u;  // casts no shadow
                    if (COUNT) nb += 96;
                } else if (LAYOUT == REF) {
                    const uint32_t* r = v.words + 16ull + (uint64_t)i * 32u;
#pragma unroll
                    for (int k = 0; k < 18; ++k) rw[k] = r[k];
                    rw[19] = any_hit ? r[27] : 0u;
                    if (COUNT) nb += any_hit ? 76 : 72;
                } else {
                    const uint4* r4 = reinterpret_cast<const uint4*>(v.inst + (uint64_t)i * 32u);
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        uint4 q = r4[k];
                        rw[4 * k] = q.x; rw[4 * k + 1] = q.y; rw[4 * k + 2] = q.z; rw[4 * k + 3] = q.w;
                    }
                    if (COUNT) nb += 80;
                }
                hold_words(rw);
                const bool skip = any_hit && (rw[19] & 1u);
                const float lim = limit(i);
                // The world-box test only culls. The 8-wide layouts skip it: their instances
                // are reached through top-level leaf boxes (their relayouts build a top level
                // for any instance count) that hold the box of the instance's triangles,
                // clipped to the record box and quantized outward, so the test would cull few
                // rays and cost more than the node steps it saves.
                const bool box_test = !W8;
                if (!skip && (!box_test || slab_near<WT>(wo, winv, mk(bf(rw[12]), bf(rw[13]), bf(rw[14])),
                                                         mk(bf(rw[15]), bf(rw[16]), bf(rw[17])), tmin, lim) <= lim)) {
                    V3 r0 = mk(bf(rw[0]), bf(rw[1]), bf(rw[2])), r1 = mk(bf(rw[4]), bf(rw[5]), bf(rw[6])),
                       r2 = mk(bf(rw[8]), bf(rw[9]), bf(rw[10]));
                    oo = mk(dot3(r0, wo) + bf(rw[3]) * 1.0f, dot3(r1, wo) + bf(rw[7]) * 1.0f,
                            dot3(r2, wo) + bf(rw[11]) * 1.0f);
                    od = mk(dot3(r0, wd), dot3(r1, wd), dot3(r2, wd));
                    oinv = LAYOUT == CW8 ? fast_inverse(od) : safe_inverse(od);
                    cur = i;
                    uint32_t root;
                    if (RSF) {
                        root = rw[18];
                    } else if (LAYOUT == REF) {
                        const uint32_t* r = v.words + 16ull + (uint64_t)i * 32u;
                        m_nodes = r[18];
                        m_leaves = r[19];
                        m_pos = r[20];
                        m_idx = r[21];
                        root = r[22];
                        if (COUNT) nb += 20;
                    } else {
                        root = rw[18];
                    }
                    floor_ = sp;
                    in_blas = true;
                    if (MERGE && !(root & LEAF_BIT)) {
                        item = root;  // the BLAS root: visited in this same step
                        node_step = true;
                    } else {
                        nxt = root;  // pushed and popped at once
                    }
                }
            }
            if (node_step && COUNT) {
                st_node++;
                st_spec += spec;
                st_top += !in_blas;
            }
            if (!node_step) {
            } else if (LAYOUT == REF || LAYOUT == RS) {
                // RS: rs_trace's top level and BLAS nodes are the same records; its
                // closest hit does not depend on the visit order (rs_limit + the
                // primitive tie rule), and any-hit rays report visibility bits
                if (!in_blas) {
                    // top-level wide node (:240-261), the reference's own sort network
                    const uint32_t* w = v.top + (uint64_t)item * 32u;
                    uint32_t tw[28];
#pragma unroll
                    for (int k = 0; k < 28; ++k) tw[k] = w[k];
                    if (COUNT) nb += 112;
                    float bound = bhit ? bf(fb(bt) + 1u) : bt;
                    float dist[4];
                    uint32_t ids[4];
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        float ax = (bf(tw[l]) - wo.x) * winv.x, bx = (bf(tw[12 + l]) - wo.x) * winv.x;
                        float ay = (bf(tw[4 + l]) - wo.y) * winv.y, by = (bf(tw[16 + l]) - wo.y) * winv.y;
                        float az = (bf(tw[8 + l]) - wo.z) * winv.z, bz = (bf(tw[20 + l]) - wo.z) * winv.z;
                        float near = fmx(fmx(fmx(fmn(ax, bx), fmn(ay, by)), fmn(az, bz)), tmin);
                        float exit = fmn(fmn(fmx(ax, bx), fmx(ay, by)), fmx(az, bz));
                        if (WT) exit *= ROBUST_FAR;
                        float far = fmn(exit, bound);
                        ids[l] = tw[24 + l];
                        dist[l] = (near <= far && ids[l] != EMPTY) ? near : FMAXV;
                    }
#define RQ_SW(A, B)                \
    {                              \
        float td = dist[A];        \
        dist[A] = dist[B];         \
        dist[B] = td;              \
        uint32_t ti = ids[A];      \
        ids[A] = ids[B];           \
        ids[B] = ti;               \
    }
                    if (dist[1] < dist[0]) RQ_SW(0, 1);
                    if (dist[3] < dist[2]) RQ_SW(2, 3);
                    if (dist[2] < dist[0]) RQ_SW(0, 2);
                    if (dist[3] < dist[1]) RQ_SW(1, 3);
                    if (dist[2] < dist[1]) RQ_SW(1, 2);
#undef RQ_SW
                    room(3);
#pragma unroll
                    for (int l = 3; l >= 1; --l) push_if(ids[l], dist[l] <= bound && sp < STACK_CAP);
                    if (dist[0] <= bound && sp < STACK_CAP) nxt = ids[0];  // the nearest, popped next
                } else {
                // BLAS node (:171-204), decoded from the quantized record
                const float lim = limit(cur);
                const uint32_t* w = v.words + m_nodes + (uint64_t)item * 16u;
                uint32_t nw[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) nw[k] = w[k];
                if (COUNT) nb += 64;
                V3 oq = mk(bf(nw[0]), bf(nw[1]), bf(nw[2])), stp = mk(bf(nw[3]), bf(nw[4]), bf(nw[5]));
                float t[4];
                uint32_t c[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    c[s] = nw[6 + s];
                    t[s] = FMAXV;
                    if (c[s] != EMPTY) {
                        uint32_t sh = 8u * s;
                        V3 lq = mk((float)((nw[10] >> sh) & 0xffu), (float)((nw[11] >> sh) & 0xffu),
                                   (float)((nw[12] >> sh) & 0xffu));
                        V3 hq = mk((float)((nw[13] >> sh) & 0xffu), (float)((nw[14] >> sh) & 0xffu),
                                   (float)((nw[15] >> sh) & 0xffu));
                        V3 lo3 = mk(oq.x + stp.x * lq.x, oq.y + stp.y * lq.y, oq.z + stp.z * lq.z);
                        V3 hi3 = mk(oq.x + stp.x * hq.x, oq.y + stp.y * hq.y, oq.z + stp.z * hq.z);
                        t[s] = slab_near<WT>(oo, oinv, lo3, hi3, tmin, lim);
                    }
                }
                uint32_t k0 = fb(t[0]), k1 = fb(t[1]), k2 = fb(t[2]), k3 = fb(t[3]);
                cas_desc(k0, c[0], k1, c[1]);
                cas_desc(k2, c[2], k3, c[3]);
                cas_desc(k0, c[0], k2, c[2]);
                cas_desc(k1, c[1], k3, c[3]);
                cas_desc(k1, c[1], k2, c[2]);
                room(3);
                push_if(c[0], bf(k0) <= lim && sp - floor_ < STACK_CAP);
                push_if(c[1], bf(k1) <= lim && sp - floor_ < STACK_CAP);
                push_if(c[2], bf(k2) <= lim && sp - floor_ < STACK_CAP);
                if (bf(k3) <= lim && sp - floor_ < STACK_CAP) nxt = c[3];  // the nearest, popped next
                }
            } else {
            // PLAIN / CW8: one node format for the top level and the BLAS; the
            // level only picks the ray, the bound and the stack budget.
            const bool blas = in_blas;
            const V3 ro = blas ? oo : wo, ri = blas ? oinv : winv;
            const float lim = blas ? limit(cur) : (bhit ? bf(fb(bt) + 1u) : bt);
            const uint32_t cap = blas ? floor_ + BLAS_CAP : TOP_CAP;
            if (LAYOUT == PLAIN) {
                const uint4* w4 = reinterpret_cast<const uint4*>((blas ? v.nodes : v.top) + (uint64_t)item * 32u);
                uint32_t nw[28];
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    uint4 q = w4[k];
                    nw[4 * k] = q.x; nw[4 * k + 1] = q.y; nw[4 * k + 2] = q.z; nw[4 * k + 3] = q.w;
                }
                if (COUNT) nb += 112;
                float t[4];
                uint32_t c[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    c[s] = nw[24 + s];
                    t[s] = FMAXV;
                    if (c[s] != EMPTY)
                        t[s] = slab_near<WT>(ro, ri, mk(bf(nw[s]), bf(nw[4 + s]), bf(nw[8 + s])),
                                         mk(bf(nw[12 + s]), bf(nw[16 + s]), bf(nw[20 + s])), tmin, lim);
                }
                uint32_t k0 = fb(t[0]), k1 = fb(t[1]), k2 = fb(t[2]), k3 = fb(t[3]);
                cas_desc(k0, c[0], k1, c[1]);
                cas_desc(k2, c[2], k3, c[3]);
                cas_desc(k0, c[0], k2, c[2]);
                cas_desc(k1, c[1], k3, c[3]);
                cas_desc(k1, c[1], k2, c[2]);
                room(3);
                push_if(c[0], bf(k0) <= lim && sp < cap);
                push_if(c[1], bf(k1) <= lim && sp < cap);
                push_if(c[2], bf(k2) <= lim && sp < cap);
                if (bf(k3) <= lim && sp < cap) nxt = c[3];  // the nearest, popped next
            } else {
                // CW8 node: 24 words (origin, exponents + valid mask, byte planes, tags).
                // Planes decode exactly (origin + 2^e * q is an f32 by construction).
                // Each bound is fma(q, 2^e / d, (origin - o) / d - err): one fma a
                // plane after a per-node setup. The roundings (origin - o, the
                // reciprocal's 1 ulp, two products, the slack subtraction and the fma)
                // stay below 7 * 2^-24 * M with M = |base| + 255 |scale|, so a slack
                // of 2^-20 * M keeps culling conservative.
                uint4 q0, q1, q2, q3, q4, q5;
                if (W8) {
                    q0 = cq0; q1 = cq1; q2 = cq2; q3 = cq3; q4 = cq4; q5 = cq5;
                } else {
                    const uint4* w4 = reinterpret_cast<const uint4*>((blas ? v.nodes : v.top) + (uint64_t)item * 24u);
                    q0 = w4[0]; q1 = w4[1]; q2 = w4[2]; q3 = w4[3]; q4 = w4[4]; q5 = w4[5];
                }
                if (COUNT) nb += 96;
                // An empty slot holds an inverted box (low codes 255, high codes 0), which fails
                // the slab test unless the ray's origin is about 2^19 node extents away (the
                // slack then exceeds the node), and a copy of a real child's tag: visiting it is extra work, never a
                // wrong answer. So no slot needs the valid mask.
                // 2^e * (1/d) per axis, e a signed byte: exact, as a product with a power of two is
                const V3 scl = mk(ldexpf(ri.x, (int8_t)(q0.w & 0xffu)), ldexpf(ri.y, (int8_t)((q0.w >> 8) & 0xffu)),
                                  ldexpf(ri.z, (int8_t)((q0.w >> 16) & 0xffu)));
                const V3 base = mk((bf(q0.x) - ro.x) * ri.x, (bf(q0.y) - ro.y) * ri.y, (bf(q0.z) - ro.z) * ri.z);
                const float err = fmx(fmx(__builtin_fmaf(255.0f, fabsf(scl.x), fabsf(base.x)),
                                          __builtin_fmaf(255.0f, fabsf(scl.y), fabsf(base.y))),
                                      __builtin_fmaf(255.0f, fabsf(scl.z), fabsf(base.z))) * 9.53674316e-7f;  // 2^-20
                const V3 bn = mk(base.x - err, base.y - err, base.z - err);  // near planes
                const V3 bfr = mk(base.x + err, base.y + err, base.z + err); // far planes
                // planes: q1 = lo x (2 words), lo y; q2 = lo z, hi x; q3 = hi y, hi z
                const bool px = ri.x >= 0.0f, py = ri.y >= 0.0f, pz = ri.z >= 0.0f;
                const uint32_t nx0 = px ? q1.x : q2.z, nx1 = px ? q1.y : q2.w, fx0 = px ? q2.z : q1.x, fx1 = px ? q2.w : q1.y;
                const uint32_t ny0 = py ? q1.z : q3.x, ny1 = py ? q1.w : q3.y, fy0 = py ? q3.x : q1.z, fy1 = py ? q3.y : q1.w;
                const uint32_t nz0 = pz ? q2.x : q3.z, nz1 = pz ? q2.y : q3.w, fz0 = pz ? q3.z : q2.x, fz1 = pz ? q3.w : q2.y;
                const uint32_t tag[8] = {q4.x, q4.y, q4.z, q4.w, q5.x, q5.y, q5.z, q5.w};
                uint32_t hit = 0;
                // Closest-hit rays visit the two nearest hit children first, in order:
                // the two smallest of (entry distance bits with the slot in the low 3
                // bits) -- entry distances are >= +-0, so clearing the sign orders them
                // as integers. The other hit children go below them in slot order.
                // Any-hit rays end at their first hit and their visibility bit does
                // not depend on the order: they push in slot order.
                const bool order = AM == 0 || (AM == 2 && !any_hit);
                uint32_t kmin = 0xffffffffu, kmin2 = 0xffffffffu;
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const int sh = 8 * (s & 3);
                    const uint32_t wnx = s < 4 ? nx0 : nx1, wny = s < 4 ? ny0 : ny1, wnz = s < 4 ? nz0 : nz1;
                    const uint32_t wfx = s < 4 ? fx0 : fx1, wfy = s < 4 ? fy0 : fy1, wfz = s < 4 ? fz0 : fz1;
                    const float tnx = __builtin_fmaf((float)((wnx >> sh) & 0xffu), scl.x, bn.x);
                    const float tny = __builtin_fmaf((float)((wny >> sh) & 0xffu), scl.y, bn.y);
                    const float tnz = __builtin_fmaf((float)((wnz >> sh) & 0xffu), scl.z, bn.z);
                    const float tfx = __builtin_fmaf((float)((wfx >> sh) & 0xffu), scl.x, bfr.x);
                    const float tfy = __builtin_fmaf((float)((wfy >> sh) & 0xffu), scl.y, bfr.y);
                    const float tfz = __builtin_fmaf((float)((wfz >> sh) & 0xffu), scl.z, bfr.z);
                    const float near = fmx(fmx(fmx(tnx, tny), tnz), tmin);
                    const float far = fmn(fmn(fmn(tfx, tfy), tfz), lim);
                    const bool h = near <= far;
                    if (h) hit |= 1u << s;
                    if (order) {
                        const uint32_t key = h ? ((fb(near) & 0x7ffffff8u) | (uint32_t)s) : 0xffffffffu;
                        kmin2 = min(kmin2, max(kmin, key));
                        kmin = min(kmin, key);
                    }
                }
                // tag[m] for a lane-varying slot m: a select tree on its bits
                auto tag_of = [&](uint32_t m) {
                    const uint32_t a0 = (m & 1u) ? tag[1] : tag[0], a1 = (m & 1u) ? tag[3] : tag[2];
                    const uint32_t a2 = (m & 1u) ? tag[5] : tag[4], a3 = (m & 1u) ? tag[7] : tag[6];
                    const uint32_t b0 = (m & 2u) ? a1 : a0, b1 = (m & 2u) ? a3 : a2;
                    return (m & 4u) ? b1 : b0;
                };
                if (order) {
                    uint32_t first = NONE, second = NONE;
                    if (hit) {
                        first = tag_of(kmin & 7u);
                        hit &= ~(1u << (kmin & 7u));
                    }
                    if (hit) {
                        second = tag_of(kmin2 & 7u);
                        hit &= ~(1u << (kmin2 & 7u));
                    }
                    // Pushes by prefix count: slot k goes above the hit slots beyond it
                    // (pushed first). A slot that is not pushed writes the word the next
                    // pushed slot overwrites, or the free word above the top. The CW8
                    // relayout keeps every tree within CW8_LEVELS levels, so a level's
                    // stack stays within 7 * CW8_LEVELS = CW8_CAP entries unchecked.
                    room(8);
#pragma unroll
                    for (int k = 7; k >= 0; --k) put(tag[k], sp + (uint32_t)__popc(hit >> (k + 1)));
                    const uint32_t n_hit = (uint32_t)__popc(hit);
                    put(second, sp + n_hit);
                    sp += n_hit + (second != NONE ? 1u : 0u);
                    nxt = first;
                } else {
                    room(7);
#pragma unroll
                    for (int k = 7; k >= 1; --k) put(tag[k], sp + (uint32_t)__popc(hit >> (k + 1)));
                    sp += (uint32_t)__popc(hit >> 1);
                    if (hit & 1u) nxt = tag[0];
                }
                (void)cap;
            }
            }
        }
        if (COUNT) {
            const bool node_lane = (kind >= 2 && (kind == pick || dual)) || spec;
            const bool tri_lane = kind == 1 && (pick == 1 || dual);
            const uint64_t nm = __ballot(node_lane), tm = __ballot(tri_lane), im = __ballot(entered);
            if (lane == (uint32_t)__ffsll((unsigned long long)__ballot(true)) - 1u) {
                st_node_it += nm != 0ull;
                st_tri_it += tm != 0ull;
                st_inst_it += im != 0ull;
                st_drain += dual;
            }
            entered = false;
        }
        if (fin) done = true;
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) {
            nb += __shfl_xor(nb, off);
            st_node += __shfl_xor(st_node, off);
            st_tri += __shfl_xor(st_tri, off);
            st_inst += __shfl_xor(st_inst, off);
            st_want_node += __shfl_xor(st_want_node, off);
            st_want_tri += __shfl_xor(st_want_tri, off);
            st_want_inst += __shfl_xor(st_want_inst, off);
            st_node_it += __shfl_xor(st_node_it, off);
            st_tri_it += __shfl_xor(st_tri_it, off);
            st_inst_it += __shfl_xor(st_inst_it, off);
            st_spec += __shfl_xor(st_spec, off);
            st_drain += __shfl_xor(st_drain, off);
            st_top += __shfl_xor(st_top, off);
        }
        if (lane == 0) {
            atomicAdd(a.bytes, nb);
            atomicAdd(a.bytes + 1, st_iter);
            atomicAdd(a.bytes + 2, st_node);
            atomicAdd(a.bytes + 3, st_tri);
            atomicAdd(a.bytes + 4, st_inst);
            atomicAdd(a.bytes + 5, st_refill);
            atomicAdd(a.bytes + 6, st_live);
            atomicAdd(a.bytes + 7, st_want_node);
            atomicAdd(a.bytes + 8, st_want_tri);
            atomicAdd(a.bytes + 9, st_want_inst);
            atomicAdd(a.bytes + 10, st_node_it);
            atomicAdd(a.bytes + 11, st_tri_it);
            atomicAdd(a.bytes + 12, st_inst_it);
            atomicAdd(a.bytes + 13, st_spec);
            atomicAdd(a.bytes + 14, st_drain);
            atomicAdd(a.bytes + 15, st_top);
        }
    }
}

// A plain streaming copy: the known-size reference for HBM byte counters.
__global__ __launch_bounds__(256) void copy_kernel(uint4* dst, const uint4* src, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ----------------------------------------------------------------- raygen
struct Mat16 {
    float m[16];
};

// shaders/primary_rays.wgsl:50-66. TILED: ray i is pixel k = i % 64 of 8x8 tile
// t = i / 64 (tiles row-major, pixels row-major within a tile; w and h multiples
// of 8), so a wave's 64 rays leave through one screen tile.
template <bool TILED>
__global__ void primary_kernel(Mat16 inv, uint32_t w, uint32_t h, rq_ray* rays) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    uint32_t px = i % w, py = i / w;
    if (TILED) {
        const uint32_t t = i >> 6, k = i & 63u, tiles_x = w >> 3;
        px = (t % tiles_x) * 8u + (k & 7u);
        py = (t / tiles_x) * 8u + (k >> 3);
    }
    float nx = __builtin_fmaf(((float)px + 0.5f) / (float)w, 2.0f, -1.0f);
    float ny = __builtin_fmaf(((float)py + 0.5f) / (float)h, -2.0f, 1.0f);
    float xy[4], A[4], B[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        xy[k] = __builtin_fmaf(inv.m[k], nx, inv.m[4 + k] * ny);
        A[k] = (xy[k] + inv.m[8 + k]) + inv.m[12 + k];
        B[k] = xy[k] + inv.m[12 + k];
    }
    float near[3], far[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        near[k] = A[k] / A[3];
        far[k] = B[k] / B[3];
    }
    float4* rp = reinterpret_cast<float4*>(rays + i);
    rp[0] = make_float4(near[0], near[1], near[2], 0.0f);
    rp[1] = make_float4(far[0] - near[0], far[1] - near[1], far[2] - near[2], 1e20f);
}

__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// reference_path_tracer.wgsl:149-167, 237-256, 466-470. Returns the valid
// mask (bit 0 shadow, bit 1+g bounce g) and writes the rays asked for.
__device__ __forceinline__ uint32_t secondary(const rq_ray* pr, const rq_hit* ph, V3 sun, uint32_t pixel, uint32_t seed,
                                              uint32_t n_gi, rq_ray* out_shadow, rq_ray* const* out_gi) {
    const float4* hp = reinterpret_cast<const float4*>(ph);
    float4 h0 = hp[0], h1 = hp[1], h2 = hp[2];
    if (!(fb(h2.z) & 1u)) return 0u;
    float t = h0.x;
    const float4* rp = reinterpret_cast<const float4*>(pr);
    float4 r0 = rp[0], r1 = rp[1];
    V3 o = mk(r0.x, r0.y, r0.z), d = mk(r1.x, r1.y, r1.z);
    V3 p = mk(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
    V3 ng = mk(h1.w, h2.x, h2.y);
    if (dot3(ng, d) > 0.0f) ng = mk(-ng.x, -ng.y, -ng.z);
    float m = fmx(fmx(fabsf(p.x), fabsf(p.y)), fabsf(p.z));
    float k = 1e-4f + 4e-7f * m;
    V3 start = mk(p.x + ng.x * k, p.y + ng.y * k, p.z + ng.z * k);
    uint32_t valid = dot3(ng, sun) > 0.0f ? 1u : 0u;
    if (out_shadow && valid) {
        float4* q = reinterpret_cast<float4*>(out_shadow);
        q[0] = make_float4(start.x, start.y, start.z, 0.0f);
        q[1] = make_float4(sun.x, sun.y, sun.z, 1e20f);
    }
    float sg = ng.z >= 0.0f ? 1.0f : -1.0f;
    float aa = -1.0f / (sg + ng.z);
    float b = ng.x * ng.y * aa;
    V3 c0 = mk(1.0f + sg * ng.x * ng.x * aa, sg * b, -sg * ng.x);
    V3 c1 = mk(b, sg + ng.y * ng.y * aa, -ng.y);
    V3 c2 = ng;
    uint32_t state = pixel ^ (seed * 0x9E3779B9u);
    for (uint32_t g = 0; g < n_gi; ++g) {
        state = pcg(state);
        float rr = sqrtf((float)(state >> 8u) * (1.0f / 16777216.0f));
        state = pcg(state);
        float phi = 6.2831855f * ((float)(state >> 8u) * (1.0f / 16777216.0f));
        V3 l = mk(rr * cosf(phi), rr * sinf(phi), sqrtf(fmx(0.0f, 1.0f - rr * rr)));
        V3 wi = mk(c0.x * l.x + c1.x * l.y + c2.x * l.z, c0.y * l.x + c1.y * l.y + c2.y * l.z,
                   c0.z * l.x + c1.z * l.y + c2.z * l.z);
        bool ok = dot3(ng, wi) > 0.0f;
        if (ok) valid |= 1u << (1u + g);
        if (out_gi && out_gi[g] && ok) {
            float4* q = reinterpret_cast<float4*>(out_gi[g]);
            q[0] = make_float4(start.x, start.y, start.z, 0.0f);
            q[1] = make_float4(wi.x, wi.y, wi.z, 1e20f);
        }
    }
    return valid;
}

constexpr uint32_t MAX_GI = 4;
constexpr uint32_t SEC_BLOCK = 256;

struct SecArgs {
    const rq_ray* primary;
    const rq_hit* hits;
    uint32_t count;
    float sx, sy, sz;
    uint32_t seed, n_gi;
    uint32_t* block_counts;  // [cat][block]
    uint32_t* counts;        // [cat] totals
    rq_ray* shadow;
    rq_ray* gi;
    uint32_t nblocks;
};

// Block-local exclusive rank of `flag` among the block's threads, and the
// block total; uses wave ballots and a small LDS table.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* wave_tot, uint32_t& total) {
    const uint64_t m = __ballot(flag);
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
    __syncthreads();
    if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < SEC_BLOCK / 64u; ++w) {
        if (w < wave) before += wave_tot[w];
        all += wave_tot[w];
    }
    total = all;
    return before + lane_rank(m);
}

__global__ __launch_bounds__(SEC_BLOCK) void secondary_count(SecArgs a) {
    __shared__ uint32_t wave_tot[SEC_BLOCK / 64u];
    uint32_t i = blockIdx.x * SEC_BLOCK + threadIdx.x;
    uint32_t valid = 0;
    if (i < a.count) valid = secondary(a.primary + i, a.hits + i, mk(a.sx, a.sy, a.sz), i, a.seed, a.n_gi, nullptr, nullptr);
    for (uint32_t c = 0; c <= a.n_gi; ++c) {
        uint32_t total;
        block_rank((valid >> c) & 1u, wave_tot, total);
        if (threadIdx.x == 0) a.block_counts[c * a.nblocks + blockIdx.x] = total;
    }
}

// One workgroup per category: exclusive scan of the block counts in place.
__global__ __launch_bounds__(1024) void secondary_scan(SecArgs a) {
    __shared__ uint32_t part[1024];
    uint32_t c = blockIdx.x;
    uint32_t* bc = a.block_counts + c * a.nblocks;
    uint32_t per = (a.nblocks + 1023u) / 1024u;
    uint32_t b0 = threadIdx.x * per, sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < a.nblocks) sum += bc[b0 + k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < a.nblocks) {
            uint32_t v = bc[b0 + k];
            bc[b0 + k] = run;
            run += v;
        }
    if (threadIdx.x == 1023) a.counts[c] = part[1023];
}

__global__ __launch_bounds__(SEC_BLOCK) void secondary_write(SecArgs a) {
    __shared__ uint32_t wave_tot[SEC_BLOCK / 64u];
    uint32_t i = blockIdx.x * SEC_BLOCK + threadIdx.x;
    uint32_t valid = 0;
    V3 sun = mk(a.sx, a.sy, a.sz);
    if (i < a.count) valid = secondary(a.primary + i, a.hits + i, sun, i, a.seed, a.n_gi, nullptr, nullptr);
    rq_ray* dst_shadow = nullptr;
    rq_ray* dst_gi[MAX_GI] = {nullptr, nullptr, nullptr, nullptr};
    for (uint32_t c = 0; c <= a.n_gi; ++c) {
        uint32_t total;
        uint32_t rank = block_rank((valid >> c) & 1u, wave_tot, total);
        uint32_t at = a.block_counts[c * a.nblocks + blockIdx.x] + rank;
        if ((valid >> c) & 1u) {
            if (c == 0)
                dst_shadow = a.shadow + at;
            else
                dst_gi[c - 1] = a.gi + (uint64_t)(c - 1) * a.count + at;
        }
    }
    if (valid) secondary(a.primary + i, a.hits + i, sun, i, a.seed, a.n_gi, dst_shadow, dst_gi);
}

// ------------------------------------------------------------ FP64 verify
struct VerifyArgs {
    const uint32_t* words;  // REF words: instance records, positions, indices
    const rq_ray* rays;
    const rq_hit* hits;
    uint32_t count;
    unsigned long long* out;  // [0] max rel t err bits, [1] max bary err bits, [2] checked, [3] max scaled t err bits
};

__global__ __launch_bounds__(256) void verify_kernel(VerifyArgs a) {
    __shared__ double st[256], sb[256], ss[256], sw[256];
    __shared__ uint32_t sn[256], sx[256], sc[256];
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    double et = 0.0, eb = 0.0, es = 0.0, ew = 0.0;
    uint32_t n = 0, nx = 0, nw = 0;
    if (i < a.count) {
        const rq_hit h = a.hits[i];
        if (h.flags & 1u) {
            const rq_ray r = a.rays[i];
            double ro[3], rd[3], va[3], vb[3], vc[3];
            double size = 0.0;  // the largest magnitude the f32 computation rounds
            const uint32_t* rec = a.words + 16ull + (uint64_t)h.inst * 32u;
            for (int k = 0; k < 3; ++k) {
                const uint32_t* w = rec + 4 * k;
                ro[k] = (double)bf(w[0]) * r.origin[0] + (double)bf(w[1]) * r.origin[1] +
                        (double)bf(w[2]) * r.origin[2] + (double)bf(w[3]);
                size = fmax(size, fabs((double)bf(w[0]) * r.origin[0]) + fabs((double)bf(w[1]) * r.origin[1]) +
                                      fabs((double)bf(w[2]) * r.origin[2]) + fabs((double)bf(w[3])));
                rd[k] = (double)bf(w[0]) * r.direction[0] + (double)bf(w[1]) * r.direction[1] +
                        (double)bf(w[2]) * r.direction[2];
            }
            {
                uint32_t pos = rec[20], idx = rec[21];
                uint64_t x = (uint64_t)idx + 3ull * h.slot;
                for (int k = 0; k < 3; ++k) {
                    va[k] = bf(a.words[pos + 3ull * a.words[x] + k]);
                    vb[k] = bf(a.words[pos + 3ull * a.words[x + 1] + k]);
                    vc[k] = bf(a.words[pos + 3ull * a.words[x + 2] + k]);
                }
            }
            double ab[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
            double ac[3] = {vc[0] - va[0], vc[1] - va[1], vc[2] - va[2]};
            double p[3] = {rd[1] * ac[2] - rd[2] * ac[1], rd[2] * ac[0] - rd[0] * ac[2], rd[0] * ac[1] - rd[1] * ac[0]};
            double det = ab[0] * p[0] + ab[1] * p[1] + ab[2] * p[2];
            if (det != 0.0) {
                double s[3] = {ro[0] - va[0], ro[1] - va[1], ro[2] - va[2]};
                double q[3] = {s[1] * ab[2] - s[2] * ab[1], s[2] * ab[0] - s[0] * ab[2], s[0] * ab[1] - s[1] * ab[0]};
                double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) / det;
                double v = (rd[0] * q[0] + rd[1] * q[1] + rd[2] * q[2]) / det;
                double t = (ac[0] * q[0] + ac[1] * q[1] + ac[2] * q[2]) / det;
                et = fabs((double)h.t - t) / fmax(fabs(t), 1e-300);
                eb = fmax(fabs((double)h.u - u), fabs((double)h.v - v));
                // The distance error against the size of the numbers it was computed
                // from (transform terms, vertices, hit distance): f32 rounding of those,
                // not of t, bounds a short hit far from the origin.
                for (int k = 0; k < 3; ++k) size = fmax(size, fmax(fabs(va[k]), fmax(fabs(vb[k]), fabs(vc[k]))));
                double dn = sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
                double scale = size + fabs(t) * dn;
                es = fabs((double)h.t - t) * dn / fmax(scale, 1e-300);
                n = 1;
                nx = et > 1e-5 ? 1u : 0u;
                // Away from grazing (cosine to the normal > 0.01) and slivers (edge angle
                // sine > 1e-3), the same error times the cosine, in units of 2^-24 of
                // that size: a count of f32 roundings, small for a sound f32 test.
                double nn[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
                double nl = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                double l1 = sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
                double l2 = sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
                double cosine = fabs(det) / fmax(dn * nl, 1e-300);
                if (cosine >= 1e-2 && nl >= 1e-3 * l1 * l2) {
                    nw = 1;
                    ew = es * cosine * 16777216.0;
                }
            }
        }
    }
    st[threadIdx.x] = et;
    sb[threadIdx.x] = eb;
    ss[threadIdx.x] = es;
    sw[threadIdx.x] = ew;
    sn[threadIdx.x] = n;
    sx[threadIdx.x] = nx;
    sc[threadIdx.x] = nw;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            st[threadIdx.x] = fmax(st[threadIdx.x], st[threadIdx.x + off]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + off]);
            ss[threadIdx.x] = fmax(ss[threadIdx.x], ss[threadIdx.x + off]);
            sw[threadIdx.x] = fmax(sw[threadIdx.x], sw[threadIdx.x + off]);
            sn[threadIdx.x] += sn[threadIdx.x + off];
            sx[threadIdx.x] += sx[threadIdx.x + off];
            sc[threadIdx.x] += sc[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // non-negative doubles order like their bit patterns
        atomicMax(a.out, (unsigned long long)__double_as_longlong(st[0]));
        atomicMax(a.out + 1, (unsigned long long)__double_as_longlong(sb[0]));
        atomicAdd(a.out + 2, (unsigned long long)sn[0]);
        atomicMax(a.out + 3, (unsigned long long)__double_as_longlong(ss[0]));
        atomicAdd(a.out + 4, (unsigned long long)sx[0]);
        atomicMax(a.out + 5, (unsigned long long)__double_as_longlong(sw[0]));
        atomicAdd(a.out + 6, (unsigned long long)sc[0]);
    }
}

}  // namespace dev

// ------------------------------------------------------------- host side
hipError_t ray_order(const rq_ray* rays, uint32_t n, const uint32_t* count_ptr, const float bounds[6], uint32_t* keys_in,
                     uint32_t* keys_out, uint32_t* vals_in, uint32_t* order, void* temp, size_t& temp_bytes,
                     hipStream_t s);

namespace {

// Launches shorter than this are traced in submission order.
constexpr uint32_t REORDER_MIN = 65536;

struct Tune {
    uint32_t block, stack, refill_min, grid_per_cu;
};

// Device memory a launch uses besides its inputs and outputs, owned per HIP
// stream: launches on one stream run in order, so they can share it, and
// launches on different streams never do.
struct StreamRes {
    hipStream_t stream = nullptr;
    uint32_t* counter = nullptr;  // the persistent kernels' ray cursor
    uint32_t* spill = nullptr;    // per-lane stack spill area
    size_t spill_lanes = 0;
    void* scratch = nullptr;      // secondary-ray block counts
    size_t scratch_bytes = 0;
    void* table = nullptr;        // batched launches: batch table and prefix
    size_t table_bytes = 0;
    void* order = nullptr;        // ray reordering: keys, indices and sort space
    size_t order_bytes = 0;
};

struct Ctx {
    std::mutex mu;
    int device = -1;
    int cus = 0;
    std::vector<StreamRes> res;
    rq_tuning tuning[LAYOUT_COUNT][2];
    rq_tuning batch_tuning[LAYOUT_COUNT];
    bool tuned = false;
};
Ctx g_ctx;

// A layout traces a scene of its kind: REF, PLAIN and CW8 the GI words, RS and
// RS_CW8 the reflection words.
bool layout_ok(const rq_scene* s, int layout) {
    if (!s || layout < 0 || layout >= LAYOUT_COUNT) return false;
    const bool rs = layout == RQ_LAYOUT_RS_BVH4Q8 || layout == RQ_LAYOUT_RS_CW8_Q8;
    return s->kind == RQ_SCENE_REFLECTION ? rs : !rs;
}

rq_status hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RQ_ERR_OUT_OF_MEMORY : RQ_ERR_DEVICE;
}
#define HIP_TRY(x, what)                         \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

void default_tuning(rq_tuning& t, int layout, int any) {
    std::memset(&t, 0, sizeof t);
    t.block = 256;
    t.stack = 16;
    t.persistent = 1;
    t.waves_per_eu = 0;
    t.grid_per_cu = 0;
    t.refill = 16u;  // refill threshold (lanes): the autotuner's pick for CW8 on the measured frames (PLAIN: 16 or 32)
    t.speculate = layout == RQ_LAYOUT_CW8_Q8 || layout == RQ_LAYOUT_RS_CW8_Q8 ? 1u : 0u;  // 2-3% on the measured frames
    t.drain = 1u;                                        // 0.4-1% faster on the measured frames
}

void init_tuning() {
    if (g_ctx.tuned) return;
    for (int l = 0; l < LAYOUT_COUNT; ++l) {
        for (int m = 0; m < 2; ++m) default_tuning(g_ctx.tuning[l][m], l, m);
        default_tuning(g_ctx.batch_tuning[l], l, 0);
    }
    g_ctx.tuned = true;
}

rq_status ensure_ctx() {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (g_ctx.device == dev) return RQ_OK;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    g_ctx.device = dev;
    g_ctx.cus = prop.multiProcessorCount;
    init_tuning();
    return RQ_OK;
}

template <int L, int B, int S, int AM, bool C = false, bool BT = false, bool WT = false>
void* kernel_ptr() {
    return reinterpret_cast<void*>(&dev::trace_kernel<L, B, S, AM, C, BT, WT>);
}

typedef void (*Launch)(dim3, dim3, hipStream_t, const dev::TraceArgs&);
template <int L, int B, int S, int AM, bool C = false, bool BT = false, bool WT = false>
void launch_one(dim3 g, dim3 b, hipStream_t s, const dev::TraceArgs& a) {
    hipLaunchKernelGGL((dev::trace_kernel<L, B, S, AM, C, BT, WT>), g, b, 0, s, a);
}

struct KernelEntry {
    uint32_t block, stack;
    Launch fn[LAYOUT_COUNT][2];  // by layout, any-hit
    void* ptr[LAYOUT_COUNT][2];
    Launch bfn[LAYOUT_COUNT];    // batched, mode per batch
    void* bptr[LAYOUT_COUNT];
    Launch rfn[LAYOUT_COUNT];    // ranges of one scene, mode per range
    void* rptr[LAYOUT_COUNT];
};

template <int L, int B, int S>
void fill_layout(KernelEntry& e) {
    e.fn[L][0] = &launch_one<L, B, S, 0>;
    e.fn[L][1] = &launch_one<L, B, S, 1>;
    e.ptr[L][0] = kernel_ptr<L, B, S, 0>();
    e.ptr[L][1] = kernel_ptr<L, B, S, 1>();
    e.bfn[L] = &launch_one<L, B, S, 2, false, true>;
    e.bptr[L] = kernel_ptr<L, B, S, 2, false, true>();
    e.rfn[L] = &launch_one<L, B, S, 2>;
    e.rptr[L] = kernel_ptr<L, B, S, 2>();
}

template <int B, int S>
KernelEntry make_entry() {
    KernelEntry e;
    e.block = B;
    e.stack = S;
    fill_layout<dev::REF, B, S>(e);
    fill_layout<dev::PLAIN, B, S>(e);
    fill_layout<dev::CW8, B, S>(e);
    fill_layout<dev::RS, B, S>(e);
    fill_layout<dev::RSW, B, S>(e);
    return e;
}

const std::vector<KernelEntry>& kernels() {
    static std::vector<KernelEntry> k = {make_entry<128, 8>(),  make_entry<128, 16>(), make_entry<128, 32>(),
                                         make_entry<256, 8>(),  make_entry<256, 16>(), make_entry<256, 32>(),
                                         make_entry<64, 16>(),  make_entry<64, 32>()};
    return k;
}

const KernelEntry& count_kernel() {
    static KernelEntry e = [] {
        KernelEntry k;
        std::memset(&k, 0, sizeof k);
        k.block = 256;
        k.stack = 16;
        auto put = [&](auto lc) {
            constexpr int L = decltype(lc)::value;
            k.fn[L][0] = &launch_one<L, 256, 16, 0, true>;
            k.fn[L][1] = &launch_one<L, 256, 16, 1, true>;
            k.ptr[L][0] = kernel_ptr<L, 256, 16, 0, true>();
            k.ptr[L][1] = kernel_ptr<L, 256, 16, 1, true>();
        };
        put(std::integral_constant<int, dev::REF>());
        put(std::integral_constant<int, dev::PLAIN>());
        put(std::integral_constant<int, dev::CW8>());
        put(std::integral_constant<int, dev::RS>());
        put(std::integral_constant<int, dev::RSW>());
        return k;
    }();
    return e;
}

// The watertight leaf test (RQ_WATERTIGHT): REF, PLAIN and CW8, closest and any hit.
const KernelEntry& watertight_kernel() {
    static KernelEntry e = [] {
        KernelEntry k;
        std::memset(&k, 0, sizeof k);
        k.block = 256;
        k.stack = 16;
        auto put = [&](auto lc) {
            constexpr int L = decltype(lc)::value;
            k.fn[L][0] = &launch_one<L, 256, 16, 0, false, false, true>;
            k.fn[L][1] = &launch_one<L, 256, 16, 1, false, false, true>;
            k.ptr[L][0] = kernel_ptr<L, 256, 16, 0, false, false, true>();
            k.ptr[L][1] = kernel_ptr<L, 256, 16, 1, false, false, true>();
        };
        put(std::integral_constant<int, dev::REF>());
        put(std::integral_constant<int, dev::PLAIN>());
        put(std::integral_constant<int, dev::CW8>());
        return k;
    }();
    return e;
}

const KernelEntry* find_kernel(uint32_t block, uint32_t stack) {
    for (const auto& k : kernels())
        if (k.block == block && k.stack == stack) return &k;
    return nullptr;
}

rq_status res_for(hipStream_t s, StreamRes** out) {
    for (auto& r : g_ctx.res)
        if (r.stream == s) {
            *out = &r;
            return RQ_OK;
        }
    StreamRes r;
    r.stream = s;
    HIP_TRY(hipMalloc(&r.counter, CURSOR_BYTES), "hipMalloc counter");
    g_ctx.res.push_back(r);
    *out = &g_ctx.res.back();
    return RQ_OK;
}

rq_status ensure_spill(StreamRes& r, size_t lanes) {
    if (r.spill && r.spill_lanes >= lanes) return RQ_OK;
    if (r.spill) {
        HIP_TRY(hipStreamSynchronize(r.stream), "hipStreamSynchronize");
        HIP_TRY(hipFree(r.spill), "hipFree spill");
        r.spill = nullptr;
    }
    HIP_TRY(hipMalloc(&r.spill, lanes * dev::SPILL_WORDS * sizeof(uint32_t)), "hipMalloc spill");
    r.spill_lanes = lanes;
    return RQ_OK;
}

rq_status ensure_scratch(StreamRes& r, size_t bytes) {
    if (r.scratch && r.scratch_bytes >= bytes) return RQ_OK;
    if (r.scratch) {
        HIP_TRY(hipStreamSynchronize(r.stream), "hipStreamSynchronize");
        HIP_TRY(hipFree(r.scratch), "hipFree scratch");
        r.scratch = nullptr;
    }
    HIP_TRY(hipMalloc(&r.scratch, bytes), "hipMalloc scratch");
    r.scratch_bytes = bytes;
    return RQ_OK;
}

rq_status ensure_table(StreamRes& r, size_t bytes) {
    if (r.table && r.table_bytes >= bytes) {
        // the previous batched launch on this stream may still read it
        HIP_TRY(hipStreamSynchronize(r.stream), "hipStreamSynchronize");
        return RQ_OK;
    }
    if (r.table) {
        HIP_TRY(hipStreamSynchronize(r.stream), "hipStreamSynchronize");
        HIP_TRY(hipFree(r.table), "hipFree table");
        r.table = nullptr;
    }
    HIP_TRY(hipMalloc(&r.table, bytes), "hipMalloc table");
    r.table_bytes = bytes;
    return RQ_OK;
}

rq_status ensure_order(StreamRes& r, size_t bytes) {
    if (r.order && r.order_bytes >= bytes) return RQ_OK;
    if (r.order) {
        HIP_TRY(hipStreamSynchronize(r.stream), "hipStreamSynchronize");
        HIP_TRY(hipFree(r.order), "hipFree order");
        r.order = nullptr;
    }
    HIP_TRY(hipMalloc(&r.order, bytes), "hipMalloc order");
    r.order_bytes = bytes;
    return RQ_OK;
}

void release_stream(hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    for (size_t i = 0; i < g_ctx.res.size(); ++i)
        if (g_ctx.res[i].stream == s) {
            StreamRes& r = g_ctx.res[i];
            (void)hipStreamSynchronize(s);
            if (r.counter) (void)hipFree(r.counter);
            if (r.spill) (void)hipFree(r.spill);
            if (r.scratch) (void)hipFree(r.scratch);
            if (r.table) (void)hipFree(r.table);
            if (r.order) (void)hipFree(r.order);
            g_ctx.res.erase(g_ctx.res.begin() + i);
            return;
        }
}

template <typename T>
rq_status upload(void* base, uint64_t& at, const std::vector<T>& v) {
    if (!v.empty()) HIP_TRY(hipMemcpy((char*)base + at, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy scene");
    at += (v.size() * sizeof(T) + 255) & ~(size_t)255;
    return RQ_OK;
}
template <typename T>
size_t padded(const std::vector<T>& v) {
    return (v.size() * sizeof(T) + 255) & ~(size_t)255;
}

}  // namespace
}  // namespace rq

using namespace rq;

extern "C" int rq_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" rq_status rq_scene_upload(rq_scene* scene, uint32_t layout_mask) {
    if (!scene) {
        set_error("rq_scene_upload: null scene");
        return RQ_ERR_INVALID_INPUT;
    }
    for (int l = 0; l < LAYOUT_COUNT; ++l)
        if ((layout_mask & (1u << l)) && !layout_ok(scene, l)) {
            set_error(scene->kind == RQ_SCENE_REFLECTION
                          ? "rq_scene_upload: a reflection scene has only the RS layouts"
                          : "rq_scene_upload: the RS layouts need a reflection scene");
            return RQ_ERR_INVALID_INPUT;
        }
    if (layout_mask & ~((1u << LAYOUT_COUNT) - 1u)) {
        set_error("rq_scene_upload: unknown layout bits");
        return RQ_ERR_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    rq_status s = ensure_ctx();
    if (s != RQ_OK) return s;
    std::string err;
    if ((s = sync_host_words(scene)) != RQ_OK) return s;
    const bool want_plain = ((layout_mask & 2u) && !scene->dev[1].ready) ||
                            ((layout_mask & 4u) && !scene->dev[2].ready && scene->cw8.inst.empty());
    if (want_plain && scene->plain.inst.empty() && scene->words[0] > 0) {
        if (!build_plain(scene->words, scene->plain, err)) {
            set_error("plain relayout: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
    }
    if ((layout_mask & 4u) && !scene->dev[2].ready && scene->cw8.inst.empty() && scene->words[0] > 0) {
        if (!build_cw8(scene->words, scene->plain, scene->cw8, err)) {
            set_error("cw8 relayout: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
    }
    if ((layout_mask & 16u) && !scene->dev[RQ_LAYOUT_RS_CW8_Q8].ready && scene->rs_cw8.inst.empty()) {
        if (!build_rs_cw8(scene->words, scene->rs_cw8, err)) {
            set_error("rs cw8 relayout: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
    }
    for (int l = 0; l < LAYOUT_COUNT; ++l) {
        if (!(layout_mask & (1u << l)) || scene->dev[l].ready) continue;
        DeviceLayout& d = scene->dev[l];
        if (l == RQ_LAYOUT_RS_CW8_Q8) {  // reflection words, CW8 top level, CW8 BLAS nodes
            HostLayout& h = scene->rs_cw8;
            if (h.roots.empty()) h.roots.assign(24, 0u);
            d.bytes = padded(h.inst) + padded(h.top) + padded(h.nodes) + padded(h.roots);
            HIP_TRY(hipMalloc(&d.base, d.bytes), "hipMalloc scene");
            uint64_t at = 0;
            d.off[0] = at; if ((s = upload(d.base, at, h.inst)) != RQ_OK) return s;
            d.off[1] = at; if ((s = upload(d.base, at, h.top)) != RQ_OK) return s;
            d.off[2] = at; if ((s = upload(d.base, at, h.nodes)) != RQ_OK) return s;
            d.off[3] = at; if ((s = upload(d.base, at, h.roots)) != RQ_OK) return s;
            d.off[4] = at;
            d.count[0] = h.instance_count;
            d.count[1] = h.top_on ? 1u : 0u;
            d.count[2] = h.top_root;
            d.count[3] = (uint32_t)h.top.size();
        } else if (l == 0 || l == 3) {  // the reference words verbatim
            d.bytes = padded(scene->words);
            HIP_TRY(hipMalloc(&d.base, d.bytes), "hipMalloc scene");
            uint64_t at = 0;
            s = upload(d.base, at, scene->words);
            if (s != RQ_OK) return s;
        } else {
            HostLayout& h = l == 1 ? scene->plain : scene->cw8;
            if (h.inst.empty()) h.inst.assign(32, 0u);
            if (h.nodes.empty()) h.nodes.assign(32, 0u);
            if (h.tris.empty()) h.tris.assign(TRI_FLOATS, 0.0f);
            if (h.tri_ids.empty()) h.tri_ids.assign(TRI_ID_WORDS, 0u);
            if (h.verts.empty()) h.verts.assign(TRI_FLOATS, 0.0f);
            if (h.roots.empty()) h.roots.assign(24, 0u);
            d.bytes = padded(h.inst) + padded(h.top) + padded(h.nodes) + padded(h.tris) + padded(h.tri_ids) +
                      padded(h.verts) + padded(h.roots);
            HIP_TRY(hipMalloc(&d.base, d.bytes), "hipMalloc scene");
            uint64_t at = 0;
            d.off[0] = at; if ((s = upload(d.base, at, h.inst)) != RQ_OK) return s;
            d.off[1] = at; if ((s = upload(d.base, at, h.top)) != RQ_OK) return s;
            d.off[2] = at; if ((s = upload(d.base, at, h.nodes)) != RQ_OK) return s;
            d.off[3] = at; if ((s = upload(d.base, at, h.tris)) != RQ_OK) return s;
            d.off[4] = at; if ((s = upload(d.base, at, h.tri_ids)) != RQ_OK) return s;
            d.off[5] = at; if ((s = upload(d.base, at, h.verts)) != RQ_OK) return s;
            d.off[6] = at; if ((s = upload(d.base, at, h.roots)) != RQ_OK) return s;
            d.count[0] = h.instance_count;
            d.count[1] = h.top_on ? 1u : 0u;
            d.count[2] = h.top_root;
            d.count[3] = (uint32_t)h.top.size();
            d.mesh_base = scene->plain.mesh_base;
        }
        d.ready = true;
    }
    scene->device = g_ctx.device;
    return RQ_OK;
}

extern "C" rq_status rq_scene_download_part(const rq_scene* scene, rq_layout layout, uint32_t part, void* dst,
                                            size_t capacity, size_t* bytes) {
    if (!scene || !bytes || (int)layout < 0 || (int)layout >= LAYOUT_COUNT || part > 6 || part == 5 ||
        (capacity && !dst)) {
        set_error("rq_scene_download_part: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    const DeviceLayout& d = scene->dev[layout];
    if (!d.ready) {
        set_error("rq_scene_download_part: layout not uploaded");
        return RQ_ERR_INVALID_INPUT;
    }
    size_t at = 0, n = d.bytes;
    if (layout == RQ_LAYOUT_PLAIN_F32 || layout == RQ_LAYOUT_CW8_Q8) {
        const uint32_t p = part == 6 ? 5u : part;  // part 6: the vertex records, stored after part 4
        at = d.off[p];
        n = d.off[p + 1] - at;  // part 6 (p 5) ends where the CW8 root copies start
    } else if (layout == RQ_LAYOUT_RS_CW8_Q8) {  // parts: words, top, nodes, root copies
        if (part > 3) {
            set_error("rq_scene_download_part: the RS CW8 layout has four parts");
            return RQ_ERR_INVALID_INPUT;
        }
        at = d.off[part];
        n = d.off[part + 1] - at;
    } else if (part != 0) {
        set_error("rq_scene_download_part: the word layouts have one part");
        return RQ_ERR_INVALID_INPUT;
    }
    *bytes = n;
    if (capacity < n) return RQ_OK;  // size query
    HIP_TRY(hipMemcpy(dst, reinterpret_cast<const char*>(d.base) + at, n, hipMemcpyDeviceToHost), "download part");
    return RQ_OK;
}

extern "C" rq_status rq_scene_device_bytes(const rq_scene* scene, rq_layout layout, size_t* bytes) {
    if (!scene || !bytes || (int)layout < 0 || (int)layout >= LAYOUT_COUNT) {
        set_error("rq_scene_device_bytes: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    *bytes = scene->dev[layout].ready ? scene->dev[layout].bytes : 0;
    return RQ_OK;
}

extern "C" void rq_scene_destroy(rq_scene* scene) {
    if (!scene) return;
    for (auto& d : scene->dev)
        if (d.base) (void)hipFree(d.base);
    for (auto& p : scene->refit_plans) {
        if (p.order) (void)hipFree(p.order);
        if (p.boxes) (void)hipFree(p.boxes);
    }
    if (scene->refit_scratch) (void)hipFree(scene->refit_scratch);
    delete scene;
}

extern "C" rq_status rq_get_tuning(rq_layout layout, uint32_t mode, rq_tuning* out) {
    if (!out || (int)layout < 0 || (int)layout >= LAYOUT_COUNT) {
        set_error("rq_get_tuning: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    init_tuning();
    *out = g_ctx.tuning[layout][mode ? 1 : 0];
    return RQ_OK;
}

extern "C" rq_status rq_set_tuning(rq_layout layout, uint32_t mode, const rq_tuning* t) {
    if (!t || (int)layout < 0 || (int)layout >= LAYOUT_COUNT || !find_kernel(t->block, t->stack) || t->refill == 0 ||
        t->refill > 64) {
        set_error("rq_set_tuning: unsupported block/stack/refill combination");
        return RQ_ERR_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    init_tuning();
    g_ctx.tuning[layout][mode ? 1 : 0] = *t;
    return RQ_OK;
}

namespace rq {
namespace {
dev::SceneView scene_view(const rq_scene* scene, rq_layout layout) {
    dev::SceneView v;
    std::memset(&v, 0, sizeof v);
    const DeviceLayout& d = scene->dev[layout];
    if (layout == RQ_LAYOUT_REF_BVH4Q8 || layout == RQ_LAYOUT_RS_BVH4Q8) {
        v.words = reinterpret_cast<const uint32_t*>(d.base);
        v.top = v.words + scene->words[1];
        v.n_inst = scene->words[0];
        v.top_on = scene->words[1] != 0;
        v.top_root = scene->words[2];
    } else if (layout == RQ_LAYOUT_RS_CW8_Q8) {
        // the reflection words (records' roots rewritten), CW8 top level, CW8 BLAS nodes
        char* b = reinterpret_cast<char*>(d.base);
        v.words = reinterpret_cast<const uint32_t*>(b + d.off[0]);
        v.top = reinterpret_cast<const uint32_t*>(b + d.off[1]);
        v.nodes = reinterpret_cast<const uint32_t*>(b + d.off[2]);
        v.roots = reinterpret_cast<const uint32_t*>(b + d.off[3]);
        v.n_inst = d.count[0];
        v.top_on = d.count[1] != 0;
        v.top_root = d.count[2];
    } else {
        char* b = reinterpret_cast<char*>(d.base);
        v.inst = reinterpret_cast<const uint32_t*>(b + d.off[0]);
        v.top = reinterpret_cast<const uint32_t*>(b + d.off[1]);
        v.nodes = reinterpret_cast<const uint32_t*>(b + d.off[2]);
        v.tris = reinterpret_cast<const float*>(b + d.off[3]);
        v.tri_ids = reinterpret_cast<const uint32_t*>(b + d.off[4]);
        v.verts = reinterpret_cast<const float*>(b + d.off[5]);
        v.roots = reinterpret_cast<const uint32_t*>(b + d.off[6]);
        v.n_inst = d.count[0];
        v.top_on = d.count[1] != 0;
        v.top_root = d.count[2];
    }
    return v;
}

}  // namespace
}  // namespace rq

// `fixed`, when set, is device memory the launch uses instead of the stream's
// own (no allocation): the ray cursor, then the stack spill area of at most
// fixed->lanes lanes; the grid is capped to that many lanes.
rq_status rq::launch_trace(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count, uint32_t mode,
                           rq_hit* hits, uint32_t* vis_bits, void* stream, unsigned long long* bytes,
                           const uint32_t* count_ptr, const rq::FixedScratch* fixed) {
    const bool reorder = (mode & RQ_REORDER) != 0;
    const bool watertight = (mode & RQ_WATERTIGHT) != 0;
    mode &= ~(RQ_REORDER | RQ_WATERTIGHT);
    if (!scene || !layout_ok(scene, layout) || mode > 2 || (count && !rays)) {
        set_error("rq_trace: bad argument (or a layout of the other scene kind)");
        return RQ_ERR_INVALID_INPUT;
    }
    if (watertight && (bytes || layout > RQ_LAYOUT_CW8_Q8)) {
        set_error("rq_trace: RQ_WATERTIGHT is for the REF, PLAIN and CW8 layouts (not counted launches)");
        return RQ_ERR_UNSUPPORTED;
    }
    if (fixed && (reorder || bytes || count_ptr)) {
        set_error("rq_trace_noalloc: RQ_REORDER is not available without allocation (its sort space grows with "
                  "the ray count)");
        return RQ_ERR_UNSUPPORTED;
    }
    if (!scene->dev[layout].ready) {
        set_error("rq_trace: layout not uploaded (rq_scene_upload)");
        return RQ_ERR_INVALID_INPUT;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    rq_status st = ensure_ctx();
    if (st != RQ_OK) return st;
    if (scene->device != g_ctx.device) {
        set_error("rq_trace: scene lives on another device");
        return RQ_ERR_INVALID_INPUT;
    }
    if (vis_bits && count) {
        uint64_t words = (count + 31u) / 32u;
        hipLaunchKernelGGL(dev::fill_u32, dim3((unsigned)std::min<uint64_t>((words + 255) / 256, 4096)), dim3(256), 0, s,
                           vis_bits, 0u, words);
    }
    if (count == 0) return RQ_OK;
    const int any = mode != 0 ? 1 : 0;
    const rq_tuning& t = g_ctx.tuning[layout][any];
    const KernelEntry* k = bytes ? &count_kernel() : watertight ? &watertight_kernel() : find_kernel(t.block, t.stack);
    if (!k) {
        set_error("rq_trace: no kernel for the tuning");
        return RQ_ERR_INVALID_INPUT;
    }
    int per_cu = (int)t.grid_per_cu;
    if (per_cu == 0) {
        int blocks = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k->ptr[layout][any], (int)k->block, 0),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        per_cu = std::max(blocks, 1);
    }
    uint64_t grid = (uint64_t)g_ctx.cus * per_cu;
    uint64_t need = ((uint64_t)count + k->block - 1) / k->block;
    grid = std::max<uint64_t>(1, std::min(grid, need));
    uint32_t* counter = nullptr;
    uint32_t* spill = nullptr;
    if (fixed) {
        if (fixed->lanes < k->block) {
            set_error("rq_trace_noalloc: the fixed scratch holds fewer lanes than one workgroup");
            return RQ_ERR_UNSUPPORTED;
        }
        grid = std::min<uint64_t>(grid, fixed->lanes / k->block);
        counter = fixed->counter;
        spill = fixed->spill;
    } else {
        StreamRes* res = nullptr;
        if ((st = res_for(s, &res)) != RQ_OK) return st;
        if ((st = ensure_spill(*res, (size_t)(grid * k->block))) != RQ_OK) return st;
        counter = res->counter;
        spill = res->spill;
    }
    hipLaunchKernelGGL(dev::fill_u32, dim3(16), dim3(256), 0, s, counter, 0u, (uint64_t)(CURSOR_BYTES / 4));

    dev::TraceArgs a;
    std::memset(&a, 0, sizeof a);
    a.sv = scene_view(scene, layout);
    a.rays = rays;
    a.hits = hits;
    a.vis = vis_bits;
    a.count = count;
    a.counter = counter;
    a.spill = spill;
    a.count_ptr = count_ptr;
    a.refill_min = t.refill ? t.refill : 1u;
    a.speculate = t.speculate;
    a.drain = t.drain;
    a.bytes = bytes;
    if (reorder && count >= REORDER_MIN) {
        // keys, indices, sorted keys, order, then the sort's own space
        const size_t arr = ((size_t)count * 4 + 255) & ~(size_t)255;
        size_t temp = 0;
        HIP_TRY(ray_order(rays, count, count_ptr, scene->bounds, nullptr, nullptr, nullptr, nullptr, nullptr, temp, s),
                "ray order (size)");
        StreamRes* res = nullptr;
        if ((st = res_for(s, &res)) != RQ_OK) return st;
        if ((st = ensure_order(*res, 4 * arr + temp)) != RQ_OK) return st;
        char* b = reinterpret_cast<char*>(res->order);
        uint32_t* keys = reinterpret_cast<uint32_t*>(b);
        uint32_t* vals = reinterpret_cast<uint32_t*>(b + arr);
        uint32_t* sorted = reinterpret_cast<uint32_t*>(b + 2 * arr);
        uint32_t* order = reinterpret_cast<uint32_t*>(b + 3 * arr);
        HIP_TRY(ray_order(rays, count, count_ptr, scene->bounds, keys, sorted, vals, order, b + 4 * arr, temp, s),
                "ray order");
        a.order = order;
    }
    k->fn[layout][any](dim3((unsigned)grid), dim3(k->block), s, a);
    HIP_TRY(hipGetLastError(), "trace launch");
    return RQ_OK;
}

extern "C" rq_status rq_trace(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count, uint32_t mode,
                              rq_hit* hits, uint32_t* vis_bits, void* stream) {
    return launch_trace(scene, layout, rays, count, mode, hits, vis_bits, stream, nullptr);
}

extern "C" rq_status rq_trace_indirect(rq_scene* scene, rq_layout layout, const rq_ray* rays,
                                       const uint32_t* count_dev, uint32_t max_count, uint32_t mode, rq_hit* hits,
                                       uint32_t* vis_bits, void* stream) {
    if (!count_dev) {
        set_error("rq_trace_indirect: null count");
        return RQ_ERR_INVALID_INPUT;
    }
    return launch_trace(scene, layout, rays, max_count, mode, hits, vis_bits, stream, nullptr, count_dev);
}

extern "C" rq_status rq_trace_ranges(rq_scene* scene, rq_layout layout, const rq_range* ranges,
                                     uint32_t range_count, void* stream) {
    if (!scene || !ranges || !layout_ok(scene, layout) || range_count == 0 ||
        range_count > RQ_MAX_RANGES) {
        set_error("rq_trace_ranges: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    if (!scene->dev[layout].ready) {
        set_error("rq_trace_ranges: layout not uploaded (rq_scene_upload)");
        return RQ_ERR_INVALID_INPUT;
    }
    dev::TraceArgs a;
    std::memset(&a, 0, sizeof a);
    uint64_t total = 0;
    for (uint32_t j = 0; j < range_count; ++j) {
        const rq_range& q = ranges[j];
        if (q.mode > 2 || (q.count && !q.rays)) {
            set_error("rq_trace_ranges: range " + std::to_string(j) + " is malformed");
            return RQ_ERR_INVALID_INPUT;
        }
        a.rrays[j] = q.rays;
        a.rhits[j] = q.hits;
        a.rvis[j] = q.vis_bits;
        a.rstart[j] = (uint32_t)total;
        if (q.mode != 0) a.rany |= 1u << j;
        total += q.count;
    }
    if (total >= 0xffffffffull) {
        set_error("rq_trace_ranges: more than 2^32 rays");
        return RQ_ERR_INVALID_INPUT;
    }
    for (uint32_t j = range_count; j < 5; ++j) a.rstart[j] = (uint32_t)total;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    rq_status st = ensure_ctx();
    if (st != RQ_OK) return st;
    if (scene->device != g_ctx.device) {
        set_error("rq_trace_ranges: scene lives on another device");
        return RQ_ERR_INVALID_INPUT;
    }
    for (uint32_t j = 0; j < range_count; ++j) {
        if (!ranges[j].vis_bits || !ranges[j].count) continue;
        uint64_t words = (ranges[j].count + 31u) / 32u;
        hipLaunchKernelGGL(dev::fill_u32, dim3((unsigned)std::min<uint64_t>((words + 255) / 256, 4096)), dim3(256), 0,
                           s, ranges[j].vis_bits, 0u, words);
    }
    if (total == 0) return RQ_OK;
    const rq_tuning& t = g_ctx.tuning[layout][0];
    const KernelEntry* k = find_kernel(t.block, t.stack);
    if (!k) {
        set_error("rq_trace_ranges: no kernel for the tuning");
        return RQ_ERR_INVALID_INPUT;
    }
    int per_cu = (int)t.grid_per_cu;
    if (per_cu == 0) {
        int blocks = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k->rptr[layout], (int)k->block, 0),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        per_cu = std::max(blocks, 1);
    }
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)g_ctx.cus * per_cu,
                                                             (total + k->block - 1) / k->block));
    StreamRes* res = nullptr;
    if ((st = res_for(s, &res)) != RQ_OK) return st;
    if ((st = ensure_spill(*res, (size_t)(grid * k->block))) != RQ_OK) return st;
    hipLaunchKernelGGL(dev::fill_u32, dim3(16), dim3(256), 0, s, res->counter, 0u, (uint64_t)(CURSOR_BYTES / 4));
    a.sv = scene_view(scene, layout);
    a.count = (uint32_t)total;
    a.counter = res->counter;
    a.spill = res->spill;
    a.refill_min = t.refill ? t.refill : 1u;
    a.speculate = t.speculate;
    a.drain = t.drain;
    k->rfn[layout](dim3((unsigned)grid), dim3(k->block), s, a);
    HIP_TRY(hipGetLastError(), "trace ranges launch");
    return RQ_OK;
}

extern "C" rq_status rq_trace_batches(rq_scene* const* scenes, uint32_t scene_count, rq_layout layout,
                                      const rq_batch* batches, uint32_t batch_count, void* stream) {
    return rq::trace_batches_impl(scenes, scene_count, layout, batches, batch_count, stream, -1);
}

rq_status rq::trace_batches_impl(rq_scene* const* scenes, uint32_t scene_count, rq_layout layout,
                                 const rq_batch* batches, uint32_t batch_count, void* stream, int speculate) {
    if (!scenes || !batches || (int)layout < 0 || (int)layout >= LAYOUT_COUNT || batch_count == 0) {
        set_error("rq_trace_batches: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    std::vector<dev::BatchDev> table(batch_count);
    std::vector<uint32_t> prefix(batch_count + 1, 0);
    uint64_t total = 0;
    for (uint32_t b = 0; b < batch_count; ++b) {
        const rq_batch& q = batches[b];
        if (q.scene >= scene_count || !layout_ok(scenes[q.scene], layout) || !scenes[q.scene]->dev[layout].ready ||
            q.mode > 2 ||
            (q.count && !q.rays)) {
            set_error("rq_trace_batches: batch " + std::to_string(b) + " names a missing scene or layout");
            return RQ_ERR_INVALID_INPUT;
        }
        table[b].sv = scene_view(scenes[q.scene], layout);
        table[b].rays = q.rays;
        table[b].hits = q.hits;
        table[b].vis = q.vis_bits;
        table[b].mode = q.mode;
        table[b].count = q.count;
        prefix[b] = (uint32_t)total;
        total += q.count;
    }
    if (total >= 0xffffffffull) {
        set_error("rq_trace_batches: more than 2^32 rays");
        return RQ_ERR_INVALID_INPUT;
    }
    prefix[batch_count] = (uint32_t)total;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    rq_status st = ensure_ctx();
    if (st != RQ_OK) return st;
    for (uint32_t b = 0; b < batch_count; ++b) {
        if (!batches[b].vis_bits || !batches[b].count) continue;
        uint64_t words = (batches[b].count + 31u) / 32u;
        hipLaunchKernelGGL(dev::fill_u32, dim3((unsigned)std::min<uint64_t>((words + 255) / 256, 4096)), dim3(256), 0,
                           s, batches[b].vis_bits, 0u, words);
    }
    if (total == 0) return RQ_OK;
    const rq_tuning& t = g_ctx.batch_tuning[layout];
    const KernelEntry* k = find_kernel(t.block, t.stack);
    if (!k) {
        set_error("rq_trace_batches: no kernel for the tuning");
        return RQ_ERR_INVALID_INPUT;
    }
    int per_cu = (int)t.grid_per_cu;
    if (per_cu == 0) {
        int blocks = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k->bptr[layout], (int)k->block, 0),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        per_cu = std::max(blocks, 1);
    }
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)g_ctx.cus * per_cu,
                                                             (total + k->block - 1) / k->block));
    StreamRes* res = nullptr;
    if ((st = res_for(s, &res)) != RQ_OK) return st;
    if ((st = ensure_spill(*res, (size_t)(grid * k->block))) != RQ_OK) return st;
    size_t tbytes = table.size() * sizeof(dev::BatchDev), pbytes = prefix.size() * sizeof(uint32_t);
    if ((st = ensure_table(*res, tbytes + pbytes + 256)) != RQ_OK) return st;
    char* tb = reinterpret_cast<char*>(res->table);
    HIP_TRY(hipMemcpyAsync(tb, table.data(), tbytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync batch table");
    size_t poff = (tbytes + 255) & ~(size_t)255;
    HIP_TRY(hipMemcpyAsync(tb + poff, prefix.data(), pbytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync prefix");
    hipLaunchKernelGGL(dev::fill_u32, dim3(16), dim3(256), 0, s, res->counter, 0u, (uint64_t)(CURSOR_BYTES / 4));
    dev::TraceArgs a;
    std::memset(&a, 0, sizeof a);
    a.count = (uint32_t)total;
    a.counter = res->counter;
    a.spill = res->spill;
    a.batches = reinterpret_cast<const dev::BatchDev*>(tb);
    a.prefix = reinterpret_cast<const uint32_t*>(tb + poff);
    a.n_batches = batch_count;
    a.refill_min = t.refill ? t.refill : 1u;
    a.speculate = speculate < 0 ? t.speculate : (uint32_t)speculate;
    a.drain = t.drain;
    k->bfn[layout](dim3((unsigned)grid), dim3(k->block), s, a);
    HIP_TRY(hipGetLastError(), "batched trace launch");
    // The table is read by the kernel: the next call on this stream may overwrite it only after.
    return RQ_OK;
}

extern "C" rq_status rq_stream_reserve(void* stream, uint32_t max_rays) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    rq_status st = ensure_ctx();
    if (st != RQ_OK) return st;
    StreamRes* res = nullptr;
    if ((st = res_for(s, &res)) != RQ_OK) return st;
    // the largest persistent grid: every wave slot of every CU
    if ((st = ensure_spill(*res, (size_t)g_ctx.cus * 2048u)) != RQ_OK) return st;
    uint32_t nblocks = (max_rays + dev::SEC_BLOCK - 1) / dev::SEC_BLOCK;
    return ensure_scratch(*res, (size_t)(dev::MAX_GI + 1) * std::max(nblocks, 1u) * sizeof(uint32_t));
}

extern "C" rq_status rq_get_batch_tuning(rq_layout layout, rq_tuning* out) {
    if (!out || (int)layout < 0 || (int)layout >= LAYOUT_COUNT) {
        set_error("rq_get_batch_tuning: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    init_tuning();
    *out = g_ctx.batch_tuning[layout];
    return RQ_OK;
}

extern "C" rq_status rq_set_batch_tuning(rq_layout layout, const rq_tuning* t) {
    if (!t || (int)layout < 0 || (int)layout >= LAYOUT_COUNT || !find_kernel(t->block, t->stack) || t->refill == 0 ||
        t->refill > 64) {
        set_error("rq_set_batch_tuning: unsupported block/stack/refill combination");
        return RQ_ERR_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    init_tuning();
    g_ctx.batch_tuning[layout] = *t;
    return RQ_OK;
}

extern "C" rq_status rq_trace_count_stats(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count,
                                          uint32_t mode, rq_hit* hits, uint32_t* vis_bits, uint64_t stats[16]) {
    if (!scene || !stats || !layout_ok(scene, layout) || (mode & ~RQ_REORDER) > 2 || (count && !rays) ||
        !scene->dev[layout].ready) {
        set_error("rq_trace_count_bytes: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    unsigned long long* acc = nullptr;
    HIP_TRY(hipMalloc(&acc, 16 * sizeof(unsigned long long)), "hipMalloc count");
    HIP_TRY(hipMemset(acc, 0, 16 * sizeof(unsigned long long)), "hipMemset count");
    rq_status st = launch_trace(scene, layout, rays, count, mode, hits, vis_bits, nullptr, acc);
    unsigned long long host[16] = {};
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(host, acc, sizeof host, hipMemcpyDeviceToHost);
    (void)hipFree(acc);
    if (st != RQ_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "count");
    for (int k = 0; k < 16; ++k) stats[k] = host[k];
    return RQ_OK;
}

extern "C" rq_status rq_trace_count_bytes(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count,
                                          uint32_t mode, rq_hit* hits, uint32_t* vis_bits, uint64_t* bytes) {
    uint64_t stats[16];
    rq_status st = rq_trace_count_stats(scene, layout, rays, count, mode, hits, vis_bits, stats);
    if (st == RQ_OK) *bytes = stats[0];
    return st;
}

extern "C" rq_status rq_verify_fp64(rq_scene* scene, const rq_ray* rays, const rq_hit* hits, uint32_t count,
                                    double out[8], void* stream) {
    if (!scene || !out || (count && (!rays || !hits))) {
        set_error("rq_verify_fp64: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    if (scene->kind != RQ_SCENE_GI || !scene->dev[0].ready) {
        set_error("rq_verify_fp64: needs a GI scene with the REF layout uploaded (the original positions)");
        return RQ_ERR_INVALID_INPUT;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    unsigned long long* acc = nullptr;
    HIP_TRY(hipMalloc(&acc, 8 * sizeof(unsigned long long)), "hipMalloc verify");
    HIP_TRY(hipMemsetAsync(acc, 0, 8 * sizeof(unsigned long long), s), "hipMemset verify");
    dev::VerifyArgs a;
    std::memset(&a, 0, sizeof a);
    a.words = reinterpret_cast<const uint32_t*>(scene->dev[0].base);
    a.rays = rays;
    a.hits = hits;
    a.count = count;
    a.out = acc;
    if (count) hipLaunchKernelGGL(dev::verify_kernel, dim3((count + 255) / 256), dim3(256), 0, s, a);
    unsigned long long host[8];
    hipError_t e = hipMemcpyAsync(host, acc, sizeof host, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(acc);
    if (e != hipSuccess) return hip_fail(e, "verify");
    auto as_double = [](unsigned long long bits) {
        double d;
        std::memcpy(&d, &bits, 8);
        return d;
    };
    out[0] = as_double(host[0]);
    out[1] = as_double(host[1]);
    out[2] = (double)host[2];
    out[3] = as_double(host[3]);
    out[4] = (double)host[4];
    out[5] = as_double(host[5]);
    out[6] = (double)host[6];
    out[7] = 0.0;
    return RQ_OK;
}

extern "C" rq_status rq_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src || bytes % 16) {
        set_error("rq_device_copy: null pointer or size not a multiple of 16");
        return RQ_ERR_INVALID_INPUT;
    }
    uint64_t n = bytes / 16;
    hipLaunchKernelGGL(dev::copy_kernel, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, 8192)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), reinterpret_cast<uint4*>(dst),
                       reinterpret_cast<const uint4*>(src), n);
    HIP_TRY(hipGetLastError(), "copy launch");
    return RQ_OK;
}

namespace rq {
namespace {
rq_status gen_primary(const float inverse[16], uint32_t width, uint32_t height, rq_ray* rays, void* stream,
                      bool tiled) {
    if (!inverse || !rays || width == 0 || height == 0 || (uint64_t)width * height >= (1ull << 32) ||
        (tiled && (width % 8 || height % 8))) {
        set_error(tiled ? "rq_gen_primary_tiled: bad argument (width and height must be multiples of 8)"
                        : "rq_gen_primary: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    dev::Mat16 m;
    std::memcpy(m.m, inverse, sizeof m.m);
    uint64_t n = (uint64_t)width * height;
    if (tiled)
        hipLaunchKernelGGL(dev::primary_kernel<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           reinterpret_cast<hipStream_t>(stream), m, width, height, rays);
    else
        hipLaunchKernelGGL(dev::primary_kernel<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           reinterpret_cast<hipStream_t>(stream), m, width, height, rays);
    HIP_TRY(hipGetLastError(), "primary launch");
    return RQ_OK;
}
}  // namespace
}  // namespace rq

extern "C" rq_status rq_gen_primary(const float inverse[16], uint32_t width, uint32_t height, rq_ray* rays,
                                    void* stream) {
    return rq::gen_primary(inverse, width, height, rays, stream, false);
}

extern "C" rq_status rq_gen_primary_tiled(const float inverse[16], uint32_t width, uint32_t height, rq_ray* rays,
                                          void* stream) {
    return rq::gen_primary(inverse, width, height, rays, stream, true);
}

extern "C" rq_status rq_gen_secondary(const rq_ray* primary, const rq_hit* hits, uint32_t count, const float sun[3],
                                      uint32_t seed, uint32_t n_gi, rq_ray* shadow, rq_ray* gi, uint32_t* counts,
                                      void* stream) {
    if (!primary || !hits || !sun || !counts || n_gi > dev::MAX_GI || !shadow || (n_gi && !gi)) {
        set_error("rq_gen_secondary: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    if (count == 0) return RQ_OK;
    std::lock_guard<std::mutex> lock(g_ctx.mu);
    rq_status st = ensure_ctx();
    if (st != RQ_OK) return st;
    uint32_t nblocks = (count + dev::SEC_BLOCK - 1) / dev::SEC_BLOCK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    StreamRes* res = nullptr;
    if ((st = res_for(s, &res)) != RQ_OK) return st;
    st = ensure_scratch(*res, (size_t)(n_gi + 1) * nblocks * sizeof(uint32_t));
    if (st != RQ_OK) return st;
    dev::SecArgs a;
    a.primary = primary;
    a.hits = hits;
    a.count = count;
    a.sx = sun[0];
    a.sy = sun[1];
    a.sz = sun[2];
    a.seed = seed;
    a.n_gi = n_gi;
    a.block_counts = reinterpret_cast<uint32_t*>(res->scratch);
    a.counts = counts;
    a.shadow = shadow;
    a.gi = gi;
    a.nblocks = nblocks;
    hipLaunchKernelGGL(dev::secondary_count, dim3(nblocks), dim3(dev::SEC_BLOCK), 0, s, a);
    hipLaunchKernelGGL(dev::secondary_scan, dim3(n_gi + 1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(dev::secondary_write, dim3(nblocks), dim3(dev::SEC_BLOCK), 0, s, a);
    HIP_TRY(hipGetLastError(), "secondary launch");
    return RQ_OK;
}

// ---- device helpers
extern "C" rq_status rq_malloc(size_t bytes, void** ptr) {
    if (!ptr) return RQ_ERR_INVALID_INPUT;
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
    return RQ_OK;
}
extern "C" rq_status rq_free(void* ptr) {
    HIP_TRY(hipFree(ptr), "hipFree");
    return RQ_OK;
}
extern "C" rq_status rq_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy h2d");
    return RQ_OK;
}
extern "C" rq_status rq_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy d2h");
    return RQ_OK;
}
extern "C" rq_status rq_memset(void* dst, int value, size_t bytes) {
    HIP_TRY(hipMemset(dst, value, bytes), "hipMemset");
    return RQ_OK;
}
extern "C" rq_status rq_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return RQ_OK;
}
extern "C" rq_status rq_stream_create(void** stream) {
    if (!stream) return RQ_ERR_INVALID_INPUT;
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = s;
    return RQ_OK;
}
extern "C" rq_status rq_stream_destroy(void* stream) {
    release_stream(reinterpret_cast<hipStream_t>(stream));
    HIP_TRY(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)), "hipStreamDestroy");
    return RQ_OK;
}
extern "C" rq_status rq_stream_synchronize(void* stream) {
    HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return RQ_OK;
}
extern "C" rq_status rq_event_create(void** event) {
    if (!event) return RQ_ERR_INVALID_INPUT;
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return RQ_OK;
}
extern "C" rq_status rq_event_destroy(void* event) {
    HIP_TRY(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)), "hipEventDestroy");
    return RQ_OK;
}
extern "C" rq_status rq_event_record(void* event, void* stream) {
    HIP_TRY(hipEventRecord(reinterpret_cast<hipEvent_t>(event), reinterpret_cast<hipStream_t>(stream)),
            "hipEventRecord");
    return RQ_OK;
}
extern "C" rq_status rq_stream_wait_event(void* stream, void* event) {
    if (!event) return RQ_ERR_INVALID_INPUT;
    HIP_TRY(hipStreamWaitEvent(reinterpret_cast<hipStream_t>(stream), reinterpret_cast<hipEvent_t>(event), 0),
            "hipStreamWaitEvent");
    return RQ_OK;
}
extern "C" rq_status rq_event_synchronize(void* event) {
    HIP_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)), "hipEventSynchronize");
    return RQ_OK;
}
extern "C" rq_status rq_event_elapsed_ms(void* start, void* stop, float* ms) {
    HIP_TRY(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)),
            "hipEventElapsedTime");
    return RQ_OK;
}
extern "C" rq_status rq_event_query(void* event) {
    hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(event));
    if (e == hipSuccess) return RQ_OK;
    if (e == hipErrorNotReady) return RQ_ERR_NOT_READY;
    return hip_fail(e, "hipEventQuery");
}
