// GPU build of every mesh's BVH4 for a whole scene, word for word the tree the
// host builder makes (near_field::build_bvh_sah, crates/asset/src/
// near_field.rs:172-430), packed into the gi_geometry words as
// rq_scene_build packs it (src/gi_bvh.rs:33-274, src/gi_cache.rs:4388-4479).
//
// The host builder recurses: a group of more than LEAF_TRIANGLES triangles is
// split twice (binned SAH, 16 bins, or the median when SAH finds no plane) and
// each of the 2-4 groups recurses; nodes are numbered in post-order, leaves in
// creation order. Here every mesh is built at once, one BVH4 level at a time:
// each level's groups are split in two passes over all groups together (a
// workgroup per group up to 2048 triangles; chunked with global bins above;
// median splits through a segmented radix sort on (total-order key, triangle)),
// partitions stable, so each group's triangle order is the host's. After the
// last level: subtree sizes bottom-up give the post-order ids, a scan over
// leaf starts gives leaf ids (leaves are contiguous in the final permutation,
// left to right), and the node boxes come from the refit kernel applied level
// by level from the deepest -- the host builder's child_bounds and
// quantized_node on the same children. A mesh whose tree is deeper than
// MAX_SAH_DEPTH is built again with median splits only (near_field.rs:184-190).
//
// Exactness notes: centroids, bins, areas and costs use the host's operations
// in the host's order (no contraction; correctly rounded division); min and
// max of centroids and boxes are taken with atomics on total-order keys,
// which differ from the host's sequential min/max only in the sign of a zero,
// and no decision or stored word depends on that sign (binning subtracts the
// bound, the extent of all-zero coordinates is +0 either way, costs compare
// equal). Duplicate index triplets stay in ascending source order through
// every stable partition and (key, triangle) sort, so the primitive table is
// the permutation itself (what map_primitives, src/gi_bvh.rs:108-136, finds).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>

#include "bvh_device.h"
#include "internal.h"

namespace rq {
namespace gbuild {

using bvhdev::bf;
using bvhdev::fb;

constexpr uint32_t BINS = 16;
constexpr uint32_t SMALL_MAX = 2048;  // triangles a workgroup splits on its own
constexpr uint32_t CHUNK = 4096;      // triangles per workgroup in the chunked path
constexpr uint32_t MODE_SAH = 0, MODE_MEDIAN = 1;

struct MeshDev {
    uint32_t tbase, tcount, vbase, vcount;
    uint32_t mode, keep, node_count, leaf_count;  // counts: set once the layout is known
    uint32_t nodes, leaves, positions, indices, prims, lbase, pad3, pad4;  // lbase: first leaf box
};

struct Seg {
    uint32_t start, count, mesh, out;
};

struct MedianReq {
    uint32_t start, count, mesh, out, axis, pad0, pad1, pad2;
};

struct Chunk {
    uint32_t seg, start, count, pad;
};

using bvhdev::okey;
using bvhdev::okey_inv;
__device__ __forceinline__ uint32_t sat_index(float f, uint32_t cap) {
    if (!(f > 0.0f)) return 0;
    if (f >= (float)cap) return cap;
    return (uint32_t)f;
}
// near_field.rs:329-331: the bin of a centroid coordinate.
__device__ __forceinline__ uint32_t bin_of(float c, float lo, float extent) {
    const float f = (c - lo) / extent * (float)BINS;
    return min(sat_index(f, 0xffffffffu), BINS - 1);
}
__device__ __forceinline__ float half_area(const float lo[3], const float hi[3]) {
    float e[3];
    for (int i = 0; i < 3; ++i) e[i] = bvhdev::mx(hi[i] - lo[i], 0.0f);
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

// near_field.rs:337-360: the sweep over one group's bins (counts and box keys
// laid out [axis][bin] and [axis][bin][lo xyz, hi xyz]).
__device__ inline bool sah_sweep(const uint32_t* cnt, const int32_t* box, const bool valid[3], uint32_t size,
                                 int& best_axis, uint32_t& best_plane) {
    bool found = false;
    float best_cost = 0.0f;
    for (int axis = 0; axis < 3; ++axis) {
        if (!valid[axis]) continue;
        float right_cost[BINS];
        float lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
            lo[i] = __builtin_inff();
            hi[i] = -__builtin_inff();
        }
        uint32_t n = 0;
        for (uint32_t b = BINS - 1; b >= 1; --b) {
            const int32_t* bb = box + (axis * BINS + b) * 6;
            for (int i = 0; i < 3; ++i) {
                lo[i] = bvhdev::mn(lo[i], okey_inv(bb[i]));
                hi[i] = bvhdev::mx(hi[i], okey_inv(bb[3 + i]));
            }
            n += cnt[axis * BINS + b];
            right_cost[b] = half_area(lo, hi) * (float)n;
        }
        for (int i = 0; i < 3; ++i) {
            lo[i] = __builtin_inff();
            hi[i] = -__builtin_inff();
        }
        n = 0;
        for (uint32_t b = 0; b < BINS - 1; ++b) {
            const int32_t* bb = box + (axis * BINS + b) * 6;
            for (int i = 0; i < 3; ++i) {
                lo[i] = bvhdev::mn(lo[i], okey_inv(bb[i]));
                hi[i] = bvhdev::mx(hi[i], okey_inv(bb[3 + i]));
            }
            n += cnt[axis * BINS + b];
            if (n == 0 || n == size) continue;
            const float cost = half_area(lo, hi) * (float)n + right_cost[b + 1];
            if (!found || cost < best_cost) {
                found = true;
                best_cost = cost;
                best_axis = axis;
                best_plane = b;
            }
        }
    }
    return found;
}

// near_field.rs:367-372: the longest centroid extent, ties to the later axis.
__device__ __forceinline__ uint32_t median_axis(const float extent[3]) {
    uint32_t axis = 0;
    for (uint32_t a = 1; a < 3; ++a)
        if (okey(extent[a]) >= okey(extent[axis])) axis = a;
    return axis;
}

// One candidate plane of the sweep above, evaluated on its own: c = axis * 15 +
// plane, the host's visiting order. The left and right boxes are the same
// min/max folds the host's two sweeps make, so the cost is the host's.
__device__ inline bool plane_cost(const uint32_t* cnt, const int32_t* box, const bool valid[3], uint32_t size,
                                  uint32_t c, float& cost) {
    const uint32_t axis = c / (BINS - 1), plane = c % (BINS - 1);
    if (c >= 3 * (BINS - 1) || !valid[axis]) return false;
    float llo[3], lhi[3], rlo[3], rhi[3];
    for (int i = 0; i < 3; ++i) {
        llo[i] = rlo[i] = __builtin_inff();
        lhi[i] = rhi[i] = -__builtin_inff();
    }
    uint32_t nl = 0, nr = 0;
    for (uint32_t b = 0; b < BINS; ++b) {
        const int32_t* bb = box + (axis * BINS + b) * 6;
        float* lo = b <= plane ? llo : rlo;
        float* hi = b <= plane ? lhi : rhi;
        for (int i = 0; i < 3; ++i) {
            lo[i] = bvhdev::mn(lo[i], okey_inv(bb[i]));
            hi[i] = bvhdev::mx(hi[i], okey_inv(bb[3 + i]));
        }
        (b <= plane ? nl : nr) += cnt[axis * BINS + b];
    }
    if (nl == 0 || nl == size) return false;
    cost = half_area(llo, lhi) * (float)nl + half_area(rlo, rhi) * (float)nr;
    return true;
}

// The sequential choice over the 45 candidates (first one taken, then only a
// strictly smaller cost replaces it) across one wave: the first candidate if
// its cost is NaN, else the lowest-indexed candidate of least non-NaN cost.
// Returns the winning lane, or -1 if no lane holds a candidate.
__device__ inline int pick_plane(bool cand, float cost) {
    const uint64_t m = __ballot(cand);
    if (m == 0) return -1;
    const int first = __builtin_ctzll(m);
    const float fc = __shfl(cost, first, 64);
    if (fc != fc) return first;
    bool ok = cand && cost == cost;
    float bc = cost;
    int bl = (int)(threadIdx.x & 63u);
    for (int o = 32; o > 0; o >>= 1) {
        const bool ook = __shfl_xor((int)ok, o, 64) != 0;
        const float oc = __shfl_xor(bc, o, 64);
        const int ol = __shfl_xor(bl, o, 64);
        if (ook && (!ok || oc < bc || (oc == bc && ol < bl))) {
            ok = true;
            bc = oc;
            bl = ol;
        }
    }
    return bl;
}

// Exclusive prefix sum of v over a 256-thread workgroup; *total gets the sum.
__device__ inline uint32_t block_exclusive(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < 4; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    *total = all;
    return before + x - v;
}

__device__ __forceinline__ void ask_median(MedianReq* med, uint32_t* med_count, const Seg& s, uint32_t axis) {
    const uint32_t slot = atomicAdd(med_count, 1u);
    MedianReq r;
    r.start = s.start;
    r.count = s.count;
    r.mesh = s.mesh;
    r.out = s.out;
    r.axis = axis;
    r.pad0 = r.pad1 = r.pad2 = 0;
    med[slot] = r;
}

// ---------------------------------------------------------------- set-up
__global__ __launch_bounds__(256) void prep_tris(const MeshDev* meshes, uint32_t mesh_count, const float* pos,
                                                 const uint32_t* idx, uint32_t tris, float* cent, float* box,
                                                 uint32_t* perm) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= tris) return;
    uint32_t lo = 0, hi = mesh_count;  // last mesh with tbase <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (meshes[mid].tbase <= t) lo = mid;
        else hi = mid;
    }
    const float* p = pos + 3ull * meshes[lo].vbase;
    const float* a = p + 3ull * idx[3ull * t];
    const float* b = p + 3ull * idx[3ull * t + 1];
    const float* c = p + 3ull * idx[3ull * t + 2];
    for (int x = 0; x < 3; ++x) {  // near_field.rs:213-219
        float s = a[x] + b[x];
        s = s + c[x];
        cent[3ull * t + x] = s / 3.0f;
        box[6ull * t + x] = bvhdev::mn(bvhdev::mn(a[x], b[x]), c[x]);
        box[6ull * t + 3 + x] = bvhdev::mx(bvhdev::mx(a[x], b[x]), c[x]);
    }
    perm[t] = t;
}

__global__ __launch_bounds__(256) void fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull) p[i] = v;
}

// Level-0 groups: one per mesh, all its triangles in source order.
__global__ __launch_bounds__(256) void root_groups(const MeshDev* meshes, uint32_t mesh_count, uint32_t* gstart,
                                                   uint32_t* gcount, uint32_t* gmesh) {
    const uint32_t m = blockIdx.x * 256u + threadIdx.x;
    if (m >= mesh_count) return;
    gstart[m] = meshes[m].tbase;
    gcount[m] = meshes[m].tcount;
    gmesh[m] = m;
}

// Appends `s` where `want` holds, one atomic per wave (list order is free:
// every entry is processed on its own). All lanes of the wave call this.
__device__ inline void append(bool want, const Seg& s, Seg* list, uint32_t* count) {
    const uint64_t m = __ballot(want);
    if (m == 0) return;
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(count, (uint32_t)__builtin_popcountll(m));
    base = __shfl(base, leader, 64);
    if (want) list[base + __builtin_popcountll(m & ((1ull << lane) - 1))] = s;
}

// ---------------------------------------------------------------- classify
// The total of an exclusive scan (last prefix + last flag) in one word, so a
// level reads it back with one copy.
__global__ void scan_total(const uint32_t* scan, const uint32_t* flag, uint32_t n, uint32_t* out) {
    if (threadIdx.x == 0) out[0] = scan[n - 1] + flag[n - 1];
}

__global__ __launch_bounds__(256) void node_flags(const uint32_t* gcount, uint32_t n, uint32_t* flag) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g < n) flag[g] = gcount[g] > LEAF_TRIANGLES ? 1u : 0u;
}

__global__ __launch_bounds__(256) void make_nodes(const uint32_t* gstart, const uint32_t* gcount, const uint32_t* gmesh,
                                                  uint32_t* gnode, uint32_t gbase, uint32_t ng, const uint32_t* noff,
                                                  uint32_t nbase, uint32_t level, uint32_t* node_group,
                                                  uint32_t* node_mesh, uint32_t* node_level, Seg* segs) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= ng) return;
    const uint32_t gi = gbase + g;
    if (gcount[gi] <= LEAF_TRIANGLES) {
        gnode[gi] = EMPTY;
        return;
    }
    const uint32_t r = noff[g], k = nbase + r;
    gnode[gi] = k;
    node_group[k] = gi;
    node_mesh[k] = gmesh[gi];
    node_level[k] = level;
    Seg s;
    s.start = gstart[gi];
    s.count = gcount[gi];
    s.mesh = gmesh[gi];
    s.out = 3 * r;
    segs[r] = s;
}

// Round 1: the halves of each node's group that still hold too many triangles.
__global__ __launch_bounds__(256) void round1_segs(const uint32_t* gstart, const uint32_t* gcount, const uint32_t* gmesh,
                                                   const uint32_t* node_group, uint32_t nbase, uint32_t nn,
                                                   const uint32_t* split_left, Seg* segs, uint32_t* seg_count) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    const bool live = r < nn;
    uint32_t start = 0, count = 0, l0 = 0, mesh = 0;
    if (live) {
        const uint32_t gi = node_group[nbase + r];
        start = gstart[gi];
        count = gcount[gi];
        l0 = split_left[3 * r];
        mesh = gmesh[gi];
    }
    const uint32_t hs[2] = {start, start + l0}, hc[2] = {l0, count - l0};
    for (int h = 0; h < 2; ++h) {
        Seg s;
        s.start = hs[h];
        s.count = hc[h];
        s.mesh = mesh;
        s.out = 3 * r + 1 + h;
        append(live && hc[h] > LEAF_TRIANGLES, s, segs, seg_count);
    }
}

__global__ __launch_bounds__(256) void child_counts(const uint32_t* gcount, const uint32_t* node_group, uint32_t nbase,
                                                    uint32_t nn, const uint32_t* split_left, uint32_t* cc) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= nn) return;
    const uint32_t count = gcount[node_group[nbase + r]], l0 = split_left[3 * r];
    cc[r] = (l0 > LEAF_TRIANGLES ? 2u : 1u) + (count - l0 > LEAF_TRIANGLES ? 2u : 1u);
}

// near_field.rs:289-305: the node's children, halves split once more in order.
__global__ __launch_bounds__(256) void emit_children(uint32_t* gstart, uint32_t* gcount, uint32_t* gmesh,
                                                     const uint32_t* node_group, uint32_t nbase, uint32_t nn,
                                                     const uint32_t* split_left, const uint32_t* coff, uint32_t gnext,
                                                     uint32_t* child_first, uint32_t* child_count) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= nn) return;
    const uint32_t k = nbase + r, gi = node_group[k];
    const uint32_t start = gstart[gi], count = gcount[gi], mesh = gmesh[gi], l0 = split_left[3 * r];
    const uint32_t hs[2] = {start, start + l0}, hc[2] = {l0, count - l0};
    uint32_t at = gnext + coff[r], n = 0;
    for (int h = 0; h < 2; ++h) {
        if (hc[h] > LEAF_TRIANGLES) {
            const uint32_t l1 = split_left[3 * r + 1 + h];
            gstart[at + n] = hs[h];
            gcount[at + n] = l1;
            gmesh[at + n] = mesh;
            ++n;
            gstart[at + n] = hs[h] + l1;
            gcount[at + n] = hc[h] - l1;
            gmesh[at + n] = mesh;
            ++n;
        } else {
            gstart[at + n] = hs[h];
            gcount[at + n] = hc[h];
            gmesh[at + n] = mesh;
            ++n;
        }
    }
    child_first[k] = at;
    child_count[k] = n;
}

// ---------------------------------------------------------------- splits
__global__ __launch_bounds__(256) void route(const Seg* segs, uint32_t n, Seg* small, Seg* large, uint32_t* counts) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool live = i < n;
    Seg s;
    if (live) s = segs[i];
    const bool big = live && s.count > SMALL_MAX;
    append(live && !big, s, small, &counts[0]);
    append(big, s, large, &counts[1]);
}

// One workgroup splits one group of at most SMALL_MAX triangles in place.
__global__ __launch_bounds__(256) void split_small(const Seg* segs, const MeshDev* meshes, uint32_t* perm,
                                                   const float* cent, const float* box, uint32_t* split_left,
                                                   MedianReq* med, uint32_t* med_count) {
    __shared__ uint32_t ids[SMALL_MAX];
    __shared__ int32_t bounds[6];
    __shared__ uint32_t cnt[3 * BINS];
    __shared__ int32_t bins[3 * BINS * 6];
    __shared__ uint32_t runs[4];
    __shared__ int s_found, s_axis;
    __shared__ uint32_t s_plane;
    const Seg sg = segs[blockIdx.x];
    const uint32_t n = sg.count, t = threadIdx.x;
    if (t < 3) {
        bounds[t] = okey(__builtin_inff());
        bounds[3 + t] = okey(-__builtin_inff());
    }
    for (uint32_t i = t; i < 3 * BINS; i += 256) {
        cnt[i] = 0;
        for (int k = 0; k < 3; ++k) {
            bins[i * 6 + k] = okey(__builtin_inff());
            bins[i * 6 + 3 + k] = okey(-__builtin_inff());
        }
    }
    __syncthreads();
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint32_t i = t; i < n; i += 256) {
        const uint32_t id = perm[sg.start + i];
        ids[i] = id;
        for (int x = 0; x < 3; ++x) {
            const float c = cent[3ull * id + x];
            if (c == c) {
                lo[x] = fminf(lo[x], c);
                hi[x] = fmaxf(hi[x], c);
            }
        }
    }
    for (int x = 0; x < 3; ++x) {
        atomicMin(&bounds[x], okey(lo[x]));
        atomicMax(&bounds[3 + x], okey(hi[x]));
    }
    __syncthreads();
    float clo[3], extent[3];
    bool valid[3];
    for (int x = 0; x < 3; ++x) {
        clo[x] = okey_inv(bounds[x]);
        extent[x] = okey_inv(bounds[3 + x]) - clo[x];
        valid[x] = extent[x] > 0.0f;  // near_field.rs:325: NaN or <= 0 skips the axis
    }
    const bool sah = meshes[sg.mesh].mode == MODE_SAH;
    if (sah) {
        for (uint32_t i = t; i < n; i += 256) {
            const uint32_t id = ids[i];
            for (int x = 0; x < 3; ++x) {
                if (!valid[x]) continue;
                const uint32_t b = x * BINS + bin_of(cent[3ull * id + x], clo[x], extent[x]);
                atomicAdd(&cnt[b], 1u);
                for (int k = 0; k < 3; ++k) {
                    atomicMin(&bins[b * 6 + k], okey(box[6ull * id + k]));
                    atomicMax(&bins[b * 6 + 3 + k], okey(box[6ull * id + 3 + k]));
                }
            }
        }
        __syncthreads();
        if (t < 64) {
            float cost = 0.0f;
            const bool cand = plane_cost(cnt, bins, valid, n, t, cost);
            const int win = pick_plane(cand, cost);
            if (t == 0) {
                s_found = win >= 0;
                s_axis = win >= 0 ? win / (int)(BINS - 1) : 0;
                s_plane = win >= 0 ? (uint32_t)win % (BINS - 1) : 0;
            }
        }
        __syncthreads();
    }
    if (!sah || !s_found) {
        if (t == 0) ask_median(med, med_count, sg, median_axis(extent));
        return;
    }
    const int ax = s_axis;
    const uint32_t plane = s_plane;
    const uint32_t per = (n + 255) / 256, b0 = min(t * per, n), b1 = min(b0 + per, n);
    uint32_t nl = 0;
    for (uint32_t i = b0; i < b1; ++i) nl += bin_of(cent[3ull * ids[i] + ax], clo[ax], extent[ax]) <= plane;
    uint32_t nleft;
    const uint32_t before = block_exclusive(nl, runs, &nleft);
    uint32_t l = before, r = b0 - before;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t id = ids[i];
        if (bin_of(cent[3ull * id + ax], clo[ax], extent[ax]) <= plane) perm[sg.start + l++] = id;
        else perm[sg.start + nleft + r++] = id;
    }
    if (t == 0) split_left[sg.out] = nleft;
}

// Chunked path for groups above SMALL_MAX triangles.
__global__ __launch_bounds__(256) void large_bounds(const Chunk* chunks, const uint32_t* perm, const float* cent,
                                                    int32_t* lb) {
    __shared__ int32_t b[6];
    const Chunk c = chunks[blockIdx.x];
    const uint32_t t = threadIdx.x;
    if (t < 3) {
        b[t] = okey(__builtin_inff());
        b[3 + t] = okey(-__builtin_inff());
    }
    __syncthreads();
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint32_t i = t; i < c.count; i += 256) {
        const uint32_t id = perm[c.start + i];
        for (int x = 0; x < 3; ++x) {
            const float v = cent[3ull * id + x];
            if (v == v) {
                lo[x] = fminf(lo[x], v);
                hi[x] = fmaxf(hi[x], v);
            }
        }
    }
    for (int x = 0; x < 3; ++x) {
        atomicMin(&b[x], okey(lo[x]));
        atomicMax(&b[3 + x], okey(hi[x]));
    }
    __syncthreads();
    if (t < 3) {
        atomicMin(&lb[6 * c.seg + t], b[t]);
        atomicMax(&lb[6 * c.seg + 3 + t], b[3 + t]);
    }
}

__device__ inline void large_axes(const int32_t* lb, float clo[3], float extent[3], bool valid[3]) {
    for (int x = 0; x < 3; ++x) {
        clo[x] = okey_inv(lb[x]);
        extent[x] = okey_inv(lb[3 + x]) - clo[x];
        valid[x] = extent[x] > 0.0f;
    }
}

__global__ __launch_bounds__(256) void large_bins(const Chunk* chunks, const Seg* large, const MeshDev* meshes,
                                                  const uint32_t* perm, const float* cent, const float* box,
                                                  const int32_t* lb, uint32_t* gcnt, int32_t* gbins) {
    __shared__ uint32_t cnt[3 * BINS];
    __shared__ int32_t bins[3 * BINS * 6];
    const Chunk c = chunks[blockIdx.x];
    if (meshes[large[c.seg].mesh].mode != MODE_SAH) return;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < 3 * BINS; i += 256) {
        cnt[i] = 0;
        for (int k = 0; k < 3; ++k) {
            bins[i * 6 + k] = okey(__builtin_inff());
            bins[i * 6 + 3 + k] = okey(-__builtin_inff());
        }
    }
    __syncthreads();
    float clo[3], extent[3];
    bool valid[3];
    large_axes(lb + 6 * c.seg, clo, extent, valid);
    for (uint32_t i = t; i < c.count; i += 256) {
        const uint32_t id = perm[c.start + i];
        for (int x = 0; x < 3; ++x) {
            if (!valid[x]) continue;
            const uint32_t b = x * BINS + bin_of(cent[3ull * id + x], clo[x], extent[x]);
            atomicAdd(&cnt[b], 1u);
            for (int k = 0; k < 3; ++k) {
                atomicMin(&bins[b * 6 + k], okey(box[6ull * id + k]));
                atomicMax(&bins[b * 6 + 3 + k], okey(box[6ull * id + 3 + k]));
            }
        }
    }
    __syncthreads();
    if (t < 3 * BINS && cnt[t]) {
        uint32_t* gc = gcnt + (size_t)c.seg * 3 * BINS;
        int32_t* gb = gbins + (size_t)c.seg * 3 * BINS * 6;
        atomicAdd(&gc[t], cnt[t]);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&gb[t * 6 + k], bins[t * 6 + k]);
            atomicMax(&gb[t * 6 + 3 + k], bins[t * 6 + 3 + k]);
        }
    }
}

__global__ __launch_bounds__(64) void large_eval(const Seg* large, uint32_t n, const MeshDev* meshes, const int32_t* lb,
                                                 const uint32_t* gcnt, const int32_t* gbins, uint32_t* best,
                                                 MedianReq* med, uint32_t* med_count) {
    const uint32_t j = blockIdx.x * 64u + threadIdx.x;
    if (j >= n) return;
    const Seg s = large[j];
    float clo[3], extent[3];
    bool valid[3];
    large_axes(lb + 6 * j, clo, extent, valid);
    int axis = 0;
    uint32_t plane = 0;
    if (meshes[s.mesh].mode == MODE_SAH &&
        sah_sweep(gcnt + (size_t)j * 3 * BINS, gbins + (size_t)j * 3 * BINS * 6, valid, s.count, axis, plane)) {
        best[j] = ((uint32_t)axis << 8) | plane;
    } else {
        best[j] = EMPTY;
        ask_median(med, med_count, s, median_axis(extent));
    }
}

__device__ inline bool large_left(const int32_t* lb, uint32_t best, const float* cent, uint32_t id) {
    float clo[3], extent[3];
    bool valid[3];
    large_axes(lb, clo, extent, valid);
    const uint32_t ax = best >> 8, plane = best & 0xffu;
    return bin_of(cent[3ull * id + ax], clo[ax], extent[ax]) <= plane;
}

__global__ __launch_bounds__(256) void large_count(const Chunk* chunks, const uint32_t* best, const uint32_t* perm,
                                                   const float* cent, const int32_t* lb, uint32_t* chunk_left) {
    __shared__ uint32_t sum;
    const Chunk c = chunks[blockIdx.x];
    const uint32_t bst = best[c.seg];
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    if (bst == EMPTY) {
        if (threadIdx.x == 0) chunk_left[blockIdx.x] = 0;
        return;
    }
    uint32_t nl = 0;
    for (uint32_t i = threadIdx.x; i < c.count; i += 256) nl += large_left(lb + 6 * c.seg, bst, cent, perm[c.start + i]);
    atomicAdd(&sum, nl);
    __syncthreads();
    if (threadIdx.x == 0) chunk_left[blockIdx.x] = sum;
}

__global__ __launch_bounds__(64) void large_total(const Seg* large, uint32_t n, const uint32_t* best,
                                                  const uint32_t* first_chunk, const uint32_t* chunk_n,
                                                  const uint32_t* chunk_left, const uint32_t* chunk_off,
                                                  uint32_t* seg_left, uint32_t* split_left) {
    const uint32_t j = blockIdx.x * 64u + threadIdx.x;
    if (j >= n || best[j] == EMPTY) return;
    const uint32_t f = first_chunk[j], l = f + chunk_n[j] - 1;
    const uint32_t nl = chunk_off[l] + chunk_left[l] - chunk_off[f];
    seg_left[j] = nl;
    split_left[large[j].out] = nl;
}

__global__ __launch_bounds__(256) void large_scatter(const Chunk* chunks, const Seg* large, const uint32_t* best,
                                                     const uint32_t* first_chunk, const uint32_t* chunk_off,
                                                     const uint32_t* seg_left, const uint32_t* perm, const float* cent,
                                                     const int32_t* lb, uint32_t* tmp) {
    __shared__ uint32_t runs[4];
    const Chunk c = chunks[blockIdx.x];
    const uint32_t bst = best[c.seg];
    if (bst == EMPTY) return;
    const Seg s = large[c.seg];
    const uint32_t t = threadIdx.x, per = (c.count + 255) / 256, b0 = min(t * per, c.count),
                   b1 = min(b0 + per, c.count);
    uint32_t nl = 0;
    for (uint32_t i = b0; i < b1; ++i) nl += large_left(lb + 6 * c.seg, bst, cent, perm[c.start + i]);
    uint32_t all;
    const uint32_t before = block_exclusive(nl, runs, &all);
    const uint32_t lefts_before = chunk_off[blockIdx.x] - chunk_off[first_chunk[c.seg]];
    const uint32_t rights_before = (c.start - s.start) - lefts_before;
    uint32_t l = lefts_before + before, r = rights_before + (b0 - before);
    const uint32_t nleft = seg_left[c.seg];
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t id = perm[c.start + i];
        if (large_left(lb + 6 * c.seg, bst, cent, id)) tmp[s.start + l++] = id;
        else tmp[s.start + nleft + r++] = id;
    }
}

__global__ __launch_bounds__(256) void large_copy(const Chunk* chunks, const uint32_t* best, const uint32_t* tmp,
                                                  uint32_t* perm) {
    const Chunk c = chunks[blockIdx.x];
    if (best[c.seg] == EMPTY) return;
    for (uint32_t i = threadIdx.x; i < c.count; i += 256) perm[c.start + i] = tmp[c.start + i];
}

// Median splits (near_field.rs:367-379): sort by (total-order key of the
// centroid on the axis, triangle), lower half left.
__global__ __launch_bounds__(256) void median_keys(const MedianReq* med, const uint32_t* off, const MeshDev* meshes,
                                                   const uint32_t* perm, const float* cent, uint64_t* keys) {
    const MedianReq q = med[blockIdx.x];
    const uint32_t tb = meshes[q.mesh].tbase;
    for (uint32_t i = threadIdx.x; i < q.count; i += 256) {
        const uint32_t id = perm[q.start + i];
        const uint32_t k = (uint32_t)okey(cent[3ull * id + q.axis]) ^ 0x80000000u;
        keys[off[blockIdx.x] + i] = ((uint64_t)k << 32) | (id - tb);
    }
}

__global__ __launch_bounds__(256) void median_put(const MedianReq* med, const uint32_t* off, const MeshDev* meshes,
                                                  const uint64_t* keys, uint32_t* perm, uint32_t* split_left) {
    const MedianReq q = med[blockIdx.x];
    const uint32_t tb = meshes[q.mesh].tbase;
    for (uint32_t i = threadIdx.x; i < q.count; i += 256)
        perm[q.start + i] = tb + (uint32_t)(keys[off[blockIdx.x] + i] & 0xffffffffu);
    if (threadIdx.x == 0) split_left[q.out] = q.count / 2;
}

// ---------------------------------------------------------------- numbering
__global__ __launch_bounds__(256) void node_sizes(uint32_t nbase, uint32_t nn, const uint32_t* child_first,
                                                  const uint32_t* child_count, const uint32_t* gnode, uint32_t* size) {
    const uint32_t k = nbase + blockIdx.x * 256u + threadIdx.x;
    if (k >= nbase + nn) return;
    uint32_t s = 1;
    for (uint32_t c = 0; c < child_count[k]; ++c) {
        const uint32_t g = gnode[child_first[k] + c];
        if (g != EMPTY) s += size[g];
    }
    size[k] = s;
}

// Post-order: each child's subtree in turn, then the node (near_field.rs:305).
__global__ __launch_bounds__(256) void node_ids(uint32_t nbase, uint32_t nn, bool roots, const uint32_t* child_first,
                                                const uint32_t* child_count, const uint32_t* gnode,
                                                const uint32_t* size, uint32_t* start, uint32_t* id) {
    const uint32_t k = nbase + blockIdx.x * 256u + threadIdx.x;
    if (k >= nbase + nn) return;
    if (roots) start[k] = 0;
    uint32_t at = start[k];
    id[k] = at + size[k] - 1;
    for (uint32_t c = 0; c < child_count[k]; ++c) {
        const uint32_t g = gnode[child_first[k] + c];
        if (g == EMPTY) continue;
        start[g] = at;
        at += size[g];
    }
}

__global__ __launch_bounds__(256) void leaf_marks(const uint32_t* gstart, const uint32_t* gcount, uint32_t n,
                                                  uint32_t* mark) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g < n && gcount[g] <= LEAF_TRIANGLES) mark[gstart[g]] = 1;
}

__global__ __launch_bounds__(256) void mesh_summary(const MeshDev* meshes, uint32_t mesh_count, const uint32_t* gnode,
                                                    const uint32_t* size, const uint32_t* id, const uint32_t* mark,
                                                    const uint32_t* lrank, uint32_t* out /* 4 per mesh */) {
    const uint32_t m = blockIdx.x * 256u + threadIdx.x;
    if (m >= mesh_count) return;
    const uint32_t root = gnode[m];  // level-0 group m is mesh m
    const uint32_t tb = meshes[m].tbase, te = tb + meshes[m].tcount - 1;
    out[4 * m + 0] = root == EMPTY ? 0u : size[root];
    out[4 * m + 1] = lrank[te] + mark[te] - lrank[tb];
    out[4 * m + 2] = root == EMPTY ? LEAF_BIT : id[root];
}

__global__ __launch_bounds__(256) void mesh_depth(const uint32_t* node_mesh, const uint32_t* node_level, uint32_t n,
                                                  uint32_t* out) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) atomicMax(&out[4 * node_mesh[k] + 3], node_level[k] + 1);
}

// ---------------------------------------------------------------- words
__device__ __forceinline__ uint32_t leaf_id(const MeshDev& m, const uint32_t* lrank, uint32_t start) {
    return lrank[start] - lrank[m.tbase];
}

__global__ __launch_bounds__(256) void write_nodes(uint32_t n, const uint32_t* node_mesh, const MeshDev* meshes,
                                                   const uint32_t* id, const uint32_t* child_first,
                                                   const uint32_t* child_count, const uint32_t* gnode,
                                                   const uint32_t* gstart, const uint32_t* lrank, uint32_t* words) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const MeshDev m = meshes[node_mesh[k]];
    if (id[k] >= m.node_count) return;  // numbering out of range: left zero, the words will not match
    uint32_t* w = words + m.nodes + 16ull * id[k];
    for (uint32_t s = 0; s < 4; ++s) {
        uint32_t tag = EMPTY;
        if (s < child_count[k]) {
            const uint32_t g = child_first[k] + s;
            if (gnode[g] != EMPTY) {
                tag = id[gnode[g]];
                if (tag >= m.node_count) tag = EMPTY;
            } else {
                const uint32_t l = leaf_id(m, lrank, gstart[g]);
                tag = l < m.leaf_count ? (LEAF_BIT | l) : EMPTY;
            }
        }
        w[6 + s] = tag;
    }
}

__global__ __launch_bounds__(256) void write_leaves(const uint32_t* gstart, const uint32_t* gcount,
                                                    const uint32_t* gmesh, uint32_t n, const MeshDev* meshes,
                                                    const uint32_t* lrank, uint32_t* words) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= n || gcount[g] > LEAF_TRIANGLES) return;
    const MeshDev m = meshes[gmesh[g]];
    const uint32_t l = leaf_id(m, lrank, gstart[g]);
    if (l >= m.leaf_count) return;
    words[m.leaves + 2ull * l] = 3 * (gstart[g] - m.tbase);
    words[m.leaves + 2ull * l + 1] = 3 * gcount[g];
}

__global__ __launch_bounds__(256) void write_tris(const MeshDev* meshes, uint32_t mesh_count, uint32_t tris,
                                                  const uint32_t* perm, const uint32_t* idx, uint32_t* words) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= tris) return;
    uint32_t lo = 0, hi = mesh_count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (meshes[mid].tbase <= i) lo = mid;
        else hi = mid;
    }
    const MeshDev m = meshes[lo];
    const uint32_t p = i - m.tbase, t = perm[i];
    for (int k = 0; k < 3; ++k) words[m.indices + 3ull * p + k] = idx[3ull * t + k];
    if (m.keep) words[m.prims + p] = t - m.tbase;
}

__global__ __launch_bounds__(256) void write_positions(const MeshDev* meshes, uint32_t mesh_count, uint32_t floats,
                                                       const float* pos, uint32_t* words) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= floats) return;
    const uint32_t v = i / 3;
    uint32_t lo = 0, hi = mesh_count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (meshes[mid].vbase <= v) lo = mid;
        else hi = mid;
    }
    const MeshDev m = meshes[lo];
    words[m.positions + (i - 3 * m.vbase)] = fb(pos[i]);
}

// Each leaf's box, folded from the written words (bvhdev::leaf_box), at its
// mesh's slice of `boxes`.
__global__ __launch_bounds__(256) void leaf_group_boxes(const uint32_t* gstart, const uint32_t* gcount,
                                                        const uint32_t* gmesh, uint32_t n, const MeshDev* meshes,
                                                        const uint32_t* lrank, const uint32_t* words, float* boxes) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= n || gcount[g] > LEAF_TRIANGLES) return;
    const MeshDev m = meshes[gmesh[g]];
    const uint32_t l = leaf_id(m, lrank, gstart[g]);
    if (l >= m.leaf_count) return;
    float lo[3], hi[3];
    bvhdev::leaf_box(words, m.leaves, m.positions, m.indices, l, lo, hi);
    for (int x = 0; x < 3; ++x) {
        boxes[6ull * (m.lbase + l) + x] = lo[x];
        boxes[6ull * (m.lbase + l) + 3 + x] = hi[x];
    }
}

__global__ __launch_bounds__(256) void refit_nodes(uint32_t nbase, uint32_t nn, const uint32_t* node_mesh,
                                                   const MeshDev* meshes, const uint32_t* id, const float* boxes,
                                                   uint32_t* words) {
    const uint32_t k = nbase + blockIdx.x * 256u + threadIdx.x;
    if (k >= nbase + nn) return;
    const MeshDev m = meshes[node_mesh[k]];
    if (id[k] < m.node_count)
        bvhdev::refit_node(words, m.nodes, m.leaves, m.positions, m.indices, id[k], boxes + 6ull * m.lbase);
}

}  // namespace gbuild
}  // namespace rq

using namespace rq;
using namespace rq::gbuild;

namespace {

rq_status hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RQ_ERR_OUT_OF_MEMORY : RQ_ERR_DEVICE;
}
#define GB_TRY(x, what)                                   \
    do {                                                  \
        hipError_t e_ = (x);                              \
        if (e_ != hipSuccess) return hip_fail(e_, what);  \
    } while (0)
#define GB_OK(x)                          \
    do {                                  \
        rq_status s_ = (x);               \
        if (s_ != RQ_OK) return s_;       \
    } while (0)

struct Buf {
    void* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    rq_status need(size_t bytes, const char* what) {
        if (bytes <= n) return RQ_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        GB_TRY(hipMalloc(&p, std::max<size_t>(bytes, 256)), what);
        n = std::max<size_t>(bytes, 256);
        return RQ_OK;
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

inline dim3 grid_for(uint64_t n, uint32_t block = 256) {
    return dim3((unsigned)std::max<uint64_t>(1, (n + block - 1) / block));
}

// The device side of one build: triangles of every mesh, the permutation, and
// the groups and nodes of every level.
struct GpuBuild {
    hipStream_t s = nullptr;
    uint32_t mesh_count = 0, tris = 0, verts = 0;
    Buf meshes, pos, idx, cent, box, perm, tmp;
    Buf gstart, gcount, gmesh, gnode;                // groups, all levels
    Buf node_group, node_mesh, node_level, child_first, child_count, size, start, id;
    Buf flag, scan, segs, small, large, split_left, counters, med, med_off, keys, keys_out;
    Buf chunks, first_chunk, chunk_n, chunk_left, chunk_off, lb, gcnt, gbins, best, seg_left, mark, lrank, summary;
    Buf cub;
    std::vector<uint32_t> level_gbase, level_nbase;  // per level; one past the end last
    uint32_t node_total = 0;

    rq_status exclusive_sum(const uint32_t* in, uint32_t* out, uint32_t n) {
        size_t bytes = 0;
        GB_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s),
               "scan size");
        GB_OK(cub.need(bytes, "hipMalloc scan"));
        GB_TRY(rocprim::exclusive_scan(cub.p, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s), "scan");
        return RQ_OK;
    }
    uint32_t* pinned = nullptr;  // host words for the small per-level readbacks
    ~GpuBuild() {
        if (pinned) (void)hipHostFree(pinned);
    }
    rq_status fetch(void* dst, const void* src, size_t bytes) {
        if (bytes <= 256 && !pinned && hipHostMalloc(reinterpret_cast<void**>(&pinned), 256, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            pinned = nullptr;
        }
        void* to = bytes <= 256 && pinned ? static_cast<void*>(pinned) : dst;
        GB_TRY(hipMemcpyAsync(to, src, bytes, hipMemcpyDeviceToHost, s), "build fetch");
        GB_TRY(hipStreamSynchronize(s), "build sync");
        if (to != dst) std::memcpy(dst, to, bytes);
        return RQ_OK;
    }
    // scan[n - 1] + flag[n - 1]: the number of flagged entries, read back once
    rq_status fetch_total(const Buf& scan_b, const Buf& flag_b, uint32_t n, uint32_t& out) {
        hipLaunchKernelGGL(scan_total, dim3(1), dim3(64), 0, s, scan_b.as<uint32_t>(), flag_b.as<uint32_t>(), n,
                           counters.as<uint32_t>() + 8);
        return fetch(&out, counters.as<uint32_t>() + 8, 4);
    }
    void fill(Buf& b, uint32_t v, uint64_t n) {
        hipLaunchKernelGGL(fill_u32, dim3((unsigned)std::min<uint64_t>((n + 255) / 256 + 1, 4096)), dim3(256), 0, s,
                           b.as<uint32_t>(), v, n);
    }

    // Splits `n` groups listed in segs (one pass: every group once).
    rq_status split_pass(uint32_t n) {
        if (!n) return RQ_OK;
        uint32_t* cnt = counters.as<uint32_t>();
        fill(counters, 0, 4);
        hipLaunchKernelGGL(route, grid_for(n), dim3(256), 0, s, segs.as<Seg>(), n, small.as<Seg>(), large.as<Seg>(),
                           cnt);
        uint32_t c[4];
        GB_OK(fetch(c, cnt, 16));
        const uint32_t n_small = c[0], n_large = c[1];
        if (n_small)
            hipLaunchKernelGGL(split_small, dim3(n_small), dim3(256), 0, s, small.as<Seg>(), meshes.as<MeshDev>(),
                               perm.as<uint32_t>(), cent.as<float>(), box.as<float>(), split_left.as<uint32_t>(),
                               med.as<MedianReq>(), cnt + 2);
        if (n_large) GB_OK(large_pass(n_large));
        GB_TRY(hipGetLastError(), "split launch");
        uint32_t n_med = 0;
        GB_OK(fetch(&n_med, cnt + 2, 4));
        if (n_med) GB_OK(median_pass(n_med));
        return RQ_OK;
    }

    rq_status large_pass(uint32_t n) {
        std::vector<Seg> segs_h(n);
        GB_OK(fetch(segs_h.data(), large.p, n * sizeof(Seg)));
        std::vector<Chunk> ch;
        std::vector<uint32_t> first(n), count(n);
        for (uint32_t j = 0; j < n; ++j) {
            first[j] = (uint32_t)ch.size();
            for (uint32_t o = 0; o < segs_h[j].count; o += CHUNK)
                ch.push_back(Chunk{j, segs_h[j].start + o, std::min(CHUNK, segs_h[j].count - o), 0});
            count[j] = (uint32_t)ch.size() - first[j];
        }
        const uint32_t nc = (uint32_t)ch.size();
        GB_OK(chunks.need(nc * sizeof(Chunk), "hipMalloc chunks"));
        GB_OK(first_chunk.need(n * 4, "hipMalloc chunks"));
        GB_OK(chunk_n.need(n * 4, "hipMalloc chunks"));
        GB_OK(chunk_left.need(nc * 4, "hipMalloc chunks"));
        GB_OK(chunk_off.need(nc * 4, "hipMalloc chunks"));
        GB_OK(lb.need(n * 6 * 4, "hipMalloc chunks"));
        GB_OK(gcnt.need((size_t)n * 3 * BINS * 4, "hipMalloc chunks"));
        GB_OK(gbins.need((size_t)n * 3 * BINS * 6 * 4, "hipMalloc chunks"));
        GB_OK(best.need(n * 4, "hipMalloc chunks"));
        GB_OK(seg_left.need(n * 4, "hipMalloc chunks"));
        GB_TRY(hipMemcpyAsync(chunks.p, ch.data(), nc * sizeof(Chunk), hipMemcpyHostToDevice, s), "chunks");
        GB_TRY(hipMemcpyAsync(first_chunk.p, first.data(), n * 4, hipMemcpyHostToDevice, s), "chunks");
        GB_TRY(hipMemcpyAsync(chunk_n.p, count.data(), n * 4, hipMemcpyHostToDevice, s), "chunks");
        std::vector<int32_t> init(n * 6);
        std::vector<int32_t> binit((size_t)n * 3 * BINS * 6);
        const int32_t pinf = (int32_t)0x7f800000, ninf = (int32_t)(0xff800000u ^ 0x7fffffffu);
        for (uint32_t j = 0; j < n; ++j)
            for (int k = 0; k < 6; ++k) init[6 * j + k] = k < 3 ? pinf : ninf;
        for (size_t b = 0; b < (size_t)n * 3 * BINS; ++b)
            for (int k = 0; k < 6; ++k) binit[b * 6 + k] = k < 3 ? pinf : ninf;
        GB_TRY(hipMemcpyAsync(lb.p, init.data(), init.size() * 4, hipMemcpyHostToDevice, s), "chunks");
        GB_TRY(hipMemcpyAsync(gbins.p, binit.data(), binit.size() * 4, hipMemcpyHostToDevice, s), "chunks");
        fill(gcnt, 0, (uint64_t)n * 3 * BINS);
        uint32_t* cnt = counters.as<uint32_t>();
        hipLaunchKernelGGL(large_bounds, dim3(nc), dim3(256), 0, s, chunks.as<Chunk>(), perm.as<uint32_t>(),
                           cent.as<float>(), lb.as<int32_t>());
        hipLaunchKernelGGL(large_bins, dim3(nc), dim3(256), 0, s, chunks.as<Chunk>(), large.as<Seg>(),
                           meshes.as<MeshDev>(), perm.as<uint32_t>(), cent.as<float>(), box.as<float>(),
                           lb.as<int32_t>(), gcnt.as<uint32_t>(), gbins.as<int32_t>());
        hipLaunchKernelGGL(large_eval, grid_for(n, 64), dim3(64), 0, s, large.as<Seg>(), n, meshes.as<MeshDev>(),
                           lb.as<int32_t>(), gcnt.as<uint32_t>(), gbins.as<int32_t>(), best.as<uint32_t>(),
                           med.as<MedianReq>(), cnt + 2);
        hipLaunchKernelGGL(large_count, dim3(nc), dim3(256), 0, s, chunks.as<Chunk>(), best.as<uint32_t>(),
                           perm.as<uint32_t>(), cent.as<float>(), lb.as<int32_t>(), chunk_left.as<uint32_t>());
        GB_OK(exclusive_sum(chunk_left.as<uint32_t>(), chunk_off.as<uint32_t>(), nc));
        hipLaunchKernelGGL(large_total, grid_for(n, 64), dim3(64), 0, s, large.as<Seg>(), n, best.as<uint32_t>(),
                           first_chunk.as<uint32_t>(), chunk_n.as<uint32_t>(), chunk_left.as<uint32_t>(),
                           chunk_off.as<uint32_t>(), seg_left.as<uint32_t>(), split_left.as<uint32_t>());
        hipLaunchKernelGGL(large_scatter, dim3(nc), dim3(256), 0, s, chunks.as<Chunk>(), large.as<Seg>(),
                           best.as<uint32_t>(), first_chunk.as<uint32_t>(), chunk_off.as<uint32_t>(),
                           seg_left.as<uint32_t>(), perm.as<uint32_t>(), cent.as<float>(), lb.as<int32_t>(),
                           tmp.as<uint32_t>());
        hipLaunchKernelGGL(large_copy, dim3(nc), dim3(256), 0, s, chunks.as<Chunk>(), best.as<uint32_t>(),
                           tmp.as<uint32_t>(), perm.as<uint32_t>());
        GB_TRY(hipGetLastError(), "large split launch");
        // the host vectors above must outlive their copies
        GB_TRY(hipStreamSynchronize(s), "large split");
        return RQ_OK;
    }

    rq_status median_pass(uint32_t n) {
        std::vector<MedianReq> q(n);
        GB_OK(fetch(q.data(), med.p, n * sizeof(MedianReq)));
        std::vector<uint32_t> off(n + 1, 0);
        for (uint32_t j = 0; j < n; ++j) off[j + 1] = off[j] + q[j].count;
        const uint32_t total = off[n];
        GB_OK(med_off.need((n + 1) * 4, "hipMalloc median"));
        GB_OK(keys.need((size_t)total * 8, "hipMalloc median"));
        GB_OK(keys_out.need((size_t)total * 8, "hipMalloc median"));
        GB_TRY(hipMemcpyAsync(med_off.p, off.data(), (n + 1) * 4, hipMemcpyHostToDevice, s), "median offsets");
        hipLaunchKernelGGL(median_keys, dim3(n), dim3(256), 0, s, med.as<MedianReq>(), med_off.as<uint32_t>(),
                           meshes.as<MeshDev>(), perm.as<uint32_t>(), cent.as<float>(), keys.as<uint64_t>());
        size_t bytes = 0;
        const uint32_t* o = med_off.as<uint32_t>();
        GB_TRY(rocprim::segmented_radix_sort_keys(nullptr, bytes, keys.as<uint64_t>(), keys_out.as<uint64_t>(), total,
                                                  n, o, o + 1, 0u, 64u, s),
               "median sort size");
        GB_OK(cub.need(bytes, "hipMalloc median sort"));
        GB_TRY(rocprim::segmented_radix_sort_keys(cub.p, bytes, keys.as<uint64_t>(), keys_out.as<uint64_t>(), total, n,
                                                  o, o + 1, 0u, 64u, s),
               "median sort");
        hipLaunchKernelGGL(median_put, dim3(n), dim3(256), 0, s, med.as<MedianReq>(), med_off.as<uint32_t>(),
                           meshes.as<MeshDev>(), keys_out.as<uint64_t>(), perm.as<uint32_t>(),
                           split_left.as<uint32_t>());
        GB_TRY(hipGetLastError(), "median launch");
        GB_TRY(hipStreamSynchronize(s), "median");
        return RQ_OK;
    }

    // The whole topology: every level's groups, nodes and their children.
    rq_status topology() {
        hipLaunchKernelGGL(prep_tris, grid_for(tris), dim3(256), 0, s, meshes.as<MeshDev>(), mesh_count,
                           pos.as<float>(), idx.as<uint32_t>(), tris, cent.as<float>(), box.as<float>(),
                           perm.as<uint32_t>());
        hipLaunchKernelGGL(root_groups, grid_for(mesh_count), dim3(256), 0, s, meshes.as<MeshDev>(), mesh_count,
                           gstart.as<uint32_t>(), gcount.as<uint32_t>(), gmesh.as<uint32_t>());
        level_gbase.assign(1, 0);
        level_nbase.assign(1, 0);
        uint32_t gb = 0, ng = mesh_count, nb = 0;
        for (uint32_t level = 0;; ++level) {
            hipLaunchKernelGGL(node_flags, grid_for(ng), dim3(256), 0, s, gcount.as<uint32_t>() + gb, ng,
                               flag.as<uint32_t>());
            GB_OK(exclusive_sum(flag.as<uint32_t>(), scan.as<uint32_t>(), ng));
            uint32_t nn = 0;
            GB_OK(fetch_total(scan, flag, ng, nn));
            hipLaunchKernelGGL(make_nodes, grid_for(ng), dim3(256), 0, s, gstart.as<uint32_t>(), gcount.as<uint32_t>(),
                               gmesh.as<uint32_t>(), gnode.as<uint32_t>(), gb, ng, scan.as<uint32_t>(), nb, level,
                               node_group.as<uint32_t>(), node_mesh.as<uint32_t>(), node_level.as<uint32_t>(),
                               segs.as<Seg>());
            level_gbase.push_back(gb + ng);
            if (nn == 0) {
                level_nbase.push_back(nb);
                break;
            }
            fill(split_left, EMPTY, 3ull * nn);
            GB_OK(split_pass(nn));  // round 0: each node's group in two
            fill(counters, 0, 4);
            hipLaunchKernelGGL(round1_segs, grid_for(nn), dim3(256), 0, s, gstart.as<uint32_t>(),
                               gcount.as<uint32_t>(), gmesh.as<uint32_t>(), node_group.as<uint32_t>(), nb, nn,
                               split_left.as<uint32_t>(), segs.as<Seg>(), counters.as<uint32_t>() + 3);
            uint32_t n1 = 0;
            GB_OK(fetch(&n1, counters.as<uint32_t>() + 3, 4));
            GB_OK(split_pass(n1));  // round 1: the halves still too large
            hipLaunchKernelGGL(child_counts, grid_for(nn), dim3(256), 0, s, gcount.as<uint32_t>(),
                               node_group.as<uint32_t>(), nb, nn, split_left.as<uint32_t>(), flag.as<uint32_t>());
            GB_OK(exclusive_sum(flag.as<uint32_t>(), scan.as<uint32_t>(), nn));
            uint32_t next = 0;
            GB_OK(fetch_total(scan, flag, nn, next));
            hipLaunchKernelGGL(emit_children, grid_for(nn), dim3(256), 0, s, gstart.as<uint32_t>(),
                               gcount.as<uint32_t>(), gmesh.as<uint32_t>(), node_group.as<uint32_t>(), nb, nn,
                               split_left.as<uint32_t>(), scan.as<uint32_t>(), gb + ng, child_first.as<uint32_t>(),
                               child_count.as<uint32_t>());
            GB_TRY(hipGetLastError(), "level launch");
            gb += ng;
            ng = next;
            nb += nn;
            level_nbase.push_back(nb);
        }
        node_total = nb;
        return RQ_OK;
    }

    // Post-order ids, leaf ranks and the per-mesh summary (nodes, leaves, root, depth).
    rq_status numbering(std::vector<uint32_t>& summary_h) {
        const size_t levels = level_nbase.size() - 1;  // levels with a node range (the last may be empty)
        for (size_t l = levels; l-- > 0;) {
            const uint32_t nb = level_nbase[l], nn = level_nbase[l + 1] - nb;
            if (nn)
                hipLaunchKernelGGL(node_sizes, grid_for(nn), dim3(256), 0, s, nb, nn, child_first.as<uint32_t>(),
                                   child_count.as<uint32_t>(), gnode.as<uint32_t>(), size.as<uint32_t>());
        }
        for (size_t l = 0; l < levels; ++l) {
            const uint32_t nb = level_nbase[l], nn = level_nbase[l + 1] - nb;
            if (nn)
                hipLaunchKernelGGL(node_ids, grid_for(nn), dim3(256), 0, s, nb, nn, l == 0, child_first.as<uint32_t>(),
                                   child_count.as<uint32_t>(), gnode.as<uint32_t>(), size.as<uint32_t>(),
                                   start.as<uint32_t>(), id.as<uint32_t>());
        }
        const uint32_t groups = level_gbase.back();
        fill(mark, 0, tris);
        hipLaunchKernelGGL(leaf_marks, grid_for(groups), dim3(256), 0, s, gstart.as<uint32_t>(), gcount.as<uint32_t>(),
                           groups, mark.as<uint32_t>());
        GB_OK(exclusive_sum(mark.as<uint32_t>(), lrank.as<uint32_t>(), tris));
        fill(summary, 0, 4ull * mesh_count);
        hipLaunchKernelGGL(mesh_summary, grid_for(mesh_count), dim3(256), 0, s, meshes.as<MeshDev>(), mesh_count,
                           gnode.as<uint32_t>(), size.as<uint32_t>(), id.as<uint32_t>(), mark.as<uint32_t>(),
                           lrank.as<uint32_t>(), summary.as<uint32_t>());
        if (node_total)
            hipLaunchKernelGGL(mesh_depth, grid_for(node_total), dim3(256), 0, s, node_mesh.as<uint32_t>(),
                               node_level.as<uint32_t>(), node_total, summary.as<uint32_t>());
        GB_TRY(hipGetLastError(), "numbering launch");
        summary_h.resize(4ull * mesh_count);
        return fetch(summary_h.data(), summary.p, summary_h.size() * 4);
    }

    rq_status allocate() {
        const uint64_t T = std::max<uint32_t>(tris, 1), G = 2 * T + mesh_count + 16, N = T + 16;
        GB_OK(cent.need(T * 12, "hipMalloc build"));
        GB_OK(box.need(T * 24, "hipMalloc build"));
        GB_OK(perm.need(T * 4, "hipMalloc build"));
        GB_OK(tmp.need(T * 4, "hipMalloc build"));
        for (Buf* b : {&gstart, &gcount, &gmesh, &gnode, &flag, &scan}) GB_OK(b->need(G * 4, "hipMalloc build"));
        for (Buf* b : {&node_group, &node_mesh, &node_level, &child_first, &child_count, &size, &start, &id})
            GB_OK(b->need(N * 4, "hipMalloc build"));
        GB_OK(segs.need(N * sizeof(Seg), "hipMalloc build"));
        GB_OK(small.need(N * sizeof(Seg), "hipMalloc build"));
        GB_OK(large.need(N * sizeof(Seg), "hipMalloc build"));
        GB_OK(med.need(N * sizeof(MedianReq), "hipMalloc build"));
        GB_OK(split_left.need(3 * N * 4, "hipMalloc build"));
        GB_OK(counters.need(64, "hipMalloc build"));  // [3] round-1 count, [8] scan totals
        GB_OK(mark.need(T * 4, "hipMalloc build"));
        GB_OK(lrank.need(T * 4, "hipMalloc build"));
        GB_OK(summary.need(16ull * mesh_count + 16, "hipMalloc build"));
        return RQ_OK;
    }
};

struct HostMesh {
    float lo[3], hi[3];
    std::string err;
    bool ok = true;
};

// Pinned host staging for the build's two uploads, kept between builds:
// pinning hundreds of megabytes costs more than a build, and an upload from
// pageable memory runs at a fraction of the pinned rate. The process keeps the
// largest buffer it has needed (one per concurrent build at most) until it ends.
struct PinnedPool {
    std::mutex mu;
    std::vector<std::pair<void*, size_t>> free;  // (pointer, bytes)
};
PinnedPool& pinned_pool() {
    static PinnedPool* pool = new PinnedPool();  // never destroyed: outlives every build
    return *pool;
}

// One build's staging: taken from the pool, returned once the stream has
// finished reading it (the destructor waits for the stream).
struct Staging {
    void* p = nullptr;
    size_t bytes = 0;
    hipStream_t s = nullptr;
    bool queued = false;  // an upload from this buffer may still be in flight
    std::vector<uint32_t> pageable;  // used instead when pinned memory is refused
    rq_status take(size_t need) {
        PinnedPool& pool = pinned_pool();
        {
            std::lock_guard<std::mutex> lock(pool.mu);
            size_t best = pool.free.size();
            for (size_t i = 0; i < pool.free.size(); ++i)
                if (pool.free[i].second >= need && (best == pool.free.size() || pool.free[i].second < pool.free[best].second))
                    best = i;
            if (best < pool.free.size()) {
                p = stage = pool.free[best].first;
                bytes = pool.free[best].second;
                pool.free.erase(pool.free.begin() + best);
                return RQ_OK;
            }
        }
        const size_t grain = need <= (1u << 20) ? (1u << 20) : (16u << 20);
        bytes = (need + grain - 1) & ~(grain - 1);
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
            // pinned memory exhausted: stage in ordinary memory (slower uploads)
            (void)hipGetLastError();
            p = nullptr;
            try {
                pageable.resize((need + 3) / 4);
            } catch (const std::bad_alloc&) {
                set_error("rq_scene_build_gpu: no host memory for the upload staging");
                return RQ_ERR_OUT_OF_MEMORY;
            }
            stage = pageable.data();
            return RQ_OK;
        }
        stage = p;
        return RQ_OK;
    }
    void* stage = nullptr;  // where the meshes are gathered: the pinned buffer or `pageable`
    ~Staging() {
        if (queued) (void)hipStreamSynchronize(s);
        if (!p) return;
        PinnedPool& pool = pinned_pool();
        std::lock_guard<std::mutex> lock(pool.mu);
        // keep this buffer; drop pooled ones it makes redundant (smaller ones)
        for (size_t i = 0; i < pool.free.size();)
            if (pool.free[i].second <= bytes) {
                (void)hipHostFree(pool.free[i].first);
                pool.free.erase(pool.free.begin() + i);
            } else {
                ++i;
            }
        pool.free.emplace_back(p, bytes);
    }
};

}  // namespace

extern "C" rq_status rq_scene_build_gpu(const rq_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                                        uint32_t instance_count, int top_level, rq_scene** out, void* stream) {
    const auto began = std::chrono::steady_clock::now();
    if (!out || (!meshes && mesh_count) || (!instances && instance_count)) {
        set_error("rq_scene_build_gpu: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    for (uint32_t i = 0; i < instance_count; ++i)
        if (instances[i].mesh >= mesh_count) {
            set_error("rq_scene_build_gpu: instance names a missing mesh");
            return RQ_ERR_INVALID_INPUT;
        }
    // The host builder's input checks (near_field.rs:206-212) and the mesh
    // bounds over all positions (src/gi_bvh.rs:64-87), one pass per mesh that
    // also gathers every mesh into one staging array each (two uploads in all).
    std::vector<uint64_t> vbase(mesh_count + 1, 0), tbase(mesh_count + 1, 0);
    for (uint32_t m = 0; m < mesh_count; ++m) {
        vbase[m + 1] = vbase[m] + meshes[m].vertex_count;
        tbase[m + 1] = tbase[m] + meshes[m].index_count / 3;
    }
    if (tbase[mesh_count] >= (1ull << 31) || 3 * vbase[mesh_count] >= (1ull << 32)) {
        set_error("rq_scene_build_gpu: too many triangles or vertices for one build");
        return RQ_ERR_INVALID_INPUT;
    }
    Staging staging;
    staging.s = reinterpret_cast<hipStream_t>(stream);
    if (vbase[mesh_count] + tbase[mesh_count]) GB_OK(staging.take(12ull * (vbase[mesh_count] + tbase[mesh_count])));
    float* const stage_pos = static_cast<float*>(staging.stage);
    uint32_t* const stage_idx = stage_pos ? reinterpret_cast<uint32_t*>(stage_pos + 3 * vbase[mesh_count]) : nullptr;
    std::vector<HostMesh> hm(mesh_count);
    {
        uint32_t threads = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), 64));
        std::vector<std::thread> pool;
        std::atomic_uint next{0};
        for (uint32_t t = 0; t < std::min(threads, std::max(mesh_count, 1u)); ++t)
            pool.emplace_back([&]() {
                for (;;) {
                    const uint32_t m = next.fetch_add(1);
                    if (m >= mesh_count) break;
                    const rq_mesh& src = meshes[m];
                    HostMesh& h = hm[m];
                    if (src.index_count == 0 || src.index_count % 3 != 0 || !src.indices) {
                        h.ok = false;
                        h.err = "BVH geometry needs a nonempty triangle list";
                        continue;
                    }
                    for (uint32_t k = 0; k < src.index_count && h.ok; ++k)
                        if (src.indices[k] >= src.vertex_count) {
                            h.ok = false;
                            h.err = "BVH geometry has an invalid index or position";
                        }
                    for (int a = 0; a < 3; ++a) {
                        h.lo[a] = __builtin_inff();
                        h.hi[a] = -__builtin_inff();
                    }
                    for (uint64_t v = 0; v < src.vertex_count && h.ok; ++v)
                        for (int a = 0; a < 3; ++a) {
                            const float p = src.positions[3 * v + a];
                            if (!std::isfinite(p)) {
                                h.ok = false;
                                h.err = "BVH geometry has an invalid index or position";
                            }
                            h.lo[a] = mn(h.lo[a], p);
                            h.hi[a] = mx(h.hi[a], p);
                        }
                    if (!h.ok) continue;
                    std::memcpy(&stage_pos[3 * vbase[m]], src.positions, 12ull * src.vertex_count);
                    std::memcpy(&stage_idx[3 * tbase[m]], src.indices, 4ull * (src.index_count / 3) * 3);
                }
            });
        for (auto& th : pool) th.join();
    }
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (!hm[m].ok) {
            set_error("indirect geometry is not a valid triangle list: " + hm[m].err);
            return RQ_ERR_INVALID_INPUT;
        }
    std::vector<float> inverses(16ull * instance_count);
    for (uint32_t i = 0; i < instance_count; ++i)
        if (!mat4_inverse(instances[i].model, &inverses[16ull * i])) {
            set_error("indirect geometry transform is singular");
            return RQ_ERR_INVALID_INPUT;
        }
    // src/gi_bvh.rs:230-274 (instance_words) + src/gi_cache.rs:4388-4479, as rq_scene_build:
    // the header, the records and the top level need only the transforms and the mesh
    // bounds, so a host thread builds them while the device builds the trees; each
    // record's mesh offsets and root go in once the device has numbered the nodes.
    std::vector<uint32_t> head(HEADER_WORDS + (size_t)instance_count * INSTANCE_WORDS, 0u);
    struct Joiner {
        std::thread t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } head_job;
    head_job.t = std::thread([&]() {
        head[0] = instance_count;
        for (uint32_t i = 0; i < instance_count; ++i) {
            const rq_instance& ins = instances[i];
            const float* inv = &inverses[16ull * i];
            const size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
            for (int row = 0; row < 3; ++row)
                for (int c = 0; c < 4; ++c) head[r + row * 4 + c] = f_bits(inv[c * 4 + row]);
            instance_box(ins.model, hm[ins.mesh].lo, hm[ins.mesh].hi, &head[r + 12]);
            head[r + 24] = ins.identity;
            head[r + 27] = ins.flags & 1u;
            head[r + 29] = ins.placement;
        }
        head[4] = instance_count;
        top_place(head, HEADER_WORDS, INSTANCE_WORDS, instance_count, nullptr, 0, top_level);
    });
    int device = 0;
    GB_TRY(hipGetDevice(&device), "hipGetDevice");
    GpuBuild b;
    b.s = reinterpret_cast<hipStream_t>(stream);
    b.mesh_count = mesh_count;
    std::vector<MeshDev> md(mesh_count);
    const uint64_t tris = tbase[mesh_count], verts = vbase[mesh_count];
    for (uint32_t m = 0; m < mesh_count; ++m) {
        std::memset(&md[m], 0, sizeof(MeshDev));
        md[m].tbase = (uint32_t)tbase[m];
        md[m].tcount = meshes[m].index_count / 3;
        md[m].vbase = (uint32_t)vbase[m];
        md[m].vcount = meshes[m].vertex_count;
        md[m].mode = MODE_SAH;
        md[m].keep = meshes[m].keep_primitives ? 1u : 0u;
    }
    b.tris = (uint32_t)tris;
    b.verts = (uint32_t)verts;
    GB_OK(b.meshes.need(std::max<size_t>(mesh_count, 1) * sizeof(MeshDev), "hipMalloc build"));
    GB_OK(b.pos.need(std::max<uint64_t>(verts, 1) * 12, "hipMalloc build"));
    GB_OK(b.idx.need(std::max<uint64_t>(tris, 1) * 12, "hipMalloc build"));
    staging.queued = true;
    if (verts)
        GB_TRY(hipMemcpyAsync(b.pos.p, stage_pos, 12ull * verts, hipMemcpyHostToDevice, b.s), "upload positions");
    if (tris)
        GB_TRY(hipMemcpyAsync(b.idx.p, stage_idx, 12ull * tris, hipMemcpyHostToDevice, b.s), "upload indices");
    GB_OK(b.allocate());
    std::vector<uint32_t> summary;
    if (mesh_count) {
        GB_TRY(hipMemcpyAsync(b.meshes.p, md.data(), mesh_count * sizeof(MeshDev), hipMemcpyHostToDevice, b.s),
               "upload meshes");
        GB_OK(b.topology());
        GB_OK(b.numbering(summary));
        // near_field.rs:184-190: too deep with SAH, median splits throughout instead
        bool again = false;
        for (uint32_t m = 0; m < mesh_count; ++m)
            if (summary[4 * m + 3] > MAX_SAH_DEPTH) {
                md[m].mode = MODE_MEDIAN;
                again = true;
            }
        if (again) {
            GB_TRY(hipMemcpyAsync(b.meshes.p, md.data(), mesh_count * sizeof(MeshDev), hipMemcpyHostToDevice, b.s),
                   "upload meshes");
            GB_OK(b.topology());
            GB_OK(b.numbering(summary));
        }
    }
    // src/gi_bvh.rs:151-179 (layout)
    const uint32_t* sm = summary.data();
    uint64_t at = HEADER_WORDS + (uint64_t)instance_count * INSTANCE_WORDS + top_reserve(instance_count);
    std::vector<uint32_t> offsets(4ull * mesh_count);
    uint64_t leaf_total = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        const uint64_t nodes = at, leaves = nodes + (uint64_t)sm[4 * m] * NODE_WORDS;
        const uint64_t positions = leaves + 2ull * sm[4 * m + 1], indices = positions + 3ull * md[m].vcount;
        const uint64_t end = indices + 3ull * md[m].tcount + (md[m].keep ? md[m].tcount : 0);
        if (end >= 0xffffffffull) {
            set_error("packed geometry exceeds 2^32 words");
            return RQ_ERR_INVALID_INPUT;
        }
        md[m].nodes = offsets[4 * m] = (uint32_t)nodes;
        md[m].leaves = offsets[4 * m + 1] = (uint32_t)leaves;
        md[m].positions = offsets[4 * m + 2] = (uint32_t)positions;
        md[m].indices = offsets[4 * m + 3] = (uint32_t)indices;
        md[m].prims = (uint32_t)(indices + 3ull * md[m].tcount);
        md[m].node_count = sm[4 * m];
        md[m].leaf_count = sm[4 * m + 1];
        md[m].lbase = (uint32_t)leaf_total;
        leaf_total += sm[4 * m + 1];
        at = end;
    }
    // The host holds the header, records and top level; the meshes' words stay on
    // the device until a host reader asks for them (sync_host_words).
    auto scene = new rq_scene();
    std::vector<uint32_t>& words = scene->words;
    const size_t head_words = HEADER_WORDS + (size_t)instance_count * INSTANCE_WORDS + top_reserve(instance_count);
    words.assign(head_words, 0u);
    scene->total_words = at;
    DeviceLayout& d = scene->dev[0];
    d.bytes = (at * 4 + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&d.base, d.bytes);
    if (e != hipSuccess) {
        delete scene;
        return hip_fail(e, "hipMalloc scene");
    }
    auto fail_scene = [&](rq_status st) {
        rq_scene_destroy(scene);
        return st;
    };
    uint32_t* w = reinterpret_cast<uint32_t*>(d.base);
    if (hipMemsetAsync(w, 0, d.bytes, b.s) != hipSuccess) return fail_scene(hip_fail(hipGetLastError(), "memset"));
    if (mesh_count) {
        if (hipMemcpyAsync(b.meshes.p, md.data(), mesh_count * sizeof(MeshDev), hipMemcpyHostToDevice, b.s) != hipSuccess)
            return fail_scene(RQ_ERR_DEVICE);
        if (b.node_total)
            hipLaunchKernelGGL(write_nodes, grid_for(b.node_total), dim3(256), 0, b.s, b.node_total,
                               b.node_mesh.as<uint32_t>(), b.meshes.as<MeshDev>(), b.id.as<uint32_t>(),
                               b.child_first.as<uint32_t>(), b.child_count.as<uint32_t>(), b.gnode.as<uint32_t>(),
                               b.gstart.as<uint32_t>(), b.lrank.as<uint32_t>(), w);
        const uint32_t groups = b.level_gbase.back();
        hipLaunchKernelGGL(write_leaves, grid_for(groups), dim3(256), 0, b.s, b.gstart.as<uint32_t>(),
                           b.gcount.as<uint32_t>(), b.gmesh.as<uint32_t>(), groups, b.meshes.as<MeshDev>(),
                           b.lrank.as<uint32_t>(), w);
        hipLaunchKernelGGL(write_tris, grid_for(b.tris), dim3(256), 0, b.s, b.meshes.as<MeshDev>(), mesh_count, b.tris,
                           b.perm.as<uint32_t>(), b.idx.as<uint32_t>(), w);
        hipLaunchKernelGGL(write_positions, grid_for(3ull * b.verts), dim3(256), 0, b.s, b.meshes.as<MeshDev>(),
                           mesh_count, 3 * b.verts, b.pos.as<float>(), w);
        // leaf boxes, then node boxes, deepest level first (children are final before parents)
        if (b.box.need(std::max<uint64_t>(leaf_total, 1) * 24, "hipMalloc build") != RQ_OK)
            return fail_scene(RQ_ERR_OUT_OF_MEMORY);
        hipLaunchKernelGGL(leaf_group_boxes, grid_for(groups), dim3(256), 0, b.s, b.gstart.as<uint32_t>(),
                           b.gcount.as<uint32_t>(), b.gmesh.as<uint32_t>(), groups, b.meshes.as<MeshDev>(),
                           b.lrank.as<uint32_t>(), w, b.box.as<float>());
        for (size_t l = b.level_nbase.size() - 1; l-- > 0;) {
            const uint32_t nb = b.level_nbase[l], nn = b.level_nbase[l + 1] - nb;
            if (nn)
                hipLaunchKernelGGL(refit_nodes, grid_for(nn), dim3(256), 0, b.s, nb, nn, b.node_mesh.as<uint32_t>(),
                                   b.meshes.as<MeshDev>(), b.id.as<uint32_t>(), b.box.as<float>(), w);
        }
        if (hipGetLastError() != hipSuccess) return fail_scene(hip_fail(hipGetLastError(), "build words launch"));
    }
    head_job.t.join();
    for (uint32_t i = 0; i < instance_count; ++i) {
        const uint32_t m = instances[i].mesh;
        const size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        for (int k = 0; k < 4; ++k) head[r + 18 + k] = offsets[4 * m + k];
        head[r + 22] = sm[4 * m + 2];
        head[r + 28] = md[m].keep ? md[m].prims : 0u;
    }
    if (head.size() != head_words) return fail_scene(RQ_ERR_DEVICE);  // top_place appends exactly the reserve
    std::copy(head.begin(), head.end(), words.begin());
    if (hipMemcpyAsync(w, words.data(), head.size() * 4, hipMemcpyHostToDevice, b.s) != hipSuccess ||
        hipStreamSynchronize(b.s) != hipSuccess)
        return fail_scene(hip_fail(hipGetLastError(), "build head"));
    if (at > head.size()) scene->stale.emplace_back(head.size(), at - head.size());
    d.ready = true;
    scene->device = device;
    uint64_t tri_inst = 0, nodes = 0;
    uint32_t depth = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        nodes += sm[4 * m];
        depth = std::max(depth, sm[4 * m + 3]);
    }
    for (uint32_t i = 0; i < instance_count; ++i) tri_inst += md[instances[i].mesh].tcount;
    scene->stats[0] = instance_count;
    scene->stats[1] = mesh_count;
    scene->stats[2] = tri_inst;
    scene->stats[3] = tris;
    scene->stats[4] = nodes;
    scene->stats[5] = words[2];
    scene->stats[6] = depth;
    scene->models.resize(16ull * instance_count);
    for (uint32_t i = 0; i < instance_count; ++i)
        std::copy(instances[i].model, instances[i].model + 16, &scene->models[16ull * i]);
    scene->mesh_nodes.resize(mesh_count);
    for (uint32_t m = 0; m < mesh_count; ++m) scene->mesh_nodes[m] = offsets[4 * m];
    scene->mesh_map.resize(9ull * mesh_count);
    for (uint32_t m = 0; m < mesh_count; ++m) {
        uint32_t* o = &scene->mesh_map[9ull * m];
        o[0] = md[m].nodes;
        o[1] = md[m].leaves;
        o[2] = md[m].positions;
        o[3] = md[m].indices;
        o[4] = md[m].keep ? md[m].prims : 0u;
        o[5] = md[m].node_count;
        o[6] = md[m].leaf_count;
        o[7] = md[m].tcount;
        o[8] = sm[4 * m + 2];  // root tag
    }
    scene->top_force = top_level;
    scene_bounds(scene);
    scene->stats[7] = (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                          std::chrono::steady_clock::now() - began).count();
    *out = scene;
    return RQ_OK;
}

// ----------------------------------------------------------------- reflection words
// rq_scene_build_reflection's words (src/reflection_software.rs:92-388) from the
// GPU build of the same meshes: the same SAH trees (node numbering unchanged),
// RS leaf tags LEAF | first triangle << 4 | count, and 12-word triangles
// (corner a, b - a, c - a, source primitive, 0, 0) gathered in BVH order.
namespace rq {
namespace gbuild {

struct RsMesh {
    uint32_t nodes, leaves, positions, indices, prims;  // GI word offsets of the mesh
    uint32_t rs_nodes, rs_tris;                          // RS word offsets
    uint32_t node_first, tri_first;                      // first global node / triangle
    uint32_t root;                                       // GI root tag
};

__device__ __forceinline__ uint32_t rs_mesh_of(const RsMesh* ms, uint32_t n, uint32_t i, bool tri) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if ((tri ? ms[mid].tri_first : ms[mid].node_first) <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// A GI child tag as the reflection words tag it (leaves by triangle range).
__device__ __forceinline__ uint32_t rs_tag(const uint32_t* gi, const RsMesh& m, uint32_t c) {
    if (c == EMPTY || !(c & LEAF_BIT)) return c;
    const uint32_t l = c & ~LEAF_BIT;
    const uint32_t first = gi[m.leaves + 2ull * l], count = gi[m.leaves + 2ull * l + 1];
    return LEAF_BIT | ((first / 3u) << 4) | (count / 3u);
}

__global__ __launch_bounds__(256) void rs_nodes(const RsMesh* ms, uint32_t mesh_count, uint32_t total,
                                                const uint32_t* gi, uint32_t* rs) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const RsMesh m = ms[rs_mesh_of(ms, mesh_count, i, false)];
    const uint64_t n = i - m.node_first;
    const uint32_t* src = gi + m.nodes + 16ull * n;
    uint32_t* dst = rs + m.rs_nodes + 16ull * n;
    for (int k = 0; k < 16; ++k) dst[k] = (k >= 6 && k < 10) ? rs_tag(gi, m, src[k]) : src[k];
}

__global__ __launch_bounds__(256) void rs_tris(const RsMesh* ms, uint32_t mesh_count, uint32_t total,
                                               const uint32_t* gi, uint32_t* rs) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const RsMesh m = ms[rs_mesh_of(ms, mesh_count, i, true)];
    const uint64_t t = i - m.tri_first;
    const float* P = reinterpret_cast<const float*>(gi + m.positions);
    const uint32_t a = gi[m.indices + 3 * t], b = gi[m.indices + 3 * t + 1], c = gi[m.indices + 3 * t + 2];
    uint32_t* dst = rs + m.rs_tris + 12ull * t;
    for (int k = 0; k < 3; ++k) {
        const float pa = P[3ull * a + k];
        dst[k] = __float_as_uint(pa);
        dst[3 + k] = __float_as_uint(P[3ull * b + k] - pa);
        dst[6 + k] = __float_as_uint(P[3ull * c + k] - pa);
    }
    dst[9] = m.prims ? gi[m.prims + t] : (uint32_t)t;
    dst[10] = 0u;
    dst[11] = 0u;
}

__global__ __launch_bounds__(256) void rs_roots(const RsMesh* ms, uint32_t mesh_count, const uint32_t* gi,
                                                uint32_t* out) {
    const uint32_t m = blockIdx.x * 256u + threadIdx.x;
    if (m < mesh_count) out[m] = rs_tag(gi, ms[m], ms[m].root);
}

}  // namespace gbuild
}  // namespace rq

extern "C" rq_status rq_scene_build_reflection_gpu(const rq_mesh* meshes, uint32_t mesh_count,
                                                   const rq_instance* instances, uint32_t instance_count,
                                                   int top_level, rq_scene** out, void* stream) {
    using namespace rq::gbuild;
    const auto began = std::chrono::steady_clock::now();
    if (!out || (!meshes && mesh_count) || (!instances && instance_count)) {
        set_error("rq_scene_build_reflection_gpu: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    for (uint32_t i = 0; i < instance_count; ++i)
        if (instances[i].mesh >= mesh_count) {
            set_error("rq_scene_build_reflection_gpu: instance names a missing mesh");
            return RQ_ERR_INVALID_INPUT;
        }
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (meshes[m].index_count / 3u >= RS_MAX_MESH_TRIS) {
            set_error("rq_scene_build_reflection_gpu: a mesh of 2^27 or more triangles (the leaf tag holds the "
                      "first triangle in 27 bits)");
            return RQ_ERR_INVALID_INPUT;
        }
    // The trees: the GI build of the meshes alone, every mesh with its primitive table.
    std::vector<rq_mesh> with_prims(meshes, meshes + mesh_count);
    for (auto& m : with_prims) m.keep_primitives = 1;
    rq_scene* gi = nullptr;
    rq_status st = rq_scene_build_gpu(with_prims.data(), mesh_count, nullptr, 0, -1, &gi, stream);
    if (st != RQ_OK) return st;
    struct Drop {
        rq_scene* s;
        ~Drop() { rq_scene_destroy(s); }
    } drop{gi};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // Object bounds over the referenced corners (bounds(indices.map(positions))), per mesh.
    std::vector<float> bounds(6ull * mesh_count);
    {
        const uint32_t threads = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), 64));
        std::vector<std::thread> pool;
        std::atomic_uint next{0};
        for (uint32_t t = 0; t < std::min(threads, std::max(mesh_count, 1u)); ++t)
            pool.emplace_back([&]() {
                for (;;) {
                    const uint32_t m = next.fetch_add(1);
                    if (m >= mesh_count) break;
                    float* b = &bounds[6ull * m];
                    for (int a = 0; a < 3; ++a) {
                        b[a] = __builtin_inff();
                        b[3 + a] = -__builtin_inff();
                    }
                    for (uint32_t k = 0; k < meshes[m].index_count; ++k)
                        for (int a = 0; a < 3; ++a) {
                            const float p = meshes[m].positions[3ull * meshes[m].indices[k] + a];
                            b[a] = mn(b[a], p);
                            b[3 + a] = mx(b[3 + a], p);
                        }
                }
            });
        for (auto& th : pool) th.join();
    }
    // pack (src/reflection_software.rs:285-305): each mesh's nodes then triangles
    // after the header, records, top level and scene bounds.
    std::vector<RsMesh> rm(mesh_count);
    uint64_t at = rs_records_words(instance_count), node_total = 0, tri_total = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        const uint32_t* g = &gi->mesh_map[9ull * m];
        RsMesh& r = rm[m];
        r.nodes = g[0];
        r.leaves = g[1];
        r.positions = g[2];
        r.indices = g[3];
        r.prims = g[4];
        r.root = g[8];
        r.node_first = (uint32_t)node_total;
        r.tri_first = (uint32_t)tri_total;
        r.rs_nodes = (uint32_t)at;
        at += 16ull * g[5];
        r.rs_tris = (uint32_t)at;
        at += 12ull * g[7];
        node_total += g[5];
        tri_total += g[7];
        if (at >= 0xffffffffull) {
            set_error("reflection geometry exceeds 2^32 words");
            return RQ_ERR_INVALID_INPUT;
        }
    }
    rq::DeviceLayout d;
    d.bytes = (at * 4 + 255) & ~(size_t)255;
    GB_TRY(hipMalloc(&d.base, d.bytes), "hipMalloc reflection scene");
    struct Own {
        void* p;
        ~Own() {
            if (p) (void)hipFree(p);
        }
    } own{d.base};
    uint32_t* rs = reinterpret_cast<uint32_t*>(d.base);
    const uint32_t* gw = reinterpret_cast<const uint32_t*>(gi->dev[0].base);
    Buf dev_meshes, roots;
    GB_OK(dev_meshes.need(std::max<size_t>(mesh_count, 1) * sizeof(RsMesh), "hipMalloc reflection meshes"));
    GB_OK(roots.need(std::max<size_t>(mesh_count, 1) * 4, "hipMalloc reflection roots"));
    GB_TRY(hipMemsetAsync(rs, 0, d.bytes, s), "memset reflection scene");
    std::vector<uint32_t> root_tags(mesh_count);
    if (mesh_count) {
        GB_TRY(hipMemcpyAsync(dev_meshes.p, rm.data(), mesh_count * sizeof(RsMesh), hipMemcpyHostToDevice, s),
               "upload reflection meshes");
        if (node_total)
            hipLaunchKernelGGL(rs_nodes, grid_for(node_total), dim3(256), 0, s, dev_meshes.as<RsMesh>(), mesh_count,
                               (uint32_t)node_total, gw, rs);
        if (tri_total)
            hipLaunchKernelGGL(rs_tris, grid_for(tri_total), dim3(256), 0, s, dev_meshes.as<RsMesh>(), mesh_count,
                               (uint32_t)tri_total, gw, rs);
        hipLaunchKernelGGL(rs_roots, grid_for(mesh_count), dim3(256), 0, s, dev_meshes.as<RsMesh>(), mesh_count, gw,
                           roots.as<uint32_t>());
        GB_TRY(hipGetLastError(), "reflection words launch");
        GB_TRY(hipMemcpyAsync(root_tags.data(), roots.p, mesh_count * 4ull, hipMemcpyDeviceToHost, s), "roots");
        GB_TRY(hipStreamSynchronize(s), "reflection words");
    }
    // instance_words (src/reflection_software.rs:318-345) as rq_scene_build_reflection:
    // singular models are skipped; then the top level and the scene bounds.
    std::vector<uint32_t> head(RS_HEADER_WORDS, 0u);
    uint32_t count = 0;
    for (uint32_t i = 0; i < instance_count; ++i) {
        const rq_instance& ins = instances[i];
        float inv[16];
        if (!mat4_inverse(ins.model, inv)) continue;
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 4; ++c) head.push_back(f_bits(inv[c * 4 + row]));
        float lo[3], hi[3];
        rs_world_bounds(ins.model, &bounds[6ull * ins.mesh], &bounds[6ull * ins.mesh + 3], lo, hi);
        for (int a = 0; a < 3; ++a) head.push_back(f_bits(lo[a]));
        for (int a = 0; a < 3; ++a) head.push_back(f_bits(hi[a]));
        head.push_back(rm[ins.mesh].rs_nodes);
        head.push_back(rm[ins.mesh].rs_tris);
        head.push_back(root_tags[ins.mesh]);
        head.push_back(ins.placement);
        head.push_back(ins.flags & 1u);
        head.push_back(0u);
        ++count;
    }
    head[0] = count;
    top_place(head, RS_HEADER_WORDS, RS_RECORD_WORDS, instance_count, nullptr, 0, top_level);
    {
        float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        for (uint32_t i = 0; i < count; ++i) {
            const size_t r = RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS;
            for (int a = 0; a < 3; ++a) {
                lo[a] = mn(lo[a], bits_f(head[r + 12 + a]));
                hi[a] = mx(hi[a], bits_f(head[r + 15 + a]));
            }
        }
        head[4] = (uint32_t)head.size();
        for (int a = 0; a < 3; ++a) head.push_back(f_bits(lo[a]));
        head.push_back(0u);
        for (int a = 0; a < 3; ++a) head.push_back(f_bits(hi[a]));
        head.push_back(0u);
    }
    if (head.size() > at) {
        set_error("rq_scene_build_reflection_gpu: header past the mesh words");
        return RQ_ERR_DEVICE;
    }
    // The host holds the header, records, top level and bounds; the mesh words stay
    // on the device until a host reader asks for them (sync_host_words).
    auto scene = new rq_scene();
    scene->kind = RQ_SCENE_REFLECTION;
    scene->words = head;
    scene->total_words = at;
    if (at > head.size()) scene->stale.emplace_back(head.size(), at - head.size());
    GB_TRY(hipMemcpyAsync(rs, head.data(), head.size() * 4, hipMemcpyHostToDevice, s), "reflection head");
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        delete scene;
        return hip_fail(e, "reflection words");
    }
    int device = 0;
    (void)hipGetDevice(&device);
    scene->device = device;
    rq::DeviceLayout& dl = scene->dev[RQ_LAYOUT_RS_BVH4Q8];
    dl.base = d.base;
    dl.bytes = d.bytes;
    dl.ready = true;
    own.p = nullptr;
    uint64_t tris = 0;
    for (uint32_t i = 0; i < instance_count; ++i) tris += gi->mesh_map[9ull * instances[i].mesh + 7];
    scene->stats[0] = count;
    scene->stats[1] = mesh_count;
    scene->stats[2] = tris;
    scene->stats[3] = tri_total;
    scene->stats[4] = node_total;
    scene->stats[5] = scene->words[2];
    scene_bounds(scene);
    scene->stats[7] = (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                          std::chrono::steady_clock::now() - began).count();
    *out = scene;
    return RQ_OK;
}
