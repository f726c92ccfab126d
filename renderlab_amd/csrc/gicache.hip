// The GI radiance cache's ray stages on either side of the traversal, read
// straight from the cache's table (the `gi_table` words, shaders/gi_cells.wgsl):
//
// - gi_ray (shaders/gi_cache.wgsl:1563-1576): one visibility ray per ray item
//   (slot << 7 | k) -- from the cell's query point lifted off its surface
//   (gi_origin_offset, :1064-1068) along direction k of the cell's hemisphere
//   (gi_direction, :1048-1054; gi_basis and gi_unpack_normal, gi_cells.wgsl:
//   463-487) -- traced with gi_visible into the staged bits at the item index.
// - gi_gather's fixed directions (shaders/gi_gather_software.wgsl:82-110):
//   for each listed cell, the directions phase + j * stride whose visibility
//   bit (GI_C_BITS) is clear, as closest-hit rays, compacted cell by cell.
//
// Table layout (gi_cells.wgsl:18-185): a 64-word header (word 1 the hot-slot
// mask, 2 flags, 5 the smallest cell edge, 21 the gather stride, 23 the cold
// record words), then 24-word hot records (word 0 the key's first half), then
// cold records of max(24, header[23]) words (0 the key's second half, 3 the
// level in the low byte, 7 the packed normal, 8-10 the query point, 12-13 the
// visibility bits).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "internal.h"

namespace rq {
namespace gic {

constexpr uint32_t HEADER = 64, HOT_WORDS = 24, MIN_COLD = 24;
constexpr uint32_t H_MASK = 1, H_FLAGS = 2, H_MIN_EDGE = 5, H_GATHER_STRIDE = 21, H_COLD_WORDS = 23;
constexpr uint32_t F_ROTATE = 4;
constexpr uint32_t KEY_BITS = 0x7fffffffu;
constexpr uint32_t C_KEY_B = 0, C_DEF = 3, C_NORMAL = 7, C_POINT = 8, C_BITS = 12;
constexpr uint32_t DIRECTIONS = 64;

struct Cell {
    bool ok;
    float n[3], q[3], edge;
    uint32_t key_x, key_y, bits[2];
};

__device__ __forceinline__ float bf(uint32_t w) { return __uint_as_float(w); }

// unpack2x16snorm, then gi_unpack_normal's octahedral fold and normalize.
__device__ inline void unpack_normal(uint32_t w, float n[3]) {
    const float qx = fminf(fmaxf((float)(int16_t)(w & 0xffffu) / 32767.0f, -1.0f), 1.0f);
    const float qy = fminf(fmaxf((float)(int16_t)(w >> 16) / 32767.0f, -1.0f), 1.0f);
    float x = qx, y = qy, z = 1.0f - fabsf(qx) - fabsf(qy);
    const float t = fmaxf(-z, 0.0f);
    x = x + (x >= 0.0f ? -t : t);
    y = y + (y >= 0.0f ? -t : t);
    const float len = sqrtf(x * x + y * y + z * z);
    n[0] = x / len;
    n[1] = y / len;
    n[2] = z / len;
}

// Reads slot's hot and cold records; ok is false when any word lies past the table.
__device__ inline Cell load_cell(const uint32_t* table, uint64_t words, uint32_t slot) {
    Cell c;
    c.ok = false;
    if (words <= HEADER) return c;
    const uint64_t slots = (uint64_t)table[H_MASK] + 1;
    const uint64_t cold_words = max(MIN_COLD, table[H_COLD_WORDS]);
    if (slot >= slots) return c;
    const uint64_t hot = HEADER + (uint64_t)slot * HOT_WORDS;
    const uint64_t cold = HEADER + slots * HOT_WORDS + (uint64_t)slot * cold_words;
    if (cold + MIN_COLD > words || hot + HOT_WORDS > words) return c;
    c.key_x = table[hot] & KEY_BITS;
    c.key_y = table[cold + C_KEY_B];
    c.edge = bf(table[H_MIN_EDGE]) * exp2f((float)(table[cold + C_DEF] & 0xffu));
    unpack_normal(table[cold + C_NORMAL], c.n);
    for (int i = 0; i < 3; ++i) c.q[i] = bf(table[cold + C_POINT + i]);
    c.bits[0] = table[cold + C_BITS];
    c.bits[1] = table[cold + C_BITS + 1];
    c.ok = true;
    return c;
}

// gi_origin_offset: at least the path tracer's offset, 1e-4 + 4e-7 |q|max.
__device__ inline float origin_offset(const float q[3], float edge) {
    const float m = fmaxf(fmaxf(fabsf(q[0]), fabsf(q[1])), fabsf(q[2]));
    const float mm = fmaxf(m, 1e-30f);
    const float ulp = __uint_as_float(__float_as_uint(mm) + 1u) - mm;
    return fmaxf(fmaxf(1e-5f, 32.0f * ulp) + 1e-4f * edge, 1e-4f + 4e-7f * m);
}

// gi_direction through gi_basis (Duff et al. 2017): equal solid angles of the
// hemisphere about n, turned by the key's angle when the table asks for it.
__device__ inline void direction(const Cell& c, bool rotate, uint32_t k, float d[3]) {
    const float z = 1.0f - ((float)k + 0.5f) / (float)DIRECTIONS;
    const float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    const float rotation = rotate ? (float)(c.key_y >> 8) * (float)(6.283185307179586 / 16777216.0) : 0.0f;
    const float phi = (float)k * 2.399963229728653f + rotation;
    const float lx = r * cosf(phi), ly = r * sinf(phi), lz = z;
    const float* n = c.n;
    const float s = n[2] >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + n[2]);
    const float b = n[0] * n[1] * a;
    const float c0[3] = {1.0f + s * n[0] * n[0] * a, s * b, -s * n[0]};
    const float c1[3] = {b, s + n[1] * n[1] * a, -n[1]};
    for (int i = 0; i < 3; ++i) d[i] = c0[i] * lx + c1[i] * ly + n[i] * lz;
}

__device__ inline rq_ray cell_ray(const Cell& c, bool rotate, uint32_t k) {
    rq_ray r;
    const float off = origin_offset(c.q, c.edge);
    float d[3];
    direction(c, rotate, k, d);
    for (int i = 0; i < 3; ++i) {
        r.origin[i] = c.q[i] + c.n[i] * off;
        r.direction[i] = d[i];
    }
    r.tmin = 0.0f;
    r.tmax = 1e20f;
    return r;
}

__device__ inline rq_ray bad_ray() {
    rq_ray r;
    for (int i = 0; i < 3; ++i) {
        r.origin[i] = __builtin_nanf("");
        r.direction[i] = 0.0f;
    }
    r.tmin = 0.0f;
    r.tmax = 1e20f;
    return r;
}

__global__ __launch_bounds__(256) void item_rays(const uint32_t* table, uint64_t words, const uint32_t* items,
                                                 uint32_t count, rq_ray* rays) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    const uint32_t item = items[i], k = item & 127u;
    const Cell c = load_cell(table, words, item >> 7);
    const bool rotate = words > H_FLAGS && (table[H_FLAGS] & F_ROTATE) != 0;
    rays[i] = (c.ok && k < DIRECTIONS) ? cell_ray(c, rotate, k) : bad_ray();
}

__device__ inline uint32_t stride_of(const uint32_t* table, uint64_t words) {
    return words > H_GATHER_STRIDE ? max(table[H_GATHER_STRIDE], 1u) : 1u;
}

// gi_gather: directions phase + j * stride (phase = key.x & (stride - 1)) not
// already known visible.
__device__ inline uint64_t gather_mask(const Cell& c, uint32_t stride) {
    uint64_t m = 0;
    if (!c.ok) return 0;
    const uint64_t open = (uint64_t)c.bits[0] | ((uint64_t)c.bits[1] << 32);
    for (uint32_t k = c.key_x & (stride - 1); k < DIRECTIONS; k += stride) m |= 1ull << k;
    return m & ~open;
}

__global__ __launch_bounds__(256) void gather_counts(const uint32_t* table, uint64_t words, const uint32_t* slots,
                                                     uint32_t n, uint32_t* counts) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    counts[i] = __popcll(gather_mask(load_cell(table, words, slots[i]), stride_of(table, words)));
}

__global__ __launch_bounds__(256) void gather_write(const uint32_t* table, uint64_t words, const uint32_t* slots,
                                                    uint32_t n, const uint32_t* counts, const uint32_t* offsets,
                                                    rq_ray* rays, uint32_t* items, uint32_t max_rays,
                                                    uint32_t* total) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == n - 1) *total = offsets[i] + counts[i];
    if (i >= n) return;
    const Cell c = load_cell(table, words, slots[i]);
    uint64_t m = gather_mask(c, stride_of(table, words));
    const bool rotate = words > H_FLAGS && (table[H_FLAGS] & F_ROTATE) != 0;
    for (uint32_t at = offsets[i]; m; m &= m - 1, ++at) {
        if (at >= max_rays) break;
        const uint32_t k = (uint32_t)__builtin_ctzll(m);
        rays[at] = cell_ray(c, rotate, k);
        items[at] = (slots[i] << 7) | k;
    }
}

// gi_near_bin (shaders/gi_rays_software.wgsl:330-337): the octahedral cell of d
// among bins x bins, row-major (near_field::bin, crates/asset/src/near_field.rs:139-145).
__device__ inline uint32_t near_bin(const float d[3], uint32_t bins) {
    const float s = fabsf(d[0]) + fabsf(d[1]) + fabsf(d[2]);
    float px = d[0] / s, py = d[1] / s;
    if (!(d[2] >= 0.0f)) {
        const float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    const float n = (float)bins;
    const float cx = fminf(fmaxf(floorf((px * 0.5f + 0.5f) * n), 0.0f), n - 1.0f);
    const float cy = fminf(fmaxf(floorf((py * 0.5f + 0.5f) * n), 0.0f), n - 1.0f);
    return (uint32_t)cy * bins + (uint32_t)cx;
}

// gi_near_blocked (:338-354): every vertex with weight at the point blocks d.
__global__ __launch_bounds__(256) void near_blocked(const uint32_t* words, uint64_t n_words, uint32_t bins,
                                                    const rq_near_query* q, uint32_t count, uint32_t* out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    bool blocked = false;
    if (i < count && bins) {
        const rq_near_query a = q[i];
        const uint64_t rec = HEADER_WORDS + (uint64_t)a.inst * INSTANCE_WORDS;
        const uint32_t base = a.inst < words[0] && rec + 24 <= n_words ? words[rec + 23] : 0u;
        if (base) {
            const uint32_t per = bins * bins / 32u, b = near_bin(a.d, bins);
            const uint64_t at = (uint64_t)base + (a.back ? per : 0u) + b / 32u, stride = 2ull * per;
            const uint32_t bit = 1u << (b % 32u);
            const float w[3] = {1.0f - a.bary[0] - a.bary[1], a.bary[0], a.bary[1]};
            float s = 0.0f;
            for (int v = 0; v < 3; ++v) {
                const uint64_t k = at + (uint64_t)a.tri[v] * stride;
                if (k < n_words && (words[k] & bit)) s += w[v];
            }
            blocked = s >= 0.999f;
        }
    }
    const uint64_t m = __ballot(blocked);
    const uint32_t lane = threadIdx.x & 63u;
    if (i < count && (lane & 31u) == 0) {
        const uint32_t word = (uint32_t)(m >> (lane & 32u));
        out[i / 32u] = word;
    }
}

__global__ void zero_word(uint32_t* p) {
    if (threadIdx.x == 0) *p = 0u;
}

// Exclusive prefix sum of n words (rocPRIM); a size query when temp is null.
hipError_t exclusive_sum(void* temp, size_t& bytes, const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t s) {
    return rocprim::exclusive_scan(temp, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s);
}

}  // namespace gic
}  // namespace rq

using namespace rq;

namespace {
rq_status gic_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RQ_ERR_OUT_OF_MEMORY : RQ_ERR_DEVICE;
}
}  // namespace

extern "C" rq_status rq_gi_item_rays(const uint32_t* table, size_t table_words, const uint32_t* items, uint32_t count,
                                     rq_ray* rays, void* stream) {
    if (!table || (count && (!items || !rays))) {
        set_error("rq_gi_item_rays: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    if (!count) return RQ_OK;
    hipLaunchKernelGGL(gic::item_rays, dim3((count + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       table, (uint64_t)table_words, items, count, rays);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RQ_OK : gic_fail(e, "rq_gi_item_rays launch");
}

extern "C" rq_status rq_gi_rays(rq_scene* scene, rq_layout layout, const uint32_t* table, size_t table_words,
                                const uint32_t* items, uint32_t count, rq_ray* workspace, uint32_t* staged_bits,
                                void* stream) {
    if (count && (!workspace || !staged_bits)) {
        set_error("rq_gi_rays: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    rq_status s = rq_gi_item_rays(table, table_words, items, count, workspace, stream);
    if (s != RQ_OK || !count) return s;
    return rq_trace(scene, layout, workspace, count, RQ_MODE_ANY, nullptr, staged_bits, stream);
}

extern "C" rq_status rq_gi_gather_rays(const uint32_t* table, size_t table_words, const uint32_t* slots,
                                       uint32_t slot_count, rq_ray* rays, uint32_t* items, uint32_t max_rays,
                                       uint32_t* count, void* workspace, size_t workspace_bytes, void* stream) {
    if (!table || !count || (slot_count && (!slots || !workspace)) || (max_rays && (!rays || !items))) {
        set_error("rq_gi_gather_rays: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!slot_count) {
        hipLaunchKernelGGL(gic::zero_word, dim3(1), dim3(64), 0, s, count);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? RQ_OK : gic_fail(e, "rq_gi_gather_rays");
    }
    size_t scan_bytes = 0;
    uint32_t* counts = nullptr;
    hipError_t e = gic::exclusive_sum(nullptr, scan_bytes, counts, counts, slot_count, s);
    if (e != hipSuccess) return gic_fail(e, "rq_gi_gather_rays scan size");
    if (workspace_bytes < rq_gi_gather_workspace(slot_count)) {
        set_error("rq_gi_gather_rays: workspace smaller than rq_gi_gather_workspace(slot_count)");
        return RQ_ERR_INVALID_INPUT;
    }
    char* w = reinterpret_cast<char*>(workspace);
    counts = reinterpret_cast<uint32_t*>(w);
    uint32_t* offsets = counts + slot_count;
    void* temp = w + ((8ull * slot_count + 255) & ~255ull);
    const dim3 grid((slot_count + 255) / 256);
    hipLaunchKernelGGL(gic::gather_counts, grid, dim3(256), 0, s, table, (uint64_t)table_words, slots, slot_count,
                       counts);
    e = gic::exclusive_sum(temp, scan_bytes, counts, offsets, slot_count, s);
    if (e != hipSuccess) return gic_fail(e, "rq_gi_gather_rays scan");
    hipLaunchKernelGGL(gic::gather_write, grid, dim3(256), 0, s, table, (uint64_t)table_words, slots, slot_count,
                       counts, offsets, rays, items, max_rays, count);
    e = hipGetLastError();
    return e == hipSuccess ? RQ_OK : gic_fail(e, "rq_gi_gather_rays launch");
}

extern "C" size_t rq_gi_gather_workspace(uint32_t slot_count) {
    size_t scan_bytes = 0;
    uint32_t* p = nullptr;
    // the size query only reads its arguments; it cannot fail for a valid count
    (void)gic::exclusive_sum(nullptr, scan_bytes, p, p, std::max<uint32_t>(slot_count, 1), (hipStream_t)0);
    return ((8ull * slot_count + 255) & ~255ull) + scan_bytes + 1024;
}

extern "C" rq_status rq_near_blocked(const rq_scene* scene, const rq_near_query* queries, uint32_t count,
                                     uint32_t* blocked_bits, void* stream) {
    if (!scene || (count && (!queries || !blocked_bits))) {
        set_error("rq_near_blocked: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    if (scene->kind != RQ_SCENE_GI || !scene->dev[0].ready) {
        set_error("rq_near_blocked: needs a GI scene with the REF layout uploaded");
        return RQ_ERR_INVALID_INPUT;
    }
    if (!count) return RQ_OK;
    hipLaunchKernelGGL(gic::near_blocked, dim3((count + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const uint32_t*>(scene->dev[0].base),
                       (uint64_t)scene_words(scene), scene->near_bins, queries, count, blocked_bits);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RQ_OK : gic_fail(e, "rq_near_blocked launch");
}
