// Refit of one mesh's BVH4 on the GPU, in place in a resident scene's REF
// words: near_field::refit_bvh (crates/asset/src/near_field.rs:445-479) --
// same topology, leaves and indices; every node's child boxes recomputed from
// the new positions (a leaf: the bounds of its triangles' corners; an inner
// child: the union of both decodes of its own, already refit, children) and
// quantized as quantized_node (:481-525) quantizes them, fused and unfused
// decodes both enclosing. Nodes are in post-order (children first), so the
// nodes of one height only read finished ones: one launch per height.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "bvh_device.h"
#include "internal.h"

namespace rq {
namespace refitk {

using bvhdev::bf;
using bvhdev::fb;

struct Args {
    uint32_t* words;
    uint32_t nodes, leaves, positions, indices;  // word offsets of the mesh's arrays
    const uint32_t* order;                       // node ids grouped by height
    uint32_t begin, end;                         // this launch's slice of `order`
    const float* boxes;                          // the mesh's leaf boxes
};

__global__ __launch_bounds__(256) void refit_level(Args a) {
    const uint32_t k = a.begin + blockIdx.x * 256u + threadIdx.x;
    if (k >= a.end) return;
    bvhdev::refit_node(a.words, a.nodes, a.leaves, a.positions, a.indices, a.order[k], a.boxes);
}

// Every leaf's box of one mesh into boxes (6 floats a leaf: lo xyz, hi xyz),
// so the lowest nodes read one box per leaf child, not its triangles.
__global__ __launch_bounds__(256) void leaf_boxes(const uint32_t* words, uint32_t leaves, uint32_t positions,
                                                  uint32_t indices, uint32_t count, float* boxes) {
    const uint32_t l = blockIdx.x * 256u + threadIdx.x;
    if (l >= count) return;
    float lo[3], hi[3];
    bvhdev::leaf_box(words, leaves, positions, indices, l, lo, hi);
    for (int x = 0; x < 3; ++x) {
        boxes[6ull * l + x] = lo[x];
        boxes[6ull * l + 3 + x] = hi[x];
    }
}

constexpr uint32_t BOUNDS_BLOCKS = 512;

// Bounds of the new positions and whether all are finite, as seven words:
// lo xyz and hi xyz as total-order keys (atomic min / max), then a non-finite
// flag. Min and max of finite floats do not depend on the order they are taken
// in but for the sign of a zero, which no instance box word depends on.
__global__ __launch_bounds__(64) void bounds_init(int32_t* out) {
    const uint32_t t = threadIdx.x;
    if (t < 7) out[t] = t < 3 ? bvhdev::okey(__builtin_inff()) : t < 6 ? bvhdev::okey(-__builtin_inff()) : 0;
}

__global__ __launch_bounds__(256) void bounds_reduce(const float* p, uint32_t vertices, int32_t* out) {
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    uint32_t bad = 0;
    for (uint32_t v = blockIdx.x * 256u + threadIdx.x; v < vertices; v += gridDim.x * 256u)
        for (int x = 0; x < 3; ++x) {
            const float f = p[3ull * v + x];
            bad |= !isfinite(f);
            lo[x] = fminf(lo[x], f);
            hi[x] = fmaxf(hi[x], f);
        }
    __shared__ int32_t s[7];
    if (threadIdx.x < 7)
        s[threadIdx.x] = threadIdx.x < 3 ? bvhdev::okey(__builtin_inff())
                                         : threadIdx.x < 6 ? bvhdev::okey(-__builtin_inff()) : 0;
    __syncthreads();
    for (int x = 0; x < 3; ++x) {
        atomicMin(&s[x], bvhdev::okey(lo[x]));
        atomicMax(&s[3 + x], bvhdev::okey(hi[x]));
    }
    if (bad) atomicOr(&s[6], 1);
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&out[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&out[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x == 6 && s[6]) atomicOr(&out[6], 1);
}

__global__ __launch_bounds__(256) void put_positions(uint32_t* words, uint32_t at, const float* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) words[at + i] = fb(p[i]);
}

// PLAIN refresh of one refit mesh: each 32-word node's child boxes are the
// unfused decodes (origin + step * q) of its reference node, tags untouched;
// each 36-byte triangle record is vertex a and the edges b - a, c - a.
__global__ __launch_bounds__(256) void plain_nodes(const uint32_t* words, uint32_t nodes, uint32_t count,
                                                   uint32_t* out) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= count) return;
    const uint32_t* w = words + nodes + 16ull * n;
    uint32_t* o = out + 32ull * n;
    for (int slot = 0; slot < 4; ++slot) {
        if (w[6 + slot] == EMPTY) continue;
        for (int x = 0; x < 3; ++x) {
            const float org = bf(w[x]), s = bf(w[3 + x]);
            const float ql = (float)((w[10 + x] >> (8 * slot)) & 0xffu);
            const float qh = (float)((w[13 + x] >> (8 * slot)) & 0xffu);
            o[x * 4 + slot] = fb(org + s * ql);
            o[12 + x * 4 + slot] = fb(org + s * qh);
        }
    }
}

__global__ __launch_bounds__(256) void plain_tris(const uint32_t* words, uint32_t indices, uint32_t positions,
                                                  uint32_t count, float* out, float* verts) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= count) return;
    float p[3][3];
    for (int v = 0; v < 3; ++v) {
        const uint32_t vi = words[indices + 3ull * k + v];
        for (int x = 0; x < 3; ++x) p[v][x] = bf(words[positions + 3ull * vi + x]);
    }
    float* o = out + (size_t)TRI_FLOATS * k;
    for (int x = 0; x < 3; ++x) o[x] = p[0][x];
    for (int e = 1; e < 3; ++e)
        for (int x = 0; x < 3; ++x) o[3 * e + x] = p[e][x] - p[0][x];
    float* q = verts + (size_t)TRI_FLOATS * k;
    for (int v = 0; v < 3; ++v)
        for (int x = 0; x < 3; ++x) q[3 * v + x] = p[v][x];
}

// Instance record words 0-17 (rows and world box) into the PLAIN records.
__global__ __launch_bounds__(256) void plain_records(const uint32_t* words, uint32_t count, uint32_t* inst) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= count * 18u) return;
    const uint32_t i = t / 18u, k = t % 18u;
    inst[(size_t)i * INSTANCE_WORDS + k] = words[HEADER_WORDS + (size_t)i * INSTANCE_WORDS + k];
}

}  // namespace refitk
}  // namespace rq

using namespace rq;

namespace {
rq_status fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RQ_ERR_OUT_OF_MEMORY : RQ_ERR_DEVICE;
}
#define RF_TRY(x, what)                       \
    do {                                      \
        hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return fail(e_, what); \
    } while (0)

// Node ids of one mesh grouped by height (leaf-only nodes first), uploaded once.
rq_status refit_plan(rq_scene* scene, const MeshView& m, RefitPlan*& out) {
    for (auto& p : scene->refit_plans)
        if (p.nodes == m.nodes) {
            out = &p;
            return RQ_OK;
        }
    const uint32_t* w = scene->words.data();
    std::vector<uint32_t> h(m.node_count, 0);
    uint32_t top = 0;
    for (uint32_t n = 0; n < m.node_count; ++n) {  // post-order: children first
        uint32_t hh = 0;
        for (int c = 0; c < 4; ++c) {
            uint32_t ch = w[m.nodes + 16ull * n + 6 + c];
            if (ch == EMPTY) continue;
            if (ch & LEAF_BIT) {
                if ((ch & ~LEAF_BIT) >= m.leaf_count) ch = n;  // out of range: refused below
                else continue;
            }
            if (ch >= n) {
                set_error("rq_scene_refit_mesh: mesh nodes are not in post-order");
                return RQ_ERR_INVALID_INPUT;
            }
            hh = std::max(hh, h[ch] + 1);
        }
        h[n] = hh;
        top = std::max(top, hh);
    }
    // Every leaf span and vertex index stays inside the mesh (the kernel reads them unchecked).
    for (uint32_t l = 0; l < m.leaf_count; ++l) {
        const uint64_t off = w[m.leaves + 2ull * l], cnt = w[m.leaves + 2ull * l + 1];
        if (off + cnt > m.index_count) {
            set_error("rq_scene_refit_mesh: leaf past the mesh's indices");
            return RQ_ERR_INVALID_INPUT;
        }
    }
    for (uint32_t i = 0; i < m.index_count; ++i)
        if (w[m.indices + i] >= m.vertex_count) {
            set_error("rq_scene_refit_mesh: index past the mesh's vertices");
            return RQ_ERR_INVALID_INPUT;
        }
    RefitPlan plan;
    plan.nodes = m.nodes;
    plan.level_start.assign(top + 2, 0);
    for (uint32_t n = 0; n < m.node_count; ++n) plan.level_start[h[n] + 1]++;
    for (uint32_t l = 0; l <= top; ++l) plan.level_start[l + 1] += plan.level_start[l];
    std::vector<uint32_t> order(m.node_count), fill(plan.level_start.begin(), plan.level_start.end() - 1);
    for (uint32_t n = 0; n < m.node_count; ++n) order[fill[h[n]]++] = n;
    RF_TRY(hipMalloc(&plan.boxes, std::max<size_t>(m.leaf_count, 1) * 6 * sizeof(float)), "hipMalloc refit plan");
    if (!order.empty()) {
        RF_TRY(hipMalloc(&plan.order, order.size() * 4), "hipMalloc refit plan");
        RF_TRY(hipMemcpy(plan.order, order.data(), order.size() * 4, hipMemcpyHostToDevice), "hipMemcpy refit plan");
    }
    scene->refit_plans.push_back(plan);
    out = &scene->refit_plans.back();
    return RQ_OK;
}
}  // namespace

namespace rq {
// Grows the host words to their full length (zeros; the new words are stale).
// The device copy of a scene's words: REF for GI scenes, RS for reflection scenes.
static int word_layout(const rq_scene* scene) {
    return scene->kind == RQ_SCENE_REFLECTION ? RQ_LAYOUT_RS_BVH4Q8 : RQ_LAYOUT_REF_BVH4Q8;
}

static void grow_host_words(const rq_scene* scene) {
    if (scene->total_words > scene->words.size())
        const_cast<rq_scene*>(scene)->words.resize(scene->total_words, 0u);
}

rq_status sync_host_words(const rq_scene* scene) {
    if (!scene || scene->stale.empty()) return RQ_OK;
    grow_host_words(scene);
    const uint32_t* words = reinterpret_cast<const uint32_t*>(scene->dev[word_layout(scene)].base);
    uint32_t* host = const_cast<uint32_t*>(scene->words.data());
    for (const auto& r : scene->stale)
        RF_TRY(hipMemcpy(host + r.first, words + r.first, r.second * 4, hipMemcpyDeviceToHost), "fetch refit words");
    scene->stale.clear();
    return RQ_OK;
}

// Fetches only the stale words inside [lo, hi), keeping the rest stale.
rq_status sync_host_range(const rq_scene* scene, uint64_t lo, uint64_t hi) {
    if (!scene || scene->stale.empty() || lo >= hi) return RQ_OK;
    grow_host_words(scene);
    const uint32_t* words = reinterpret_cast<const uint32_t*>(scene->dev[word_layout(scene)].base);
    uint32_t* host = const_cast<uint32_t*>(scene->words.data());
    std::vector<std::pair<uint64_t, uint64_t>> keep;
    for (const auto& r : scene->stale) {
        const uint64_t a = r.first, b = r.first + r.second;
        const uint64_t x = std::max(a, lo), y = std::min(b, hi);
        if (x >= y) {
            keep.push_back(r);
            continue;
        }
        RF_TRY(hipMemcpy(host + x, words + x, (y - x) * 4, hipMemcpyDeviceToHost), "fetch words");
        if (a < x) keep.emplace_back(a, x - a);
        if (y < b) keep.emplace_back(y, b - y);
    }
    scene->stale.swap(keep);
    return RQ_OK;
}
}  // namespace rq

namespace {
void mark_stale(rq_scene* scene, uint64_t at, uint64_t count) {
    for (const auto& r : scene->stale)
        if (r.first == at && r.second == count) return;
    scene->stale.emplace_back(at, count);
}
}  // namespace

extern "C" rq_status rq_scene_refit_mesh(rq_scene* scene, uint32_t mesh, const float* positions,
                                         uint32_t vertex_count, void* stream) {
    if (!scene || !positions || scene->kind != RQ_SCENE_GI) {
        set_error("rq_scene_refit_mesh: needs a GI scene and device positions");
        return RQ_ERR_INVALID_INPUT;
    }
    if (mesh >= scene->mesh_nodes.size()) {
        set_error("rq_scene_refit_mesh: no such mesh (refits need a scene made by rq_scene_build)");
        return RQ_ERR_INVALID_INPUT;
    }
    if (!scene->dev[0].ready) {
        set_error("rq_scene_refit_mesh: the REF layout is not uploaded");
        return RQ_ERR_INVALID_INPUT;
    }
    int dev = 0;
    RF_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev != scene->device) {
        set_error("rq_scene_refit_mesh: scene lives on another device");
        return RQ_ERR_INVALID_INPUT;
    }
    std::vector<uint32_t>& w = scene->words;  // header, records, tags and indices are current on the host
    const uint32_t count = w[0];
    uint32_t record = count;
    for (uint32_t i = 0; i < count && record == count; ++i)
        if (w[HEADER_WORDS + (size_t)i * INSTANCE_WORDS + 18] == scene->mesh_nodes[mesh]) record = i;
    if (record == count) return RQ_OK;  // no instance draws this mesh: nothing packed to refit
    // The plan reads the child tags, leaves and indices, which a refit never
    // changes: fetch them once if a GPU build left them on the device only.
    const size_t rec = HEADER_WORDS + (size_t)record * INSTANCE_WORDS;
    bool planned = false;
    for (const auto& p : scene->refit_plans) planned = planned || p.nodes == w[rec + 18];
    if (!planned && w[rec + 18] <= w[rec + 20] && w[rec + 20] <= scene_words(scene)) {
        rq_status st = sync_host_range(scene, w[rec + 18], w[rec + 20]);  // nodes and leaves
        if (st != RQ_OK) return st;
    }
    MeshView m;
    std::string err;
    if (!mesh_view(w, record, m, err)) {
        set_error("rq_scene_refit_mesh: " + err);
        return RQ_ERR_INVALID_INPUT;
    }
    if (!planned) {
        rq_status st = sync_host_range(scene, m.indices, (uint64_t)m.indices + m.index_count);
        if (st != RQ_OK) return st;
    }
    if (vertex_count != m.vertex_count || vertex_count == 0) {
        set_error("BVH refit needs finite positions of the same vertices");
        return RQ_ERR_INVALID_INPUT;
    }
    DeviceLayout& d = scene->dev[1];
    size_t k = 0;
    if (d.ready) {
        while (k + 3 <= d.mesh_base.size() && d.mesh_base[k] != m.nodes) k += 3;
        if (k + 3 > d.mesh_base.size()) {
            set_error("rq_scene_refit_mesh: the PLAIN layout does not hold this mesh");
            return RQ_ERR_INVALID_INPUT;
        }
    }
    RefitPlan* plan = nullptr;
    rq_status st = refit_plan(scene, m, plan);
    if (st != RQ_OK) return st;
    if (!scene->refit_scratch) RF_TRY(hipMalloc(&scene->refit_scratch, 8 * sizeof(int32_t)), "hipMalloc refit");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint32_t* words = reinterpret_cast<uint32_t*>(scene->dev[0].base);
    int32_t* bnd = reinterpret_cast<int32_t*>(scene->refit_scratch);
    // 1. bounds of the new positions, and whether they are all finite
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((vertex_count + 255) / 256, refitk::BOUNDS_BLOCKS));
    hipLaunchKernelGGL(refitk::bounds_init, dim3(1), dim3(64), 0, s, bnd);
    hipLaunchKernelGGL(refitk::bounds_reduce, dim3(blocks), dim3(256), 0, s, positions, vertex_count, bnd);
    int32_t keys[7];
    RF_TRY(hipMemcpyAsync(keys, bnd, 7 * sizeof(int32_t), hipMemcpyDeviceToHost, s), "refit bounds");
    RF_TRY(hipStreamSynchronize(s), "refit bounds");
    float b[6];
    for (int k = 0; k < 6; ++k) b[k] = bits_f((uint32_t)(keys[k] ^ (int32_t)(((uint32_t)(keys[k] >> 31)) >> 1)));
    if (keys[6]) {
        set_error("BVH refit needs finite positions of the same vertices");
        return RQ_ERR_INVALID_INPUT;
    }
    // 2. positions into the words, then the nodes bottom-up, one launch per height
    const uint32_t nf = 3u * vertex_count;
    hipLaunchKernelGGL(refitk::put_positions, dim3(std::min<uint32_t>((nf + 255) / 256, 2048)), dim3(256), 0, s,
                       words, m.positions, positions, nf);
    refitk::Args a;
    a.words = words;
    a.nodes = m.nodes;
    a.leaves = m.leaves;
    a.positions = m.positions;
    a.indices = m.indices;
    a.order = reinterpret_cast<const uint32_t*>(plan->order);
    a.boxes = reinterpret_cast<const float*>(plan->boxes);
    hipLaunchKernelGGL(refitk::leaf_boxes, dim3((m.leaf_count + 255) / 256), dim3(256), 0, s, words, m.leaves,
                       m.positions, m.indices, m.leaf_count, reinterpret_cast<float*>(plan->boxes));
    for (size_t l = 0; l + 1 < plan->level_start.size(); ++l) {
        a.begin = plan->level_start[l];
        a.end = plan->level_start[l + 1];
        if (a.end > a.begin)
            hipLaunchKernelGGL(refitk::refit_level, dim3((a.end - a.begin + 255) / 256), dim3(256), 0, s, a);
    }
    RF_TRY(hipGetLastError(), "refit launch");
    mark_stale(scene, m.positions, nf);
    mark_stale(scene, m.nodes, (uint64_t)m.node_count * NODE_WORDS);
    // 3. instance records over the new mesh bounds, then the top level placed
    //    over the previous one (refit while only boxes moved), as a repack would
    const size_t base = HEADER_WORDS + (size_t)count * INSTANCE_WORDS, reserve = top_reserve(count);
    std::vector<uint32_t> head(w.begin(), w.begin() + base);
    for (uint32_t i = 0; i < count; ++i) {
        const size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        if (head[r + 18] == m.nodes) instance_box(&scene->models[16ull * i], b, b + 3, &head[r + 12]);
    }
    top_place(head, HEADER_WORDS, INSTANCE_WORDS, count, w.data(), base + reserve, scene->top_force);
    std::copy(head.begin(), head.end(), w.begin());
    RF_TRY(hipMemcpyAsync(words, w.data(), head.size() * 4, hipMemcpyHostToDevice, s), "refit records");
    scene_bounds(scene);
    // 4. PLAIN follows in place (same nodes, same triangle order); CW8, whose
    //    collapse depends on the boxes, is dropped and derived again on upload
    if (d.ready) {
        char* pb = reinterpret_cast<char*>(d.base);
        uint32_t* pnodes = reinterpret_cast<uint32_t*>(pb + d.off[2]) + 32ull * d.mesh_base[k + 1];
        float* ptris = reinterpret_cast<float*>(pb + d.off[3]) + (size_t)TRI_FLOATS * d.mesh_base[k + 2];
        float* pverts = reinterpret_cast<float*>(pb + d.off[5]) + (size_t)TRI_FLOATS * d.mesh_base[k + 2];
        const uint32_t tris = m.index_count / 3;
        hipLaunchKernelGGL(refitk::plain_nodes, dim3((m.node_count + 255) / 256), dim3(256), 0, s, words, m.nodes,
                           m.node_count, pnodes);
        if (tris)
            hipLaunchKernelGGL(refitk::plain_tris, dim3((tris + 255) / 256), dim3(256), 0, s, words, m.indices,
                               m.positions, tris, ptris, pverts);
        hipLaunchKernelGGL(refitk::plain_records, dim3((count * 18 + 255) / 256), dim3(256), 0, s, words, count,
                           reinterpret_cast<uint32_t*>(pb + d.off[0]));
        if (d.count[1] && d.count[3])
            RF_TRY(hipMemcpyAsync(pb + d.off[1], words + w[1], (size_t)d.count[3] * 4, hipMemcpyDeviceToDevice, s),
                   "refit plain top");
        RF_TRY(hipGetLastError(), "refit plain launch");
    }
    scene->plain = HostLayout();  // host images are derived again when asked for
    scene->cw8 = HostLayout();
    if (scene->dev[2].base) (void)hipFree(scene->dev[2].base);
    scene->dev[2] = DeviceLayout();
    RF_TRY(hipStreamSynchronize(s), "refit");  // `head` is copied from before it goes
    return RQ_OK;
}
