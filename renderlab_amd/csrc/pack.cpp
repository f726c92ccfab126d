// The packed gi_geometry words: per-mesh records, instance records, header and
// top level, exactly as shaders/gi_rays_software.wgsl reads them.
//
// Follows src/gi_bvh.rs:33-274 (build, Blas::with_primitives, map_primitives,
// bounds, layout, mesh_words, instance_words) and the per-record extra words
// of src/gi_cache.rs:4388-4479 (identity, flags, primitive base, placement,
// header words 0 and 4, then instance_bvh::place).
#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "internal.h"

namespace rq {

// src/gi_bvh.rs:108-136: each BVH triangle's source primitive, through a
// head table and next links so duplicate triplets get distinct ids.
bool map_primitives(const std::vector<uint32_t>& bvh_indices, const uint32_t* source, uint32_t n,
                    std::vector<uint32_t>& out) {
    out.clear();
    uint32_t tri_count = n / 3, bvh_tris = (uint32_t)(bvh_indices.size() / 3);
    if (tri_count == 0 || bvh_tris == 0) return true;
    struct Key {
        uint32_t a, b, c;
        bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
    };
    struct Hash {
        size_t operator()(const Key& k) const {
            uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
            h ^= (k.b + 0x632BE59BD9B4E019ull) + (h << 6) + (h >> 2);
            h ^= (k.c + 0x85EBCA77C2B2AE63ull) + (h << 6) + (h >> 2);
            return (size_t)h;
        }
    };
    std::unordered_map<Key, uint32_t, Hash> head;
    head.reserve(tri_count * 2);
    std::vector<uint32_t> next(tri_count, 0xffffffffu);
    for (uint32_t id = tri_count; id-- > 0;) {
        Key k{source[3 * id], source[3 * id + 1], source[3 * id + 2]};
        auto it = head.find(k);
        if (it != head.end()) {
            next[id] = it->second;
            it->second = id;
        } else {
            head.emplace(k, id);
        }
    }
    out.reserve(bvh_tris);
    for (uint32_t t = 0; t < bvh_tris; ++t) {
        Key k{bvh_indices[3 * t], bvh_indices[3 * t + 1], bvh_indices[3 * t + 2]};
        auto it = head.find(k);
        if (it == head.end() || it->second == 0xffffffffu) {
            out.clear();
            return true;
        }
        uint32_t id = it->second;
        it->second = next[id];
        out.push_back(id);
    }
    return true;
}

// Inverse of a column-major 4x4 f32 matrix by cofactors; false when singular
// or not finite. (The reference calls glam's Mat4::inverse, whose exact f32
// bits are not restated here.)
bool mat4_inverse(const float m[16], float out[16]) {
    float inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    float r = 1.0f / det;
    for (int i = 0; i < 16; ++i) {
        out[i] = inv[i] * r;
        if (!std::isfinite(out[i])) return false;
    }
    return true;
}

// One instance record's mesh facts, recovered from record words 18-21 and
// the last leaf's span (leaves run contiguously over the index array).
bool mesh_view(const std::vector<uint32_t>& words, uint32_t record, MeshView& m, std::string& err) {
    size_t r = HEADER_WORDS + (size_t)record * INSTANCE_WORDS;
    if (r + INSTANCE_WORDS > words.size()) {
        err = "instance record past the end of the buffer";
        return false;
    }
    m.nodes = words[r + 18];
    m.leaves = words[r + 19];
    m.positions = words[r + 20];
    m.indices = words[r + 21];
    if (!(m.nodes <= m.leaves && m.leaves <= m.positions && m.positions <= m.indices &&
          m.indices <= words.size()) ||
        (m.leaves - m.nodes) % NODE_WORDS || (m.positions - m.leaves) % 2 || (m.indices - m.positions) % 3) {
        err = "instance record's mesh offsets are not a mesh";
        return false;
    }
    m.node_count = (m.leaves - m.nodes) / NODE_WORDS;
    m.leaf_count = (m.positions - m.leaves) / 2;
    m.vertex_count = (m.indices - m.positions) / 3;
    if (m.leaf_count == 0) {
        err = "mesh has no leaves";
        return false;
    }
    size_t last = m.leaves + 2ull * (m.leaf_count - 1);
    m.index_count = words[last] + words[last + 1];
    if ((size_t)m.indices + m.index_count > words.size()) {
        err = "mesh indices past the end of the buffer";
        return false;
    }
    return true;
}

void scene_bounds(rq_scene* scene) {
    const auto& w = scene->words;
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    uint32_t n = w.size() > 0 ? w[0] : 0;
    const bool rs = scene->kind == RQ_SCENE_REFLECTION;
    for (uint32_t i = 0; i < n; ++i) {
        size_t r = rs ? RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS : HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        if (r + 18 > w.size()) break;
        for (int a = 0; a < 3; ++a) {
            lo[a] = mn(lo[a], bits_f(w[r + 12 + a]));
            hi[a] = mx(hi[a], bits_f(w[r + 15 + a]));
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (!(lo[a] <= hi[a]) || !std::isfinite(lo[a]) || !std::isfinite(hi[a])) {
            lo[a] = 0.0f;
            hi[a] = 1.0f;
        }
        scene->bounds[a] = lo[a];
        scene->bounds[3 + a] = hi[a];
    }
}

void instance_box(const float model[16], const float blo[3], const float bhi[3], uint32_t out[6]) {
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    const float* M = model;
    for (int c = 0; c < 8; ++c) {
        float p[3];
        for (int a = 0; a < 3; ++a) p[a] = (c & (1 << a)) ? bhi[a] : blo[a];
        for (int a = 0; a < 3; ++a) {
            // glam transform_point3: ((x_axis*x + y_axis*y) + z_axis*z) + w_axis
            float wv = M[a] * p[0] + M[4 + a] * p[1];
            wv = wv + M[8 + a] * p[2];
            wv = wv + M[12 + a];
            lo[a] = mn(lo[a], wv);
            hi[a] = mx(hi[a], wv);
        }
    }
    for (int a = 0; a < 3; ++a) {
        float pad = 1e-5f * (1.0f + mx(std::fabs(lo[a]), std::fabs(hi[a])));
        out[a] = f_bits(lo[a] - pad);
        out[3 + a] = f_bits(hi[a] + pad);
    }
}

// src/gi_bvh.rs:151-274 (layout, mesh_words, instance_words) + src/gi_cache.rs:4388-4479:
// the packed words of BLASes already built, their instances and the top level.
rq_status pack_scene(std::vector<Blas>& blas, const rq_instance* instances, uint32_t instance_count,
                     int top_level, uint32_t near_bins, std::chrono::steady_clock::time_point began,
                     rq_scene** out) {
    const uint32_t mesh_count = (uint32_t)blas.size();
    // layout: each mesh's arrays, then each mesh's near-field mask blocks
    uint64_t at = HEADER_WORDS + (uint64_t)instance_count * INSTANCE_WORDS + top_reserve(instance_count);
    std::vector<uint32_t> offsets(4ull * mesh_count);
    for (uint32_t m = 0; m < mesh_count; ++m) {
        const Bvh& b = blas[m].bvh;
        uint64_t nodes = at, leaves = nodes + b.nodes.size() * NODE_WORDS;
        uint64_t positions = leaves + b.leaf_offset.size() * 2, indices = positions + b.positions.size();
        if (indices + b.indices.size() + blas[m].primitives.size() >= 0xffffffffull) {
            set_error("packed geometry exceeds 2^32 words");
            return RQ_ERR_INVALID_INPUT;
        }
        offsets[4 * m] = (uint32_t)nodes;
        offsets[4 * m + 1] = (uint32_t)leaves;
        offsets[4 * m + 2] = (uint32_t)positions;
        offsets[4 * m + 3] = (uint32_t)indices;
        at += blas[m].words();
    }
    std::vector<std::vector<uint32_t>> near_base(mesh_count);
    for (uint32_t m = 0; m < mesh_count; ++m)
        for (const NearMasks& n : blas[m].near) {
            near_base[m].push_back((uint32_t)at);
            at += n.masks.size();
        }
    if (at >= 0xffffffffull) {
        set_error("packed geometry exceeds 2^32 words");
        return RQ_ERR_INVALID_INPUT;
    }
    auto scene = new rq_scene();
    std::vector<uint32_t>& words = scene->words;
    words.assign(at, 0u);
    // src/gi_bvh.rs:182-226 (mesh_words)
    for (uint32_t m = 0; m < mesh_count; ++m) {
        const Bvh& b = blas[m].bvh;
        uint64_t w = offsets[4 * m];
        for (const Node& n : b.nodes) {
            std::memcpy(&words[w], &n, sizeof(Node));
            w += NODE_WORDS;
        }
        for (size_t l = 0; l < b.leaf_offset.size(); ++l) {
            words[w++] = b.leaf_offset[l];
            words[w++] = b.leaf_count[l];
        }
        for (float p : b.positions) words[w++] = f_bits(p);
        for (uint32_t i : b.indices) words[w++] = i;
        for (uint32_t p : blas[m].primitives) words[w++] = p;
        for (size_t k = 0; k < blas[m].near.size(); ++k)
            std::copy(blas[m].near[k].masks.begin(), blas[m].near[k].masks.end(), words.begin() + near_base[m][k]);
    }
    // src/gi_bvh.rs:230-274 (instance_words) + src/gi_cache.rs:4388-4479.
    std::vector<uint32_t> head(HEADER_WORDS + (size_t)instance_count * INSTANCE_WORDS, 0u);
    head[0] = instance_count;
    for (uint32_t i = 0; i < instance_count; ++i) {
        const rq_instance& ins = instances[i];
        const Blas& b = blas[ins.mesh];
        float inv[16];
        if (!mat4_inverse(ins.model, inv)) {
            delete scene;
            set_error("indirect geometry transform is singular");
            return RQ_ERR_INVALID_INPUT;
        }
        size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 4; ++c) head[r + row * 4 + c] = f_bits(inv[c * 4 + row]);
        instance_box(ins.model, b.lo, b.hi, &head[r + 12]);
        for (int k = 0; k < 4; ++k) head[r + 18 + k] = offsets[4 * ins.mesh + k];
        head[r + 22] = b.bvh.root;
        // src/gi_bvh.rs:97-104 (near_base): the masks cooked under exactly this model's bits
        uint32_t base = 0;
        for (size_t k = 0; k < b.near.size() && !base; ++k)
            if (std::memcmp(b.near[k].model, ins.model, sizeof ins.model) == 0) base = near_base[ins.mesh][k];
        head[r + 23] = base;
        uint32_t prim_base =
            b.primitives.empty() ? 0u : offsets[4 * ins.mesh + 3] + (uint32_t)b.bvh.indices.size();
        head[r + 24] = ins.identity;
        head[r + 27] = ins.flags & 1u;
        head[r + 28] = prim_base;
        head[r + 29] = ins.placement;
    }
    head[4] = instance_count;
    top_place(head, HEADER_WORDS, INSTANCE_WORDS, instance_count, nullptr, 0, top_level);
    std::copy(head.begin(), head.end(), words.begin());
    // stats
    uint64_t tris = 0, uniq = 0, nodes = 0;
    uint32_t depth = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        uniq += blas[m].bvh.indices.size() / 3;
        nodes += blas[m].bvh.nodes.size();
        depth = std::max(depth, bvh_depth(blas[m].bvh, blas[m].bvh.root));
    }
    for (uint32_t i = 0; i < instance_count; ++i) tris += blas[instances[i].mesh].bvh.indices.size() / 3;
    scene->stats[0] = instance_count;
    scene->stats[1] = mesh_count;
    scene->stats[2] = tris;
    scene->stats[3] = uniq;
    scene->stats[4] = nodes;
    scene->stats[5] = words[2];
    scene->stats[6] = depth;
    scene->models.resize(16ull * instance_count);
    for (uint32_t i = 0; i < instance_count; ++i) std::copy(instances[i].model, instances[i].model + 16, &scene->models[16ull * i]);
    scene->mesh_nodes.resize(mesh_count);
    for (uint32_t m = 0; m < mesh_count; ++m) scene->mesh_nodes[m] = offsets[4 * m];
    scene->top_force = top_level;
    scene->near_bins = near_bins;
    scene_bounds(scene);
    scene->stats[7] = (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                          std::chrono::steady_clock::now() - began).count();
    *out = scene;
    return RQ_OK;
}

// Bounds over all positions (src/gi_bvh.rs:64-87).
void blas_bounds(Blas& b) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = __builtin_inff();
        b.hi[a] = -__builtin_inff();
    }
    for (size_t v = 0; v + 2 < b.bvh.positions.size(); v += 3)
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = mn(b.lo[a], b.bvh.positions[v + a]);
            b.hi[a] = mx(b.hi[a], b.bvh.positions[v + a]);
        }
}

}  // namespace rq

using namespace rq;

extern "C" rq_status rq_scene_build(const rq_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                                    uint32_t instance_count, int top_level, rq_scene** out) {
    const auto began = std::chrono::steady_clock::now();
    if (!out || (!meshes && mesh_count) || (!instances && instance_count)) {
        set_error("rq_scene_build: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    for (uint32_t i = 0; i < instance_count; ++i)
        if (instances[i].mesh >= mesh_count) {
            set_error("rq_scene_build: instance names a missing mesh");
            return RQ_ERR_INVALID_INPUT;
        }
    // src/gi_bvh.rs:33-36, 64-87: SAH BLAS per mesh, bounds over all positions.
    std::vector<Blas> blas(mesh_count);
    std::vector<std::string> errs(mesh_count);
    std::vector<char> ok(mesh_count, 1);
    {
        uint32_t threads = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), 64));
        std::vector<std::thread> pool;
        std::atomic_uint next{0};
        for (uint32_t t = 0; t < std::min(threads, std::max(mesh_count, 1u)); ++t)
            pool.emplace_back([&]() {
                for (;;) {
                    uint32_t m = next.fetch_add(1);
                    if (m >= mesh_count) break;
                    const rq_mesh& src = meshes[m];
                    if (!build_bvh_sah(src.positions, src.vertex_count, src.indices, src.index_count, blas[m].bvh,
                                       errs[m])) {
                        ok[m] = 0;
                        continue;
                    }
                    blas_bounds(blas[m]);
                    if (src.keep_primitives)
                        map_primitives(blas[m].bvh.indices, src.indices, src.index_count, blas[m].primitives);
                }
            });
        for (auto& th : pool) th.join();
    }
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (!ok[m]) {
            set_error("indirect geometry is not a valid triangle list: " + errs[m]);
            return RQ_ERR_INVALID_INPUT;
        }
    return pack_scene(blas, instances, instance_count, top_level, 0, began, out);
}

namespace {

// Words from outside this library: every traversal reads node, leaf, index and
// position words by the ids it finds, and ends only if each tree is finite.
// Checks that each mesh's tree under an instance root and the top level reach
// no node twice, that every id stays inside its array, and that each reachable
// leaf's indices stay inside the buffer and name vertices of the mesh.
bool check_mesh_tree(const std::vector<uint32_t>& w, const MeshView& m, uint32_t root, std::string& err) {
    std::vector<uint8_t> seen(m.node_count, 0);
    std::vector<uint32_t> todo{root};
    while (!todo.empty()) {
        const uint32_t tag = todo.back();
        todo.pop_back();
        if (tag == EMPTY) continue;
        if (tag & LEAF_BIT) {
            const uint32_t l = tag & ~LEAF_BIT;
            if (l >= m.leaf_count) {
                err = "a leaf id past the mesh's leaves";
                return false;
            }
            const uint64_t first = w[m.leaves + 2ull * l], count = w[m.leaves + 2ull * l + 1];
            if (first % 3u || count % 3u) {
                // gi_rays_software.wgsl:127 reads indices first + k; every builder writes whole
                // triangles, and the device layouts step by triangles
                err = "a leaf's index offset or count is not a multiple of 3";
                return false;
            }
            if (m.indices + first + count > w.size()) {
                err = "a leaf's indices past the end of the buffer";
                return false;
            }
            if (first + count > m.index_count) {
                // the relayouts gather each mesh's index_count / 3 triangles and address a
                // leaf's triangles inside that range
                err = "a leaf's indices past the mesh's index count";
                return false;
            }
            for (uint64_t k = 0; k < count; ++k)
                if (w[m.indices + first + k] >= m.vertex_count) {
                    err = "an index past the mesh's vertices";
                    return false;
                }
            continue;
        }
        if (tag >= m.node_count || seen[tag]) {
            err = tag >= m.node_count ? "a node id past the mesh's nodes" : "a mesh tree reaches a node twice";
            return false;
        }
        seen[tag] = 1;
        for (int c = 0; c < 4; ++c) todo.push_back(w[m.nodes + (size_t)tag * NODE_WORDS + 6 + c]);
    }
    return true;
}

bool check_top_tree(const std::vector<uint32_t>& w, std::string& err) {
    const uint32_t count = w[0];
    if (w[1] == 0) return true;
    const uint64_t nodes = top_reserve(count) / TOP_WORDS;
    if ((uint64_t)w[1] + nodes * TOP_WORDS > w.size()) {
        err = "top level past the end of the buffer";
        return false;
    }
    std::vector<uint8_t> seen(nodes, 0);
    std::vector<uint32_t> todo{w[2]};
    while (!todo.empty()) {
        const uint32_t tag = todo.back();
        todo.pop_back();
        if (tag == EMPTY) continue;
        if (tag & LEAF_BIT) {
            if ((tag & ~LEAF_BIT) >= count) {
                err = "a top-level leaf past the instance records";
                return false;
            }
            continue;
        }
        if (tag >= nodes || seen[tag]) {
            err = tag >= nodes ? "a top-level node id past the top level" : "the top level reaches a node twice";
            return false;
        }
        seen[tag] = 1;
        for (int c = 0; c < 4; ++c) todo.push_back(w[w[1] + (size_t)tag * TOP_WORDS + 24 + c]);
    }
    return true;
}

}  // namespace

extern "C" rq_status rq_scene_from_gi_words(const uint32_t* w, size_t n, rq_scene** out) {
    if (!w || !out || n < HEADER_WORDS) {
        set_error("rq_scene_from_gi_words: buffer too small or null");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    auto scene = new rq_scene();
    scene->words.assign(w, w + n);
    uint32_t count = w[0];
    if (HEADER_WORDS + (uint64_t)count * INSTANCE_WORDS > n) {
        delete scene;
        set_error("rq_scene_from_gi_words: instance records past the end");
        return RQ_ERR_INVALID_INPUT;
    }
    uint64_t tris = 0, nodes = 0;
    std::vector<uint32_t> seen;
    std::unordered_set<uint64_t> checked_roots;  // node offset << 32 | root of each mesh tree checked
    std::string err;
    if (!check_top_tree(scene->words, err)) {
        delete scene;
        set_error("rq_scene_from_gi_words: " + err);
        return RQ_ERR_INVALID_INPUT;
    }
    for (uint32_t i = 0; i < count; ++i) {
        MeshView m;
        if (!mesh_view(scene->words, i, m, err)) {
            delete scene;
            set_error("rq_scene_from_gi_words: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
        const uint32_t root = scene->words[HEADER_WORDS + (size_t)i * INSTANCE_WORDS + 22];
        if (checked_roots.insert((uint64_t)m.nodes << 32 | root).second &&
            !check_mesh_tree(scene->words, m, root, err)) {
            delete scene;
            set_error("rq_scene_from_gi_words: instance " + std::to_string(i) + ": " + err);
            return RQ_ERR_INVALID_INPUT;
        }
        // Words 24-29 read after a hit (gi_rays_software.wgsl:297-316, the closest-hit
        // primitive remap :147-148): the primitive table holds one word per triangle, and a
        // merged world's identity table (word 24 zero) holds word-26 pairs in position order.
        const size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        const uint64_t prim_base = scene->words[r + 28];
        if (prim_base != 0 && prim_base + m.index_count / 3 > n) {
            err = "instance " + std::to_string(i) + ": its primitive table (word 28) runs past the buffer";
        } else if (scene->words[r + 24] == 0 && scene->words[r + 26] != 0) {
            const uint64_t map = scene->words[r + 25], pairs = scene->words[r + 26];
            if (map + 2 * pairs > n) {
                err = "instance " + std::to_string(i) + ": its identity table (words 25-26) runs past the buffer";
            } else {
                for (uint64_t k = 1; k < pairs; ++k)
                    if (scene->words[map + 2 * k] < scene->words[map + 2 * k - 2]) {
                        err = "instance " + std::to_string(i) + ": its identity table is not in position order";
                        break;
                    }
            }
        }
        if (!err.empty()) {
            delete scene;
            set_error("rq_scene_from_gi_words: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
        tris += m.index_count / 3;
        if (std::find(seen.begin(), seen.end(), m.nodes) == seen.end()) {
            seen.push_back(m.nodes);
            nodes += m.node_count;
        }
    }
    scene->stats[0] = count;
    scene->stats[1] = seen.size();
    scene->stats[2] = tris;
    scene->stats[4] = nodes;
    scene->stats[5] = w[2];
    scene_bounds(scene);
    *out = scene;
    return RQ_OK;
}
