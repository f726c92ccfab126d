// The reflection rays' software geometry (rl_geometry): the buffer that
// src/reflection_software.rs packs and shaders/reflections_software_bvh4.wgsl
// traverses. Same SAH BVH4 per mesh as the GI path (near_field::build_bvh_sah)
// and the same top level (instance_bvh::place); its own words:
//   header (8): [0] instance count, [1] top-level base word, [2] top-level root,
//               [4] word offset of the scene bounds (lo xyz, 0, hi xyz, 0);
//   instance record (24): 0-11 world-to-object rows, 12-17 world lo/hi (padded),
//               18 node base word, 19 triangle base word, 20 root tag,
//               21 placement slot, 22 1 if it casts no shadow, 23 zero;
//   node (16): the GI BVH4 record with leaf tags LEAF | first triangle << 4 | count;
//   triangle (12): corner a, edges b - a and c - a, source primitive, 0, 0.
// Every record starts on a 16-byte boundary (src/reflection_software.rs:26-44).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_set>

#include "internal.h"

namespace rq {
namespace {

struct RsBlas {
    std::vector<uint32_t> nodes;  // RS_NODE_WORDS a node
    std::vector<uint32_t> tris;   // RS_TRI_WORDS a triangle
    uint32_t root = 0;
    float lo[3], hi[3];
};

// Blas::from_bvh (src/reflection_software.rs:110-167).
bool rs_blas(const rq_mesh& m, RsBlas& out, std::string& err) {
    Bvh bvh;
    if (!build_bvh_sah(m.positions, m.vertex_count, m.indices, m.index_count, bvh, err)) return false;
    auto tag = [&](uint32_t child) -> uint32_t {
        if (child == EMPTY || !(child & LEAF_BIT)) return child;
        uint32_t l = child & ~LEAF_BIT;
        return LEAF_BIT | ((bvh.leaf_offset[l] / 3u) << 4) | (bvh.leaf_count[l] / 3u);
    };
    out.nodes.resize(bvh.nodes.size() * RS_NODE_WORDS);
    for (size_t i = 0; i < bvh.nodes.size(); ++i) {
        Node n = bvh.nodes[i];
        for (int c = 0; c < 4; ++c) n.children[c] = tag(n.children[c]);
        std::memcpy(&out.nodes[i * RS_NODE_WORDS], &n, sizeof(Node));
    }
    // Each BVH triangle's source primitive (ascending among identical triplets).
    std::vector<uint32_t> prims;
    map_primitives(bvh.indices, m.indices, m.index_count, prims);
    const size_t tris = bvh.indices.size() / 3;
    if (prims.size() != tris) {
        err = "reflection geometry: a BVH triangle has no source primitive";
        return false;
    }
    out.tris.reserve(tris * RS_TRI_WORDS);
    for (size_t t = 0; t < tris; ++t) {
        float p[3][3];
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) p[v][a] = m.positions[3ull * bvh.indices[3 * t + v] + a];
        for (int a = 0; a < 3; ++a) out.tris.push_back(f_bits(p[0][a]));
        for (int e = 1; e < 3; ++e)
            for (int a = 0; a < 3; ++a) out.tris.push_back(f_bits(p[e][a] - p[0][a]));
        out.tris.push_back(prims[t]);
        out.tris.push_back(0u);
        out.tris.push_back(0u);
    }
    // Object bounds over the referenced corners (bounds(indices.map(positions))).
    for (int a = 0; a < 3; ++a) {
        out.lo[a] = __builtin_inff();
        out.hi[a] = -__builtin_inff();
    }
    for (uint32_t i = 0; i < m.index_count; ++i)
        for (int a = 0; a < 3; ++a) {
            out.lo[a] = mn(out.lo[a], m.positions[3ull * m.indices[i] + a]);
            out.hi[a] = mx(out.hi[a], m.positions[3ull * m.indices[i] + a]);
        }
    out.root = tag(bvh.root);
    return true;
}

}  // namespace

// world_bounds (src/reflection_software.rs:372-388): the 8 transformed corners
// (glam transform_point3 order) widened by 1e-5 of the largest of |lo|, |hi|
// and the extent over the axes.
void rs_world_bounds(const float* M, const float lo_in[3], const float hi_in[3], float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = __builtin_inff();
        hi[a] = -__builtin_inff();
    }
    for (int c = 0; c < 8; ++c) {
        float p[3];
        for (int a = 0; a < 3; ++a) p[a] = (c & (1 << a)) ? hi_in[a] : lo_in[a];
        for (int a = 0; a < 3; ++a) {
            float wv = M[a] * p[0] + M[4 + a] * p[1];
            wv = wv + M[8 + a] * p[2];
            wv = wv + M[12 + a];
            lo[a] = mn(lo[a], wv);
            hi[a] = mx(hi[a], wv);
        }
    }
    float margin = 0.0f;
    for (int a = 0; a < 3; ++a) margin = mx(margin, mx(mx(std::fabs(hi[a]), std::fabs(lo[a])), hi[a] - lo[a]));
    margin = margin * 1e-5f;
    for (int a = 0; a < 3; ++a) {
        lo[a] = lo[a] - margin;
        hi[a] = hi[a] + margin;
    }
}

uint32_t rs_records_words(uint32_t instances) {
    return RS_HEADER_WORDS + instances * RS_RECORD_WORDS + top_reserve(instances) + 8u;
}

}  // namespace rq

using namespace rq;

extern "C" rq_status rq_scene_build_reflection(const rq_mesh* meshes, uint32_t mesh_count,
                                               const rq_instance* instances, uint32_t instance_count, int top_level,
                                               rq_scene** out) {
    const auto began = std::chrono::steady_clock::now();
    if (!out || (!meshes && mesh_count) || (!instances && instance_count)) {
        set_error("rq_scene_build_reflection: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    for (uint32_t i = 0; i < instance_count; ++i)
        if (instances[i].mesh >= mesh_count) {
            set_error("rq_scene_build_reflection: instance names a missing mesh");
            return RQ_ERR_INVALID_INPUT;
        }
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (meshes[m].index_count / 3u >= RS_MAX_MESH_TRIS) {
            set_error("rq_scene_build_reflection: a mesh of 2^27 or more triangles (the leaf tag holds the first "
                      "triangle in 27 bits)");
            return RQ_ERR_INVALID_INPUT;
        }
    std::vector<RsBlas> blas(mesh_count);
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
                    ok[m] = rs_blas(meshes[m], blas[m], errs[m]) ? 1 : 0;
                }
            });
        for (auto& th : pool) th.join();
    }
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (!ok[m]) {
            set_error("reflection geometry: " + errs[m]);
            return RQ_ERR_INVALID_INPUT;
        }
    // pack (src/reflection_software.rs:285-305): each mesh's nodes then triangles
    // after the header, records, top level and scene bounds.
    uint64_t at = rs_records_words(instance_count);
    std::vector<uint32_t> bases(2ull * mesh_count);
    for (uint32_t m = 0; m < mesh_count; ++m) {
        bases[2 * m] = (uint32_t)at;
        at += blas[m].nodes.size();
        bases[2 * m + 1] = (uint32_t)at;
        at += blas[m].tris.size();
        if (at >= 0xffffffffull) {
            set_error("reflection geometry exceeds 2^32 words");
            return RQ_ERR_INVALID_INPUT;
        }
    }
    // instance_words (src/reflection_software.rs:318-345): singular models are skipped.
    std::vector<uint32_t> head(RS_HEADER_WORDS, 0u);
    head[4] = rs_records_words(instance_count) - 8u;
    uint32_t count = 0;
    for (uint32_t i = 0; i < instance_count; ++i) {
        const rq_instance& ins = instances[i];
        const RsBlas& b = blas[ins.mesh];
        float inv[16];
        if (!mat4_inverse(ins.model, inv)) continue;
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 4; ++c) head.push_back(f_bits(inv[c * 4 + row]));
        float lo[3], hi[3];
        rs_world_bounds(ins.model, b.lo, b.hi, lo, hi);
        for (int a = 0; a < 3; ++a) head.push_back(f_bits(lo[a]));
        for (int a = 0; a < 3; ++a) head.push_back(f_bits(hi[a]));
        head.push_back(bases[2 * ins.mesh]);
        head.push_back(bases[2 * ins.mesh + 1]);
        head.push_back(b.root);
        head.push_back(ins.placement);
        head.push_back(ins.flags & 1u);
        head.push_back(0u);
        ++count;
    }
    head[0] = count;
    // the top level (src/instance_bvh.rs place) over the records, room for every slot
    top_place(head, RS_HEADER_WORDS, RS_RECORD_WORDS, instance_count, nullptr, 0, top_level);
    // finish_scene_bounds (src/reflection_software.rs:263-282)
    {
        float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        for (uint32_t i = 0; i < count; ++i) {
            size_t r = RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS;
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
    auto scene = new rq_scene();
    scene->kind = RQ_SCENE_REFLECTION;
    std::vector<uint32_t>& words = scene->words;
    words.assign(at, 0u);
    std::copy(head.begin(), head.end(), words.begin());
    uint64_t tris = 0, uniq = 0, nodes = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        std::copy(blas[m].nodes.begin(), blas[m].nodes.end(), words.begin() + bases[2 * m]);
        std::copy(blas[m].tris.begin(), blas[m].tris.end(), words.begin() + bases[2 * m + 1]);
        uniq += blas[m].tris.size() / RS_TRI_WORDS;
        nodes += blas[m].nodes.size() / RS_NODE_WORDS;
    }
    for (uint32_t i = 0; i < instance_count; ++i) tris += blas[instances[i].mesh].tris.size() / RS_TRI_WORDS;
    scene->stats[0] = count;
    scene->stats[1] = mesh_count;
    scene->stats[2] = tris;
    scene->stats[3] = uniq;
    scene->stats[4] = nodes;
    scene->stats[5] = words[2];
    scene_bounds(scene);
    scene->stats[7] = (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                          std::chrono::steady_clock::now() - began).count();
    *out = scene;
    return RQ_OK;
}

namespace {

// Words from outside this library: traversal reads node and triangle words by
// the ids it finds and ends only if each tree is finite. Each instance's tree
// under its root and the top level must reach no node twice and stay inside
// the buffer (nodes at their base, leaves' triangles at the triangle base,
// top-level leaves among the records).
bool check_rs_tree(const uint32_t* w, size_t n, uint64_t base, uint32_t node_words, uint32_t root, uint64_t tri_base,
                   uint32_t child_word, bool top, std::string& err) {
    std::unordered_set<uint32_t> seen;
    std::vector<uint32_t> todo{root};
    while (!todo.empty()) {
        const uint32_t tag = todo.back();
        todo.pop_back();
        if (tag == EMPTY) continue;
        if (tag & LEAF_BIT) {
            if (top) {
                if ((tag & ~LEAF_BIT) >= w[0]) {
                    err = "a top-level leaf past the instance records";
                    return false;
                }
            } else if (tri_base + (uint64_t)RS_TRI_WORDS * (((tag & ~LEAF_BIT) >> 4) + (tag & 15u)) > n) {
                err = "a leaf's triangles past the end of the buffer";
                return false;
            }
            continue;
        }
        if (base + (uint64_t)node_words * ((uint64_t)tag + 1) > n || !seen.insert(tag).second) {
            err = seen.count(tag) ? "a tree reaches a node twice" : "a node past the end of the buffer";
            return false;
        }
        for (int c = 0; c < 4; ++c) todo.push_back(w[base + (uint64_t)node_words * tag + child_word + c]);
    }
    return true;
}

}  // namespace

extern "C" rq_status rq_scene_from_rl_words(const uint32_t* w, size_t n, rq_scene** out) {
    if (!out || !w || n < RS_HEADER_WORDS) {
        set_error("rq_scene_from_rl_words: needs at least the 8 header words");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    const uint64_t count = w[0];
    if (RS_HEADER_WORDS + count * RS_RECORD_WORDS > n || (count && ((uint64_t)w[4] + 8 > n || (w[4] & 3u)))) {
        set_error("rq_scene_from_rl_words: records or scene bounds past the end");
        return RQ_ERR_INVALID_INPUT;
    }
    if (count && w[1] && ((uint64_t)w[1] + RS_TOP_NODE_WORDS > n || (w[1] & 3u))) {
        set_error("rq_scene_from_rl_words: top level past the end");
        return RQ_ERR_INVALID_INPUT;
    }
    for (uint64_t i = 0; i < count; ++i) {
        const uint32_t* r = w + RS_HEADER_WORDS + i * RS_RECORD_WORDS;
        if (r[18] > n || r[19] > n || (r[18] & 3u) || (r[19] & 3u)) {
            set_error("rq_scene_from_rl_words: instance " + std::to_string(i) + " names arrays past the end");
            return RQ_ERR_INVALID_INPUT;
        }
    }
    {
        std::string err;
        std::unordered_set<uint64_t> checked;  // node base << 32 | root of each tree checked
        for (uint64_t i = 0; i < count; ++i) {
            const uint32_t* r = w + RS_HEADER_WORDS + i * RS_RECORD_WORDS;
            if (checked.insert((uint64_t)r[18] << 32 | r[20]).second &&
                !check_rs_tree(w, n, r[18], RS_NODE_WORDS, r[20], r[19], 6, false, err)) {
                set_error("rq_scene_from_rl_words: instance " + std::to_string(i) + ": " + err);
                return RQ_ERR_INVALID_INPUT;
            }
        }
        if (count && w[1] && !check_rs_tree(w, n, w[1], RS_TOP_NODE_WORDS, w[2], 0, 24, true, err)) {
            set_error("rq_scene_from_rl_words: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
    }
    auto scene = new rq_scene();
    scene->kind = RQ_SCENE_REFLECTION;
    scene->words.assign(w, w + n);
    scene->stats[0] = count;
    scene->stats[5] = w[2];
    scene_bounds(scene);
    *out = scene;
    return RQ_OK;
}
