// Cooked software-ray geometry read from disk, and scenes packed from it:
//
// - the cluster BVH sidecar, "RLB\0" (crates/asset/src/cluster_bvh/codec.rs:4-188):
//   80-byte nodes (the 16 runtime words plus four child errors), 32-byte leaves,
//   24-byte levels (legal cuts), level leaf lists, vec4 positions, indices; checked
//   by Bvh::validate (crates/asset/src/cluster_bvh.rs:170-333);
// - the near field, "RNF\0" (crates/asset/src/near_field.rs:540-680): settings,
//   source digest, one single-level cluster BVH per mesh, and per placement the
//   transform it was cooked under and two octahedral visibility masks per vertex
//   (NearField::validate, :106-135);
// - rq_scene_build_cooked: meshes given as triangle lists (set_indirect_geometry),
//   cooked cluster BVHs at a level (set_indirect_bvh, src/gi_cache.rs:3209-3217) or
//   near-field meshes (set_indirect_near_field, :3218-3270), packed like
//   rq_scene_build with each placement's masks after the meshes and their base in
//   instance word 23 for the instances with the cooked transform's exact bits
//   (src/gi_bvh.rs:95-104, 151-226).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

#include "internal.h"

namespace rq {

constexpr size_t RLB_NODE_BYTES = 80, RLB_LEAF_BYTES = 32, RLB_LEVEL_BYTES = 24;
constexpr size_t CLUSTER_MAX_BYTES = 150000000, NEAR_MAX_BYTES = 160000000, RNF_HEADER_BYTES = 64;

struct ClusterLeaf {
    uint32_t mesh, cluster, page, source_index_offset, offset, count;
    float error;
    uint32_t material;
};
struct ClusterLevel {
    uint32_t mesh, root;
    float requested_error, max_error;
    uint32_t offset, count;
};
struct ClusterBvh {
    std::vector<Node> nodes;
    std::vector<std::array<float, 4>> errors;  // per node, per slot
    std::vector<ClusterLeaf> leaves;
    std::vector<ClusterLevel> levels;
    std::vector<uint32_t> level_leaves;
    std::vector<float> positions;  // x, y, z, 0 per vertex
    std::vector<uint32_t> indices;
    // cluster_bvh.rs:134-141
    size_t gpu_bytes() const {
        return nodes.size() * RLB_NODE_BYTES + leaves.size() * RLB_LEAF_BYTES + levels.size() * RLB_LEVEL_BYTES +
               level_leaves.size() * 4 + positions.size() * 4 + indices.size() * 4;
    }
};

struct NearPlacement {
    uint32_t mesh;
    float transform[16];  // column-major
    std::vector<uint32_t> masks;
};

namespace {

// Span::range (crates/asset/src/virtual_mesh.rs:75-82).
bool span_range(uint32_t offset, uint32_t count, size_t length, size_t& begin, size_t& end, std::string& err) {
    const uint64_t e = (uint64_t)offset + count;
    if (e > 0xffffffffull) {
        err = "virtual span overflow";
        return false;
    }
    if (e > length) {
        err = "virtual span out of range";
        return false;
    }
    begin = offset;
    end = (size_t)e;
    return true;
}

float rmax(float a, float b) { return std::isnan(a) ? b : std::isnan(b) ? a : (a > b ? a : b); }  // f32::max

void decode_plane(const Node& n, int slot, bool fused, double lo[3], double hi[3]) {
    for (int a = 0; a < 3; ++a) {
        const float o = n.origin[a], s = n.step[a];
        const float ql = (float)n.lower[a][slot], qh = (float)n.upper[a][slot];
        lo[a] = fused ? (double)std::fmaf(s, ql, o) : (double)(o + s * ql);
        hi[a] = fused ? (double)std::fmaf(s, qh, o) : (double)(o + s * qh);
    }
}

}  // namespace

// Bvh::validate (crates/asset/src/cluster_bvh.rs:172-333).
bool validate_cluster(const ClusterBvh& b, size_t max_bytes, std::string& err) {
    if (b.gpu_bytes() > max_bytes || b.nodes.size() >= LEAF_BIT || b.leaves.size() >= LEAF_BIT - 1 ||
        b.levels.empty()) {
        err = "cluster BVH byte/count limit or empty levels";
        return false;
    }
    for (size_t v = 0; v < b.positions.size(); v += 4) {
        if (!std::isfinite(b.positions[v]) || !std::isfinite(b.positions[v + 1]) ||
            !std::isfinite(b.positions[v + 2]) || !std::isfinite(b.positions[v + 3]) || b.positions[v + 3] != 0.0f) {
            err = "invalid cluster BVH position";
            return false;
        }
    }
    const size_t verts = b.positions.size() / 4;
    for (uint32_t i : b.indices)
        if (i >= verts) {
            err = "cluster BVH vertex index";
            return false;
        }
    size_t end = 0;
    for (const ClusterLeaf& l : b.leaves) {
        size_t r0, r1;
        if (!span_range(l.offset, l.count, b.indices.size(), r0, r1, err)) return false;
        if (r0 != end || r1 == r0 || (r1 - r0) % 3 || !std::isfinite(l.error) || l.error < 0.0f) {
            err = "invalid cluster BVH leaf";
            return false;
        }
        end = r1;
    }
    if (end != b.indices.size()) {
        err = "unowned cluster BVH indices";
        return false;
    }
    struct Box {
        double lo[3], hi[3];
        float error;
    };
    std::vector<Box> boxes;
    boxes.reserve(b.nodes.size());
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const Node& n = b.nodes[i];
        bool frame = true, any = false;
        for (int a = 0; a < 3; ++a)
            frame = frame && std::isfinite(n.origin[a]) && std::isfinite(n.step[a]) && n.step[a] > 0.0f;
        for (uint32_t c : n.children) any = any || c != EMPTY;
        if (!frame || !any) {
            err = "invalid cluster BVH node frame";
            return false;
        }
        Box box;
        for (int a = 0; a < 3; ++a) {
            box.lo[a] = INFINITY;
            box.hi[a] = -INFINITY;
        }
        box.error = 0.0f;
        for (int slot = 0; slot < 4; ++slot) {
            const uint32_t child = n.children[slot];
            if (child == EMPTY) {
                bool zero = b.errors[i][slot] == 0.0f;
                for (int a = 0; a < 3; ++a) zero = zero && n.lower[a][slot] == 0 && n.upper[a][slot] == 0;
                if (!zero) {
                    err = "noncanonical empty BVH child";
                    return false;
                }
                continue;
            }
            double clo[3], chi[3];
            float ce;
            if (child & LEAF_BIT) {
                const uint32_t id = child & ~LEAF_BIT;
                if (id >= b.leaves.size()) {
                    err = "BVH leaf reference";
                    return false;
                }
                const ClusterLeaf& l = b.leaves[id];
                for (int a = 0; a < 3; ++a) {
                    clo[a] = INFINITY;
                    chi[a] = -INFINITY;
                }
                for (uint32_t k = l.offset; k < l.offset + l.count; ++k)
                    for (int a = 0; a < 3; ++a) {
                        const double p = b.positions[4ull * b.indices[k] + a];
                        clo[a] = std::min(clo[a], p);
                        chi[a] = std::max(chi[a], p);
                    }
                ce = l.error;
            } else {
                if (child >= i) {
                    err = "BVH unordered/cyclic node";
                    return false;
                }
                for (int a = 0; a < 3; ++a) {
                    clo[a] = boxes[child].lo[a];
                    chi[a] = boxes[child].hi[a];
                }
                ce = boxes[child].error;
            }
            double pa[3], pb[3], fa[3], fb[3];
            decode_plane(n, slot, false, pa, pb);
            decode_plane(n, slot, true, fa, fb);
            bool ok = std::isfinite(b.errors[i][slot]) && b.errors[i][slot] == ce;
            for (int a = 0; a < 3 && ok; ++a)
                ok = std::isfinite(pa[a]) && std::isfinite(pb[a]) && std::isfinite(fa[a]) && std::isfinite(fb[a]) &&
                     pa[a] <= clo[a] && pb[a] >= chi[a] && fa[a] <= clo[a] && fb[a] >= chi[a];
            if (!ok) {
                err = "BVH child bounds/error do not enclose geometry";
                return false;
            }
            for (int a = 0; a < 3; ++a) {
                box.lo[a] = std::min(std::min(box.lo[a], pa[a]), fa[a]);
                box.hi[a] = std::max(std::max(box.hi[a], pb[a]), fb[a]);
            }
            box.error = rmax(box.error, ce);
        }
        boxes.push_back(box);
    }
    std::vector<char> used_nodes(b.nodes.size(), 0), used_leaves(b.leaves.size(), 0);
    size_t cut_end = 0;
    bool have_previous = false;
    uint32_t prev_mesh = 0;
    float prev_error = 0.0f;
    for (const ClusterLevel& level : b.levels) {
        if (!std::isfinite(level.requested_error) || level.requested_error < 0.0f ||
            !std::isfinite(level.max_error) || level.max_error < 0.0f || level.max_error > level.requested_error ||
            (have_previous && (level.mesh < prev_mesh ||
                               (level.mesh == prev_mesh && level.requested_error <= prev_error)))) {
            err = "invalid BVH level/error order";
            return false;
        }
        have_previous = true;
        prev_mesh = level.mesh;
        prev_error = level.requested_error;
        size_t r0, r1;
        if (!span_range(level.offset, level.count, b.level_leaves.size(), r0, r1, err)) return false;
        if (r0 != cut_end || r1 == r0) {
            err = "invalid BVH cut span";
            return false;
        }
        cut_end = r1;
        for (size_t k = r0 + 1; k < r1; ++k)
            if (b.level_leaves[k - 1] >= b.level_leaves[k]) {
                err = "BVH duplicate/unordered cut leaf";
                return false;
            }
        std::vector<uint32_t> found, stack{level.root};
        size_t visits = 0;
        float error = 0.0f;
        while (!stack.empty()) {
            const uint32_t child = stack.back();
            stack.pop_back();
            if (++visits > b.nodes.size() + b.leaves.size()) {
                err = "BVH repeated subtree";
                return false;
            }
            if (child == EMPTY) {
                err = "BVH empty root";
                return false;
            }
            if (child & LEAF_BIT) {
                const uint32_t id = child & ~LEAF_BIT;
                if (id >= b.leaves.size()) {
                    err = "BVH cut leaf reference";
                    return false;
                }
                if (b.leaves[id].mesh != level.mesh) {
                    err = "BVH cross-mesh cut";
                    return false;
                }
                used_leaves[id] = 1;
                error = rmax(error, b.leaves[id].error);
                found.push_back(id);
            } else {
                if (child >= b.nodes.size()) {
                    err = "BVH root reference";
                    return false;
                }
                used_nodes[child] = 1;
                for (uint32_t c : b.nodes[child].children)
                    if (c != EMPTY) stack.push_back(c);
            }
        }
        std::sort(found.begin(), found.end());
        if (found.size() != r1 - r0 || !std::equal(found.begin(), found.end(), b.level_leaves.begin() + r0) ||
            error != level.max_error) {
            err = "BVH cut/root disagreement";
            return false;
        }
    }
    if (cut_end != b.level_leaves.size() || std::find(used_nodes.begin(), used_nodes.end(), 0) != used_nodes.end() ||
        std::find(used_leaves.begin(), used_leaves.end(), 0) != used_leaves.end()) {
        err = "unreachable cluster BVH records";
        return false;
    }
    return true;
}

namespace {

struct Writer {
    std::vector<uint8_t> out;
    void u32(uint32_t v) {
        for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(v >> (8 * i)));
    }
    void u64(uint64_t v) {
        for (int i = 0; i < 8; ++i) out.push_back((uint8_t)(v >> (8 * i)));
    }
    void f32(float f) { u32(f_bits(f)); }
    void bytes(const uint8_t* p, size_t n) { out.insert(out.end(), p, p + n); }
};

struct Reader {
    const uint8_t* p;
    size_t n, at = 0;
    bool take(size_t k, const uint8_t*& q) {
        if (k > n - at) return false;
        q = p + at;
        at += k;
        return true;
    }
    bool u32(uint32_t& v) {
        const uint8_t* q;
        if (!take(4, q)) return false;
        v = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
        return true;
    }
    bool u64(uint64_t& v) {
        uint32_t lo, hi;
        if (!u32(lo) || !u32(hi)) return false;
        v = (uint64_t)lo | ((uint64_t)hi << 32);
        return true;
    }
    bool f32(float& f) {
        uint32_t v;
        if (!u32(v)) return false;
        f = bits_f(v);
        return true;
    }
};

}  // namespace

// codec.rs:4-75
bool encode_cluster(const ClusterBvh& b, size_t max_bytes, std::vector<uint8_t>& out, std::string& err) {
    if (!validate_cluster(b, max_bytes, err)) return false;
    const size_t counts[6] = {b.nodes.size(),        b.leaves.size(),        b.levels.size(),
                              b.level_leaves.size(), b.positions.size() / 4, b.indices.size()};
    Writer w;
    w.bytes((const uint8_t*)"RLB\0", 4);
    w.u32(1);
    for (size_t c : counts) {
        if (c > 0xffffffffull) {
            err = "BVH encoded count overflow";
            return false;
        }
        w.u32((uint32_t)c);
    }
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const Node& n = b.nodes[i];
        for (float f : n.origin) w.f32(f);
        for (float f : n.step) w.f32(f);
        for (uint32_t c : n.children) w.u32(c);
        for (int a = 0; a < 3; ++a) w.bytes(n.lower[a], 4);
        for (int a = 0; a < 3; ++a) w.bytes(n.upper[a], 4);
        for (float f : b.errors[i]) w.f32(f);
    }
    for (const ClusterLeaf& l : b.leaves) {
        for (uint32_t v : {l.mesh, l.cluster, l.page, l.source_index_offset, l.offset, l.count}) w.u32(v);
        w.f32(l.error);
        w.u32(l.material);
    }
    for (const ClusterLevel& l : b.levels) {
        w.u32(l.mesh);
        w.u32(l.root);
        w.f32(l.requested_error);
        w.f32(l.max_error);
        w.u32(l.offset);
        w.u32(l.count);
    }
    for (uint32_t v : b.level_leaves) w.u32(v);
    for (float f : b.positions) w.f32(f);
    for (uint32_t v : b.indices) w.u32(v);
    if (w.out.size() > max_bytes) {
        err = "BVH encoded byte limit";
        return false;
    }
    out.swap(w.out);
    return true;
}

// codec.rs:108-188
bool decode_cluster(const uint8_t* bytes, size_t n, size_t max_bytes, ClusterBvh& b, std::string& err) {
    if (n > max_bytes) {
        err = "BVH file byte limit";
        return false;
    }
    Reader r{bytes, n};
    const uint8_t* magic;
    uint32_t version;
    if (!r.take(4, magic)) {
        err = "truncated cluster BVH";
        return false;
    }
    if (std::memcmp(magic, "RLB\0", 4) != 0) {
        err = "unknown cluster BVH schema";
        return false;
    }
    if (!r.u32(version)) {
        err = "truncated cluster BVH";
        return false;
    }
    if (version != 1) {
        err = "unknown cluster BVH schema";
        return false;
    }
    const size_t strides[6] = {RLB_NODE_BYTES, RLB_LEAF_BYTES, RLB_LEVEL_BYTES, 4, 16, 4};
    uint64_t counts[6], length = 32;
    for (int k = 0; k < 6; ++k) {
        uint32_t c;
        if (!r.u32(c)) {
            err = "truncated cluster BVH";
            return false;
        }
        counts[k] = c;
        length += counts[k] * strides[k];
    }
    if (length != n || length > max_bytes) {
        err = "BVH extent/count limit";
        return false;
    }
    b = ClusterBvh();
    b.nodes.resize(counts[0]);
    b.errors.resize(counts[0]);
    for (uint64_t i = 0; i < counts[0]; ++i) {
        Node& nd = b.nodes[i];
        for (float& f : nd.origin) r.f32(f);
        for (float& f : nd.step) r.f32(f);
        for (uint32_t& c : nd.children) r.u32(c);
        const uint8_t* q = nullptr;  // the length check above covers every read
        for (int a = 0; a < 3; ++a) {
            r.take(4, q);
            std::memcpy(nd.lower[a], q, 4);
        }
        for (int a = 0; a < 3; ++a) {
            r.take(4, q);
            std::memcpy(nd.upper[a], q, 4);
        }
        for (float& f : b.errors[i]) r.f32(f);
    }
    b.leaves.resize(counts[1]);
    for (ClusterLeaf& l : b.leaves) {
        r.u32(l.mesh);
        r.u32(l.cluster);
        r.u32(l.page);
        r.u32(l.source_index_offset);
        r.u32(l.offset);
        r.u32(l.count);
        r.f32(l.error);
        r.u32(l.material);
    }
    b.levels.resize(counts[2]);
    for (ClusterLevel& l : b.levels) {
        r.u32(l.mesh);
        r.u32(l.root);
        r.f32(l.requested_error);
        r.f32(l.max_error);
        r.u32(l.offset);
        r.u32(l.count);
    }
    b.level_leaves.resize(counts[3]);
    for (uint32_t& v : b.level_leaves) r.u32(v);
    b.positions.resize(4 * counts[4]);
    for (float& f : b.positions) r.f32(f);
    b.indices.resize(counts[5]);
    for (uint32_t& v : b.indices) r.u32(v);
    return validate_cluster(b, max_bytes, err);
}

// near_field.rs:172-262 (build): the one-level, zero-error cluster BVH of a
// triangle list -- the runtime builder's tree with the cook's leaf records.
bool cluster_from_mesh(const float* positions, uint32_t vertex_count, const uint32_t* indices, uint32_t index_count,
                       bool sah, ClusterBvh& out, std::string& err) {
    Bvh bvh;
    if (!(sah ? build_bvh_sah(positions, vertex_count, indices, index_count, bvh, err)
              : build_bvh(positions, vertex_count, indices, index_count, false, bvh, err)))
        return false;
    out = ClusterBvh();
    out.nodes = bvh.nodes;
    out.errors.assign(bvh.nodes.size(), std::array<float, 4>{0.0f, 0.0f, 0.0f, 0.0f});
    for (size_t l = 0; l < bvh.leaf_offset.size(); ++l)
        out.leaves.push_back(ClusterLeaf{0, (uint32_t)l, 0, 3 * bvh.leaf_first[l], bvh.leaf_offset[l],
                                         bvh.leaf_count[l], 0.0f, 0});
    const uint32_t leaves = (uint32_t)out.leaves.size();
    for (uint32_t l = 0; l < leaves; ++l) out.level_leaves.push_back(l);
    out.levels.push_back(ClusterLevel{0, bvh.root, 0.0f, 0.0f, 0, leaves});
    out.positions.reserve(4ull * vertex_count);
    for (uint32_t v = 0; v < vertex_count; ++v) {
        for (int a = 0; a < 3; ++a) out.positions.push_back(positions[3ull * v + a]);
        out.positions.push_back(0.0f);
    }
    out.indices = bvh.indices;
    return validate_cluster(out, SIZE_MAX, err);
}

// Blas::new (src/gi_bvh.rs:59-80): level `level`'s root over the whole record set.
bool blas_from_cluster(const ClusterBvh& c, uint32_t level, Blas& out, std::string& err) {
    if (level >= c.levels.size()) {
        err = "indirect geometry level";
        return false;
    }
    out = Blas();
    out.bvh.nodes = c.nodes;
    for (const ClusterLeaf& l : c.leaves) {
        out.bvh.leaf_offset.push_back(l.offset);
        out.bvh.leaf_count.push_back(l.count);
    }
    out.bvh.positions.reserve(3 * (c.positions.size() / 4));
    for (size_t v = 0; v < c.positions.size(); v += 4)
        for (int a = 0; a < 3; ++a) out.bvh.positions.push_back(c.positions[v + a]);
    out.bvh.indices = c.indices;
    out.bvh.root = c.levels[level].root;
    blas_bounds(out);
    return true;
}

}  // namespace rq

struct rq_cluster_bvh {
    rq::ClusterBvh b;
};

struct rq_near_field {
    float lod_error = 0, reach = 0, skip = 0;
    uint32_t bins = 0;
    uint8_t source[32];
    std::vector<rq::ClusterBvh> meshes;
    std::vector<rq::NearPlacement> placements;
};

using namespace rq;

namespace {

// Settings::validate (near_field.rs:39-52).
bool near_settings_ok(const rq_near_field& f) {
    return std::isfinite(f.lod_error) && f.lod_error >= 0.0f && std::isfinite(f.reach) && f.reach > 0.0f &&
           std::isfinite(f.skip) && f.skip >= 0.0f && f.skip < f.reach &&
           (f.bins == 0 || f.bins == 8 || f.bins == 16 || f.bins == 32);
}

// NearField::validate (near_field.rs:106-135).
bool validate_near(const rq_near_field& f, size_t max_bytes, std::string& err) {
    if (!near_settings_ok(f)) {
        err = "invalid near-field settings";
        return false;
    }
    size_t bytes = 0;
    for (const ClusterBvh& b : f.meshes) bytes += b.gpu_bytes();
    for (const NearPlacement& p : f.placements) bytes += p.masks.size() * 4;
    if (bytes > max_bytes || f.meshes.empty()) {
        err = "near field byte limit or no meshes";
        return false;
    }
    for (const ClusterBvh& b : f.meshes) {
        if (!validate_cluster(b, max_bytes, err)) return false;
        if (b.levels.size() != 1) {
            err = "near-field mesh must have one level";
            return false;
        }
    }
    const size_t words = (size_t)f.bins * f.bins / 32;
    for (const NearPlacement& p : f.placements) {
        if (p.mesh >= f.meshes.size()) {
            err = "near-field placement mesh";
            return false;
        }
        bool finite = true;
        for (float v : p.transform) finite = finite && std::isfinite(v);
        if (p.masks.size() != f.meshes[p.mesh].positions.size() / 4 * 2 * words || !finite) {
            err = "near-field placement masks or transform";
            return false;
        }
    }
    return true;
}

rq_status fail(const std::string& what) {
    set_error(what);
    return RQ_ERR_INVALID_INPUT;
}

rq_status copy_out(const std::vector<uint8_t>& bytes, uint8_t* out, size_t capacity, size_t* length) {
    *length = bytes.size();
    if (out && capacity >= bytes.size()) std::memcpy(out, bytes.data(), bytes.size());
    return RQ_OK;
}

}  // namespace

extern "C" rq_status rq_cluster_bvh_decode(const uint8_t* bytes, size_t length, size_t max_bytes,
                                           rq_cluster_bvh** out) {
    if (!out || (!bytes && length)) return fail("rq_cluster_bvh_decode: null argument");
    *out = nullptr;
    auto c = new rq_cluster_bvh();
    std::string err;
    if (!decode_cluster(bytes, length, max_bytes ? max_bytes : CLUSTER_MAX_BYTES, c->b, err)) {
        delete c;
        return fail(err);
    }
    *out = c;
    return RQ_OK;
}

extern "C" rq_status rq_cluster_bvh_from_mesh(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                                              uint32_t index_count, int sah, rq_cluster_bvh** out) {
    if (!out || (!positions && vertex_count) || (!indices && index_count))
        return fail("rq_cluster_bvh_from_mesh: null argument");
    *out = nullptr;
    auto c = new rq_cluster_bvh();
    std::string err;
    if (!cluster_from_mesh(positions, vertex_count, indices, index_count, sah != 0, c->b, err)) {
        delete c;
        return fail(err);
    }
    *out = c;
    return RQ_OK;
}

extern "C" rq_status rq_cluster_bvh_encode(const rq_cluster_bvh* bvh, size_t max_bytes, uint8_t* out,
                                           size_t capacity, size_t* length) {
    if (!bvh || !length) return fail("rq_cluster_bvh_encode: null argument");
    std::vector<uint8_t> bytes;
    std::string err;
    if (!encode_cluster(bvh->b, max_bytes ? max_bytes : CLUSTER_MAX_BYTES, bytes, err)) return fail(err);
    return copy_out(bytes, out, capacity, length);
}

extern "C" rq_status rq_cluster_bvh_counts(const rq_cluster_bvh* bvh, uint32_t counts[6]) {
    if (!bvh || !counts) return fail("rq_cluster_bvh_counts: null argument");
    const ClusterBvh& b = bvh->b;
    const size_t c[6] = {b.nodes.size(),        b.leaves.size(),        b.levels.size(),
                         b.level_leaves.size(), b.positions.size() / 4, b.indices.size()};
    for (int k = 0; k < 6; ++k) counts[k] = (uint32_t)c[k];
    return RQ_OK;
}

extern "C" void rq_cluster_bvh_destroy(rq_cluster_bvh* bvh) { delete bvh; }

// near_field.rs:585-648
extern "C" rq_status rq_near_field_decode(const uint8_t* bytes, size_t length, size_t max_bytes,
                                          rq_near_field** out) {
    if (!out || (!bytes && length)) return fail("rq_near_field_decode: null argument");
    *out = nullptr;
    const size_t cap = max_bytes ? max_bytes : NEAR_MAX_BYTES;
    if (length < RNF_HEADER_BYTES || std::memcmp(bytes, "RNF\0", 4) != 0 || length > cap + (1u << 24))
        return fail("near-field header or size");
    Reader r{bytes, length, 4};
    auto f = std::make_unique<rq_near_field>();
    uint32_t version = 0, meshes = 0, placements = 0, bits[3];
    r.u32(version);
    if (version != 1) return fail("unsupported near-field version");
    r.u32(meshes);
    r.u32(placements);
    r.u32(f->bins);
    for (uint32_t& b : bits) r.u32(b);
    f->lod_error = bits_f(bits[0]);
    f->reach = bits_f(bits[1]);
    f->skip = bits_f(bits[2]);
    if (!near_settings_ok(*f)) return fail("invalid near-field settings");
    const uint8_t* src;
    r.take(32, src);
    std::memcpy(f->source, src, 32);
    if (meshes > length / 8 || placements > length / 76) return fail("near-field counts exceed the file");
    for (uint32_t m = 0; m < meshes; ++m) {
        uint64_t n;
        const uint8_t* blob;
        if (!r.u64(n) || n > length || !r.take((size_t)n, blob)) return fail("near-field truncated");
        ClusterBvh b;
        std::string err;
        if (!decode_cluster(blob, (size_t)n, cap, b, err)) return fail(err);
        f->meshes.push_back(std::move(b));
    }
    for (uint32_t p = 0; p < placements; ++p) {
        NearPlacement pl;
        uint64_t words;
        bool ok = r.u32(pl.mesh);
        for (float& v : pl.transform) ok = ok && r.f32(v);
        ok = ok && r.u64(words);
        if (!ok) return fail("near-field truncated");
        if (words > (length - r.at) / 4) return fail("near-field truncated");
        pl.masks.resize((size_t)words);
        for (uint32_t& w : pl.masks) r.u32(w);
        f->placements.push_back(std::move(pl));
    }
    if (r.at != length) return fail("near-field trailing bytes");
    std::string err;
    if (!validate_near(*f, cap, err)) return fail(err);
    *out = f.release();
    return RQ_OK;
}

// near_field.rs:546-583
extern "C" rq_status rq_near_field_encode(const rq_near_field* field, size_t max_bytes, uint8_t* out,
                                          size_t capacity, size_t* length) {
    if (!field || !length) return fail("rq_near_field_encode: null argument");
    const size_t cap = max_bytes ? max_bytes : NEAR_MAX_BYTES;
    std::string err;
    if (!validate_near(*field, cap, err)) return fail(err);
    Writer w;
    w.bytes((const uint8_t*)"RNF\0", 4);
    for (uint32_t v : {1u, (uint32_t)field->meshes.size(), (uint32_t)field->placements.size(), field->bins,
                       f_bits(field->lod_error), f_bits(field->reach), f_bits(field->skip)})
        w.u32(v);
    w.bytes(field->source, 32);
    for (const ClusterBvh& b : field->meshes) {
        std::vector<uint8_t> blob;
        if (!encode_cluster(b, cap, blob, err)) return fail(err);
        w.u64(blob.size());
        w.bytes(blob.data(), blob.size());
    }
    for (const NearPlacement& p : field->placements) {
        w.u32(p.mesh);
        for (float v : p.transform) w.f32(v);
        w.u64(p.masks.size());
        for (uint32_t m : p.masks) w.u32(m);
    }
    if (w.out.size() > cap + RNF_HEADER_BYTES + 4096) return fail("near-field encoded size");
    return copy_out(w.out, out, capacity, length);
}

extern "C" rq_status rq_near_field_info(const rq_near_field* field, uint32_t counts[2], uint32_t* bins,
                                        float settings[3], uint8_t source[32]) {
    if (!field) return fail("rq_near_field_info: null argument");
    if (counts) {
        counts[0] = (uint32_t)field->meshes.size();
        counts[1] = (uint32_t)field->placements.size();
    }
    if (bins) *bins = field->bins;
    if (settings) {
        settings[0] = field->lod_error;
        settings[1] = field->reach;
        settings[2] = field->skip;
    }
    if (source) std::memcpy(source, field->source, 32);
    return RQ_OK;
}

extern "C" void rq_near_field_destroy(rq_near_field* field) { delete field; }

extern "C" rq_status rq_scene_build_cooked(const rq_cooked_mesh* meshes, uint32_t mesh_count,
                                           const rq_instance* instances, uint32_t instance_count, int top_level,
                                           rq_scene** out) {
    const auto began = std::chrono::steady_clock::now();
    if (!out || (!meshes && mesh_count) || (!instances && instance_count))
        return fail("rq_scene_build_cooked: null argument");
    *out = nullptr;
    for (uint32_t i = 0; i < instance_count; ++i)
        if (instances[i].mesh >= mesh_count) return fail("rq_scene_build_cooked: instance names a missing mesh");
    uint32_t bins = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        const rq_cooked_mesh& c = meshes[m];
        if ((c.mesh != nullptr) + (c.bvh != nullptr) + (c.field != nullptr) != 1)
            return fail("rq_scene_build_cooked: each mesh names exactly one source");
        // src/gi_cache.rs:3228-3233: every cooked mesh with masks shares the bins
        if (c.field && c.field->bins) {
            if (bins && bins != c.field->bins) return fail("near-field masks must share their bins");
            bins = c.field->bins;
        }
    }
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
                    const uint32_t m = next.fetch_add(1);
                    if (m >= mesh_count) break;
                    const rq_cooked_mesh& c = meshes[m];
                    Blas& b = blas[m];
                    if (c.mesh) {  // set_indirect_geometry
                        const rq_mesh& src = *c.mesh;
                        if (!build_bvh_sah(src.positions, src.vertex_count, src.indices, src.index_count, b.bvh,
                                           errs[m])) {
                            errs[m] = "indirect geometry is not a valid triangle list: " + errs[m];
                            ok[m] = 0;
                            continue;
                        }
                        blas_bounds(b);
                        if (src.keep_primitives)
                            map_primitives(b.bvh.indices, src.indices, src.index_count, b.primitives);
                    } else if (c.bvh) {  // set_indirect_bvh
                        ok[m] = blas_from_cluster(c.bvh->b, c.index, b, errs[m]);
                    } else {  // set_indirect_near_field
                        if (c.index >= c.field->meshes.size()) {
                            errs[m] = "near-field mesh index";
                            ok[m] = 0;
                            continue;
                        }
                        if (!(ok[m] = blas_from_cluster(c.field->meshes[c.index], 0, b, errs[m]))) continue;
                        for (const NearPlacement& p : c.field->placements)
                            if (p.mesh == c.index && !p.masks.empty()) {
                                NearMasks n;
                                std::memcpy(n.model, p.transform, sizeof n.model);
                                n.masks = p.masks;
                                b.near.push_back(std::move(n));
                            }
                    }
                }
            });
        for (auto& th : pool) th.join();
    }
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (!ok[m]) return fail(errs[m]);
    return pack_scene(blas, instances, instance_count, top_level, bins, began, out);
}
