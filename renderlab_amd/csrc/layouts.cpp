// Device relayouts of a packed gi_geometry buffer. Both keep the reference's
// hit semantics: the same instance records (rows and world boxes bit for
// bit), the same top level, the same triangles in the same leaf order.
//
// PLAIN_F32: each BLAS node becomes a 32-word node of f32 child boxes, the
//   axis-major format the top level already uses (src/instance_bvh.rs:10-18);
//   each box is the unfused decode origin + step*q of the reference record
//   (shaders/gi_rays_software.wgsl:184-186), so every slab test sees the same
//   floats as the reference traversal. Triangles are gathered from the
//   index/position arrays into 36-byte records (vertex a and the f32 edges
//   b - a, c - a) so a leaf costs no dependent index load; the position
//   indices, BVH-order slot and primitive id move to a side array read once
//   per hit.
// CW8_Q8: 8-wide nodes with 8-bit child boxes on a power-of-two grid (after
//   Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs
//   Through Compressed Wide BVHs", HPG 2017), for the BLAS and the top level
//   alike. Each mesh's tree and the top level are built again from the
//   triangles and the instances' world boxes (binned SAH, then that paper's
//   cost-optimal collapse into 8-wide nodes), or, for words whose trees do not
//   hold each triangle or instance once, collapsed from the 4-wide trees. A
//   child's box encloses its subtree's triangles (or instance boxes), rounded
//   outward, and decodes exactly as origin + 2^e*q with no rounding. The
//   triangles are PLAIN's, reordered within their mesh so each leaf is a
//   contiguous range; the side records keep each triangle's slot and indices.
#include <algorithm>
#include <array>
#include <cmath>
#include <atomic>
#include <limits>
#include <map>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "internal.h"

namespace rq {
namespace {

constexpr uint32_t LEAF_COUNT_SHIFT = 27;   // leaf tag: LEAF | (n-1)<<27 | first triangle

// Exponent of the spacing of f32 values at |x| (ulp), -149 at zero.
int ulp_exp(float x) {
    if (x == 0.0f) return -149;
    int ex;
    std::frexp(std::fabs(x), &ex);
    return std::max(ex - 24, -149);
}
constexpr uint32_t MAX_LEAF_TRIS = 16;
constexpr uint32_t MAX_FIRST_TRI = 1u << 27;
constexpr uint32_t W8_WORDS = 24;

struct MeshMap {
    uint32_t nodes_at;  // key: the mesh's node word offset
    MeshView view;
    uint32_t node_base = 0;
    uint32_t tri_base = 0;
    uint32_t root = 0;
    bool prims = false;
    uint32_t prim_base = 0;
};

bool collect_meshes(const std::vector<uint32_t>& words, std::vector<MeshMap>& meshes,
                    std::vector<uint32_t>& instance_mesh, std::string& err) {
    uint32_t count = words[0];
    uint32_t total = std::max(words[0], words[4]);
    (void)total;
    instance_mesh.assign(count, 0);
    for (uint32_t i = 0; i < count; ++i) {
        MeshView v;
        if (!mesh_view(words, i, v, err)) return false;
        size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        uint32_t root = words[r + 22];
        uint32_t prim_base = words[r + 28];
        size_t m = 0;
        for (; m < meshes.size(); ++m)
            if (meshes[m].nodes_at == v.nodes && meshes[m].root == root) break;
        if (m == meshes.size()) {
            MeshMap mm;
            mm.nodes_at = v.nodes;
            mm.view = v;
            mm.root = root;
            meshes.push_back(mm);
        }
        if (prim_base) {
            meshes[m].prims = true;
            meshes[m].prim_base = prim_base;
        }
        instance_mesh[i] = (uint32_t)m;
    }
    return true;
}

uint32_t leaf_tag(uint32_t first_tri, uint32_t n, std::string& err) {
    if (n == 0 || n > MAX_LEAF_TRIS || first_tri >= MAX_FIRST_TRI) {
        err = "leaf too large for the relayout (more than 16 triangles or 2^27 triangles)";
        return EMPTY;
    }
    return LEAF_BIT | ((n - 1) << LEAF_COUNT_SHIFT) | first_tri;
}

void decode_unfused(const uint32_t* w, int slot, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        float o = bits_f(w[a]), s = bits_f(w[3 + a]);
        float ql = (float)((w[10 + a] >> (8 * slot)) & 0xffu);
        float qh = (float)((w[13 + a] >> (8 * slot)) & 0xffu);
        lo[a] = o + s * ql;
        hi[a] = o + s * qh;
    }
}

void gather_triangles(const std::vector<uint32_t>& words, const MeshMap& m, uint32_t slot_first,
                      uint32_t n, std::vector<float>& tris, std::vector<uint32_t>& ids, std::vector<float>& verts) {
    for (uint32_t k = slot_first; k < slot_first + n; ++k) {
        // Record (TRI_FLOATS): vertex a, then the edges b - a and c - a rounded to
        // f32 -- the same roundings the intersection test would make, done once.
        float p[3][3];
        for (int v = 0; v < 3; ++v) {
            uint32_t vi = words[m.view.indices + 3ull * k + v];
            for (int a = 0; a < 3; ++a) p[v][a] = bits_f(words[m.view.positions + 3ull * vi + a]);
            ids.push_back(vi);
        }
        for (int a = 0; a < 3; ++a) tris.push_back(p[0][a]);
        for (int e = 1; e < 3; ++e)
            for (int a = 0; a < 3; ++a) tris.push_back(p[e][a] - p[0][a]);
        // The vertices themselves, for the watertight test: a vertex shared by two
        // triangles must be the same floats in both (a + (b - a) need not be b).
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) verts.push_back(p[v][a]);
        // Side record (TRI_ID_WORDS), read once per hit: the three position
        // indices above, the BVH-order slot, the primitive id, zero.
        ids.push_back(k);
        ids.push_back(m.prims ? words[m.prim_base + k] : 0u);
        ids.push_back(0u);
    }
}

void instance_records(const std::vector<uint32_t>& words, const std::vector<MeshMap>& meshes,
                      const std::vector<uint32_t>& instance_mesh, const std::vector<uint32_t>& roots,
                      std::vector<uint32_t>& inst) {
    uint32_t count = words[0];
    inst.assign((size_t)count * INSTANCE_WORDS, 0u);
    for (uint32_t i = 0; i < count; ++i) {
        size_t r = HEADER_WORDS + (size_t)i * INSTANCE_WORDS;
        uint32_t* o = &inst[(size_t)i * INSTANCE_WORDS];
        for (int k = 0; k < 18; ++k) o[k] = words[r + k];
        const MeshMap& m = meshes[instance_mesh[i]];
        // 18 root tag, 19 flags (bit 0 casts no shadow, bit 1 primitive ids),
        // 20 the mesh's first triangle, 21 mesh, 22 the reference record index.
        o[18] = roots[instance_mesh[i]];
        o[19] = (words[r + 27] & 1u) | (words[r + 28] ? 2u : 0u);
        o[20] = m.tri_base;
        o[21] = instance_mesh[i];
        o[22] = i;
    }
}

void copy_top(const std::vector<uint32_t>& words, HostLayout& out) {
    uint32_t count = words[0];
    out.instance_count = count;
    out.top_on = words[1] != 0;
    out.top_root = words[2];
    out.top.clear();
    if (out.top_on) {
        size_t base = words[1];
        size_t n = std::min<size_t>(top_reserve(count), words.size() - std::min(words.size(), base));
        out.top.assign(words.begin() + base, words.begin() + base + n);
    }
    if (out.top.empty()) out.top.assign(TOP_WORDS, 0u);
}

}  // namespace

bool build_plain(const std::vector<uint32_t>& words, HostLayout& out, std::string& err) {
    out = HostLayout();
    std::vector<MeshMap> meshes;
    std::vector<uint32_t> instance_mesh;
    if (!collect_meshes(words, meshes, instance_mesh, err)) return false;
    uint32_t node_total = 0, tri_total = 0;
    for (auto& m : meshes) {
        m.node_base = node_total;
        m.tri_base = tri_total;
        node_total += m.view.node_count;
        tri_total += m.view.index_count / 3;
    }
    if (node_total >= LEAF_BIT || tri_total >= MAX_FIRST_TRI) {
        err = "scene too large for the relayout";
        return false;
    }
    out.nodes.assign((size_t)node_total * TOP_WORDS, 0u);
    out.tris.reserve((size_t)tri_total * TRI_FLOATS);
    out.tri_ids.reserve((size_t)tri_total * TRI_ID_WORDS);
    out.verts.reserve((size_t)tri_total * TRI_FLOATS);
    std::vector<uint32_t> roots(meshes.size());
    for (size_t mi = 0; mi < meshes.size(); ++mi) {
        const MeshMap& m = meshes[mi];
        auto tag = [&](uint32_t child) -> uint32_t {
            if (child == EMPTY) return EMPTY;
            if (child & LEAF_BIT) {
                uint32_t l = child & ~LEAF_BIT;
                if (l >= m.view.leaf_count) return EMPTY;
                uint32_t first = words[m.view.leaves + 2ull * l], cnt = words[m.view.leaves + 2ull * l + 1];
                if ((uint64_t)first / 3 + cnt / 3 > m.view.index_count / 3) {
                    err = "a leaf's triangles past the mesh's index count";
                    return EMPTY;
                }
                return leaf_tag(m.tri_base + first / 3, cnt / 3, err);
            }
            if (child >= m.view.node_count) return EMPTY;
            return m.node_base + child;
        };
        for (uint32_t n = 0; n < m.view.node_count; ++n) {
            const uint32_t* w = &words[m.view.nodes + (size_t)n * NODE_WORDS];
            uint32_t* o = &out.nodes[(size_t)(m.node_base + n) * TOP_WORDS];
            for (int s = 0; s < 4; ++s) {
                uint32_t c = w[6 + s];
                if (c == EMPTY) {
                    o[24 + s] = EMPTY;
                    continue;
                }
                float lo[3], hi[3];
                decode_unfused(w, s, lo, hi);
                for (int a = 0; a < 3; ++a) {
                    o[a * 4 + s] = f_bits(lo[a]);
                    o[12 + a * 4 + s] = f_bits(hi[a]);
                }
                o[24 + s] = tag(c);
                if (o[24 + s] == EMPTY) {
                    if (err.empty()) err = "node child out of range";
                    return false;
                }
            }
        }
        roots[mi] = tag(m.root);
        if (roots[mi] == EMPTY) {
            if (err.empty()) err = "mesh root out of range";
            return false;
        }
        gather_triangles(words, m, 0, m.view.index_count / 3, out.tris, out.tri_ids, out.verts);
    }
    instance_records(words, meshes, instance_mesh, roots, out.inst);
    copy_top(words, out);
    for (const MeshMap& m : meshes) {
        out.mesh_base.push_back(m.nodes_at);
        out.mesh_base.push_back(m.node_base);
        out.mesh_base.push_back(m.tri_base);
    }
    return true;
}

namespace {

struct Cand {
    uint32_t child;  // tag in the source tree: node index or leaf tag
    Bounds box;
};

float half_area(const Bounds& b) {
    float e[3];
    for (int a = 0; a < 3; ++a) e[a] = mx(b.hi[a] - b.lo[a], 0.0f);
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

// Children of a 32-word 4-wide f32 node (the PLAIN BLAS node and the top-level
// node share this format).
std::vector<Cand> wide_children(const std::vector<uint32_t>& nodes, uint32_t node) {
    std::vector<Cand> c;
    const uint32_t* w = &nodes[(size_t)node * TOP_WORDS];
    for (int s = 0; s < 4; ++s) {
        if (w[24 + s] == EMPTY) continue;
        Cand k;
        k.child = w[24 + s];
        for (int a = 0; a < 3; ++a) {
            k.box.lo[a] = bits_f(w[a * 4 + s]);
            k.box.hi[a] = bits_f(w[12 + a * 4 + s]);
        }
        c.push_back(k);
    }
    return c;
}

// Collapses a 4-wide tree into 8-wide nodes of 24 words (96 B):
//   0-2 origin (f32), 3 exponents x|y<<8|z<<16 (signed bytes) and valid-slot mask <<24,
//   4-15 byte planes (lo x, lo y, lo z, hi x, hi y, hi z; 8 codes, 2 words each),
//   16-23 child tags (node index, leaf tag unchanged from the source); an empty
//   slot holds an inverted box and a copy of the first child's tag.
// A box decodes as origin + 2^e*q, which is a representable f32 for every
// code (the origin is aligned and the step chosen for that).
struct W8Builder {
    const std::vector<uint32_t>* src;
    std::vector<uint32_t>* out;
    std::string* err;
    // GI mesh trees: the CW8 copy of the triangle records, whose leaves may be split
    // (null for the top level and for reflection trees, whose leaf tags differ)
    HostLayout* tris = nullptr;
    // Trees built from the triangles: each source node is a list of at most 8
    // children (leaf tags, or indices of further lists), taken as they are.
    const std::vector<std::vector<Cand>>* lists = nullptr;
    std::vector<Cand> cands_by_slot_;
    std::vector<char> slots_used_;

    // Exact bounds of triangles [first, first + n) of the CW8 copy (their vertices).
    Bounds tri_bounds(uint32_t first, uint32_t n) const {
        Bounds b = empty_bounds();
        for (uint32_t k = first; k < first + n; ++k)
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    const float x = tris->verts[(size_t)TRI_FLOATS * k + 3 * v + a];
                    b.lo[a] = mn(b.lo[a], x);
                    b.hi[a] = mx(b.hi[a], x);
                }
        return b;
    }

    // An 8-wide node tests all 8 slots whatever it holds, so free slots cost nothing
    // more when filled. While slots are free, the leaf with the most triangles (3 or
    // more) is split in two at the median of its triangles' centroids along their
    // longest axis, each part with the exact box of its vertices: a ray then tests
    // only the part it meets. The triangles move within the leaf's range of this
    // layout's own triangle records (the side record keeps each triangle's slot and
    // indices for the hit record). The first part has an even count, so a ray that
    // meets both parts takes no more pair steps than it did. Within an instance a
    // closest hit can then differ only between equal-distance triangles.
    void split_leaves(std::vector<Cand>& cands) {
        while (cands.size() < 8) {
            int pick = -1;
            uint32_t best = 2;
            for (int i = 0; i < (int)cands.size(); ++i) {
                const uint32_t t = cands[i].child;
                if (!(t & LEAF_BIT)) continue;
                const uint32_t n = ((t >> 27) & 0xfu) + 1u;
                if (n > best || (n == best && pick >= 0 && half_area(cands[i].box) > half_area(cands[pick].box))) {
                    pick = i;
                    best = n;
                }
            }
            if (pick < 0) break;
            const uint32_t tag = cands[pick].child, first = tag & ((1u << 27) - 1u), n = ((tag >> 27) & 0xfu) + 1u;
            if ((size_t)first + n > tris->verts.size() / TRI_FLOATS) break;  // the relayout checked every leaf
            std::vector<std::array<float, 3>> ctr(n);
            Bounds cb = empty_bounds();
            for (uint32_t k = 0; k < n; ++k) {
                const float* p = &tris->verts[(size_t)TRI_FLOATS * (first + k)];
                for (int a = 0; a < 3; ++a) {
                    ctr[k][a] = (p[a] + p[3 + a] + p[6 + a]) / 3.0f;
                    cb.lo[a] = mn(cb.lo[a], ctr[k][a]);
                    cb.hi[a] = mx(cb.hi[a], ctr[k][a]);
                }
            }
            int axis = 0;
            for (int a = 1; a < 3; ++a)
                if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
            std::vector<uint32_t> order(n);
            for (uint32_t k = 0; k < n; ++k) order[k] = k;
            std::stable_sort(order.begin(), order.end(),
                             [&](uint32_t x, uint32_t y) { return ctr[x][axis] < ctr[y][axis]; });
            auto permute = [&](auto& v, size_t width) {
                std::vector<typename std::decay_t<decltype(v)>::value_type> tmp(v.begin() + width * first,
                                                                                v.begin() + width * (first + n));
                for (uint32_t k = 0; k < n; ++k)
                    std::copy(tmp.begin() + width * order[k], tmp.begin() + width * (order[k] + 1),
                              v.begin() + width * (first + k));
            };
            permute(tris->tris, TRI_FLOATS);
            permute(tris->tri_ids, TRI_ID_WORDS);
            permute(tris->verts, TRI_FLOATS);
            const uint32_t m = 2u * ((n + 2u) / 4u);  // even, about half: 3 -> 2+1, 5 -> 2+3, 6 -> 4+2, 8 -> 4+4
            Cand a, b;
            a.child = LEAF_BIT | ((m - 1u) << 27) | first;
            a.box = tri_bounds(first, m);
            b.child = LEAF_BIT | ((n - m - 1u) << 27) | (first + m);
            b.box = tri_bounds(first + m, n - m);
            cands[pick] = a;
            cands.insert(cands.begin() + pick + 1, b);
        }
    }

    uint32_t alloc() {
        uint32_t index = (uint32_t)(out->size() / W8_WORDS);
        out->resize(out->size() + W8_WORDS, 0u);
        return index;
    }

    std::vector<Cand> source_children(uint32_t node) const {
        return lists ? (*lists)[node] : wide_children(*src, node);
    }

    bool fill(uint32_t index, std::vector<Cand> cands) {
        // Open the largest inner candidate while its children fit in 8 slots.
        for (; !lists;) {
            int pick = -1;
            for (int i = 0; i < (int)cands.size(); ++i) {
                if (cands[i].child & LEAF_BIT) continue;
                if (cands.size() - 1 + wide_children(*src, cands[i].child).size() > 8) continue;
                if (pick < 0 || half_area(cands[i].box) > half_area(cands[pick].box)) pick = i;
            }
            if (pick < 0) break;
            auto kids = wide_children(*src, cands[pick].child);
            cands.erase(cands.begin() + pick);
            cands.insert(cands.begin() + pick, kids.begin(), kids.end());
        }
        if (tris && !lists) split_leaves(cands);
        Bounds all = empty_bounds();
        for (auto& c : cands) all = merge(all, c.box);
        // Slot assignment by ray octant (Ylitie et al. 2017, sec. 3.2): slot s
        // takes the child that comes first for rays whose direction signs are
        // the bits of s, so visiting slots k ^ octant for k = 0..7 runs near to
        // far without sorting. Greedy over (cost, child, slot) ascending; cost
        // is the child centre along the octant's diagonal.
        {
            std::vector<Cand> slots(8);
            std::vector<char> used(8, 0), placed(cands.size(), 0);
            std::vector<std::tuple<float, int, int>> costs;
            for (size_t c = 0; c < cands.size(); ++c)
                for (int sl = 0; sl < 8; ++sl) {
                    float cost = 0.0f;
                    for (int a = 0; a < 3; ++a) {
                        float ctr = (cands[c].box.lo[a] + cands[c].box.hi[a]) * 0.5f -
                                    (all.lo[a] + all.hi[a]) * 0.5f;
                        cost += ((sl >> a) & 1) ? -ctr : ctr;
                    }
                    costs.emplace_back(cost, (int)c, sl);
                }
            std::stable_sort(costs.begin(), costs.end(),
                             [](const auto& x, const auto& y) { return std::get<0>(x) < std::get<0>(y); });
            std::vector<int> slot_of(cands.size(), -1);
            for (auto& [cost, c, sl] : costs) {
                (void)cost;
                if (placed[c] || used[sl]) continue;
                placed[c] = 1;
                used[sl] = 1;
                slot_of[c] = sl;
            }
            std::vector<Cand> ordered(8);
            std::vector<char> occupied(8, 0);
            for (size_t c = 0; c < cands.size(); ++c) {
                ordered[slot_of[c]] = cands[c];
                occupied[slot_of[c]] = 1;
            }
            slots_used_.assign(occupied.begin(), occupied.end());
            cands_by_slot_ = ordered;
        }
        uint32_t w[W8_WORDS] = {};
        int exps[3];
        for (int a = 0; a < 3; ++a) {
            const float lo = all.lo[a], hi = all.hi[a];
            if (!std::isfinite(lo) || !std::isfinite(hi)) {
                *err = "box too large to quantize";
                return false;
            }
            int e = -126;
            if (hi - lo > 0.0f) {
                int ex;
                std::frexp((hi - lo) / 255.0f, &ex);
                e = std::max(ex - 1, -126);
            }
            // Every code must decode exactly: the origin sits on a grid of 2^ue,
            // where 2^ue is the largest ulp over [origin, origin + 255 * 2^e],
            // and the step 2^e is at least 2^ue. Then origin + q * 2^e is a
            // representable multiple of 2^ue for every q.
            int ue = -149;
            float p = lo;
            bool ok = false;
            for (int guard = 0; guard < 1024 && e <= 127; ++guard) {
                p = ue <= ulp_exp(lo) ? lo : std::floor(lo / std::ldexp(1.0f, ue)) * std::ldexp(1.0f, ue);
                const float top = p + std::ldexp(1.0f, e) * 255.0f;
                const int need = std::max(ulp_exp(p), ulp_exp(top));
                if (!std::isfinite(top)) break;
                if (need > ue) { ue = need; continue; }
                if (e < ue) { e = ue; continue; }
                if (top < hi) { ++e; continue; }
                ok = true;
                break;
            }
            if (!ok) {
                *err = "box too large to quantize";
                return false;
            }
            exps[a] = e;
            w[a] = f_bits(p);
        }
        uint32_t valid = 0;
        std::vector<std::pair<uint32_t, uint32_t>> todo;  // (new index, source node)
        const std::vector<Cand> by_slot = cands_by_slot_;
        const std::vector<char> occupied = slots_used_;
        for (size_t s = 0; s < 8; ++s) {
            if (!occupied[s]) continue;
            const Cand& c = by_slot[s];
            valid |= 1u << s;
            for (int a = 0; a < 3; ++a) {
                float p = bits_f(w[a]);
                float sc = std::ldexp(1.0f, exps[a]);
                uint32_t ql = (uint32_t)mn(mx(std::floor((c.box.lo[a] - p) / sc), 0.0f), 255.0f);
                while (ql > 0 && p + sc * (float)ql > c.box.lo[a]) ql--;
                uint32_t qh = (uint32_t)mn(mx(std::ceil((c.box.hi[a] - p) / sc), 0.0f), 255.0f);
                while (qh < 255 && p + sc * (float)qh < c.box.hi[a]) qh++;
                if (p + sc * (float)ql > c.box.lo[a] || p + sc * (float)qh < c.box.hi[a]) {
                    *err = "quantization failed to enclose";
                    return false;
                }
                w[4 + 2 * a + s / 4] |= ql << (8 * (s % 4));
                w[10 + 2 * a + s / 4] |= qh << (8 * (s % 4));
            }
            if (c.child & LEAF_BIT) {
                w[16 + s] = c.child;
            } else {
                uint32_t ni = alloc();
                w[16 + s] = ni;
                todo.push_back({ni, c.child});
            }
        }
        // Empty slots: an inverted box (low codes 255, high codes 0) that fails the
        // traversal's slab test but for rays whose origin is about 2^19 node extents away
        // (the culling slack then exceeds the node), and a copy of the first child's tag -- if such a ray enters it,
        // it visits that child twice, which costs time and changes no result. The
        // traversal then needs no per-slot valid test; the valid mask stays in word 3.
        uint32_t dup = EMPTY;
        for (size_t s = 0; s < 8 && dup == EMPTY; ++s)
            if (occupied[s]) dup = w[16 + s];
        for (size_t s = 0; s < 8; ++s)
            if (!occupied[s]) {
                w[16 + s] = dup;
                for (int a = 0; a < 3; ++a) w[4 + 2 * a + s / 4] |= 0xffu << (8 * (s % 4));
            }
        w[3] = (uint32_t)(uint8_t)(int8_t)exps[0] | ((uint32_t)(uint8_t)(int8_t)exps[1] << 8) |
               ((uint32_t)(uint8_t)(int8_t)exps[2] << 16) | (valid << 24);
        std::copy(w, w + W8_WORDS, out->begin() + (size_t)index * W8_WORDS);
        for (auto& t : todo)
            if (!fill(t.first, source_children(t.second))) return false;
        return true;
    }

    // The 8-wide tree under a source root tag; returns the new root tag.
    bool root(uint32_t src_root, uint32_t& out_root) {
        if (src_root & LEAF_BIT) {
            out_root = src_root;
            return true;
        }
        uint32_t index = alloc();
        out_root = index;
        return fill(index, source_children(src_root));
    }
};

// ---- CW8 mesh trees built from the triangles
// The 4-wide trees of the words are the reference builder's; the CW8 layout
// needs only their triangles. A mesh's CW8 tree is built again from them: a
// binary tree by SAH over the triangles' box centres (32 bins on every axis;
// every split position of the sorted centres for 256 triangles or fewer), then
// collapsed into nodes of at most 8 children by the dynamic program
// of Ylitie, Karras and Laine (HPG 2017, sec. 3.1), which finds the collapse of
// least SAH cost. C(n, i) is the cost of binary node n spread over at most i
// child slots:
//   C(n, 1) = min(leaf: A(n) * C_PAIR * ceil(T(n) / 2) when T(n) <= LEAF_MAX,
//                 node: A(n) * C_NODE + D(n, 8))
//   C(n, i) = min(D(n, i), C(n, i - 1)),   D(n, i) = min_k C(l, k) + C(r, i - k)
// with A the box's half area and T the triangle count. Leaves are priced per
// triangle-pair step (the traversal tests two triangles a step); a node step
// costs about twice a pair step on the kernel's measured instruction counts.
// The mesh's triangles are reordered so each leaf is a contiguous range.
constexpr int SAH_BINS = 32;
constexpr uint32_t SAH_SWEEP_MAX = 256;  // smaller nodes: every split position (1.7% on the 4M frame)
constexpr uint32_t SAH_DEPTH = 64;     // deeper splits take the median (bounds the depth)
constexpr uint32_t DP_LEAF_MAX = 8;
// C_PAIR = C_NODE: a pair step and a node step each take a wave iteration; the
// measured optimum on the 4M-triangle frame was flat over C_PAIR in [1, 3].
// Top-level leaves are instances (one a leaf), priced like a node step (the
// kernel enters an instance and visits its root in one step).
constexpr double DP_C_NODE = 1.0, DP_C_PAIR = 1.0, DP_C_INST = 1.0;

struct BinNode {
    Bounds box;
    uint32_t left = 0, right = 0;   // 0: a single triangle
    uint32_t first = 0, count = 0;  // triangles order[first, first + count)
};

double area2(const Bounds& b) {
    double e[3];
    for (int a = 0; a < 3; ++a) e[a] = std::max((double)b.hi[a] - (double)b.lo[a], 0.0);
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

// Bounds of triangles [0, n) of `verts` (TRI_FLOATS a triangle).
std::vector<Bounds> triangle_boxes(const float* verts, uint32_t n) {
    std::vector<Bounds> tb(n);
    for (uint32_t k = 0; k < n; ++k) {
        Bounds b = empty_bounds();
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) {
                const float x = verts[(size_t)TRI_FLOATS * k + 3 * v + a];
                b.lo[a] = mn(b.lo[a], x);
                b.hi[a] = mx(b.hi[a], x);
            }
        tb[k] = b;
    }
    return tb;
}

// Binary tree over primitives with boxes `tb` (finite); `order` receives the
// primitive order its nodes' ranges refer to. Node 0 is the root; children
// always follow their parent.
std::vector<BinNode> sah_binary(const std::vector<Bounds>& tb, std::vector<uint32_t>& order) {
    const uint32_t n = (uint32_t)tb.size();
    std::vector<std::array<float, 3>> c(n);
    for (uint32_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) c[k][a] = 0.5f * tb[k].lo[a] + 0.5f * tb[k].hi[a];
    order.resize(n);
    for (uint32_t k = 0; k < n; ++k) order[k] = k;
    std::vector<BinNode> nodes(1);
    nodes.reserve(2 * (size_t)n);
    nodes[0].count = n;
    std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, 0u}};  // (node, depth)
    while (!todo.empty()) {
        const auto [x, depth] = todo.back();
        todo.pop_back();
        const uint32_t first = nodes[x].first, cnt = nodes[x].count;
        Bounds box = empty_bounds(), cb = empty_bounds();
        for (uint32_t k = first; k < first + cnt; ++k) {
            box = merge(box, tb[order[k]]);
            for (int a = 0; a < 3; ++a) {
                cb.lo[a] = mn(cb.lo[a], c[order[k]][a]);
                cb.hi[a] = mx(cb.hi[a], c[order[k]][a]);
            }
        }
        nodes[x].box = box;
        if (cnt == 1) continue;
        int axis = -1, split = 0;
        uint32_t sweep_mid = 0;
        if (depth < SAH_DEPTH && cnt <= SAH_SWEEP_MAX) {
            // every split position of the primitives sorted by centre, on each axis
            double best = std::numeric_limits<double>::infinity();
            std::vector<uint32_t> tmp(order.begin() + first, order.begin() + first + cnt), best_order;
            std::vector<double> right(cnt);
            for (int a = 0; a < 3; ++a) {
                std::sort(tmp.begin(), tmp.end(), [&](uint32_t p, uint32_t q) {
                    return c[p][a] < c[q][a] || (c[p][a] == c[q][a] && p < q);
                });
                Bounds acc = empty_bounds();
                for (uint32_t k = cnt; k-- > 1;) {
                    acc = merge(acc, tb[tmp[k]]);
                    right[k] = area2(acc) * (cnt - k);
                }
                acc = empty_bounds();
                bool better = false;
                for (uint32_t k = 1; k < cnt; ++k) {
                    acc = merge(acc, tb[tmp[k - 1]]);
                    const double cost = area2(acc) * k + right[k];
                    if (cost < best) {
                        best = cost;
                        axis = a;
                        sweep_mid = k;
                        better = true;
                    }
                }
                if (better) best_order = tmp;
            }
            if (axis >= 0) std::copy(best_order.begin(), best_order.end(), order.begin() + first);
        } else if (depth < SAH_DEPTH) {
            double best = std::numeric_limits<double>::infinity();
            for (int a = 0; a < 3; ++a) {
                const double lo = cb.lo[a], ext = (double)cb.hi[a] - lo;
                if (!(ext > 0.0)) continue;
                const double scale = SAH_BINS / ext;
                Bounds bb[SAH_BINS];
                uint32_t bn[SAH_BINS] = {};
                for (auto& b : bb) b = empty_bounds();
                for (uint32_t k = first; k < first + cnt; ++k) {
                    const int i = std::min(SAH_BINS - 1, (int)(((double)c[order[k]][a] - lo) * scale));
                    bb[i] = merge(bb[i], tb[order[k]]);
                    bn[i]++;
                }
                double right_area[SAH_BINS];
                uint32_t right_n[SAH_BINS];
                Bounds acc = empty_bounds();
                uint32_t m = 0;
                for (int i = SAH_BINS - 1; i > 0; --i) {
                    acc = merge(acc, bb[i]);
                    m += bn[i];
                    right_area[i] = area2(acc);
                    right_n[i] = m;
                }
                acc = empty_bounds();
                m = 0;
                for (int i = 1; i < SAH_BINS; ++i) {
                    acc = merge(acc, bb[i - 1]);
                    m += bn[i - 1];
                    if (m == 0 || right_n[i] == 0) continue;
                    const double cost = area2(acc) * m + right_area[i] * right_n[i];
                    if (cost < best) {
                        best = cost;
                        axis = a;
                        split = i;
                    }
                }
            }
        }
        uint32_t mid;
        if (axis >= 0 && cnt <= SAH_SWEEP_MAX) {
            mid = sweep_mid;
        } else if (axis >= 0) {
            const double lo = cb.lo[axis], scale = SAH_BINS / ((double)cb.hi[axis] - lo);
            auto it = std::partition(order.begin() + first, order.begin() + first + cnt, [&](uint32_t k) {
                return std::min(SAH_BINS - 1, (int)(((double)c[k][axis] - lo) * scale)) < split;
            });
            mid = (uint32_t)(it - (order.begin() + first));
        } else {
            // no usable SAH split (equal centres, or past SAH_DEPTH): the median along the widest axis
            int a = 0;
            for (int b = 1; b < 3; ++b)
                if (cb.hi[b] - cb.lo[b] > cb.hi[a] - cb.lo[a]) a = b;
            mid = cnt / 2;
            std::stable_sort(order.begin() + first, order.begin() + first + cnt,
                             [&](uint32_t p, uint32_t q) { return c[p][a] < c[q][a]; });
        }
        const uint32_t l = (uint32_t)nodes.size();
        nodes.resize(nodes.size() + 2);
        nodes[l].first = first;
        nodes[l].count = mid;
        nodes[l + 1].first = first + mid;
        nodes[l + 1].count = cnt - mid;
        nodes[x].left = l;
        nodes[x].right = l + 1;
        todo.push_back({l + 1, depth + 1});
        todo.push_back({l, depth + 1});
    }
    return nodes;
}

// The collapse of a binary tree into lists of at most 8 children (leaf tags from
// `leaf_tag` for binary nodes taken whole as leaves, of at most `leaf_max`
// primitives priced `leaf_price` (count) each per unit of area; list indices for
// nodes). Returns the root tag: a leaf tag, or list 0.
template <typename Tag, typename Price>
uint32_t dp_collapse(const std::vector<BinNode>& nodes, uint32_t leaf_max, Price leaf_price, Tag leaf_tag,
                     std::vector<std::vector<Cand>>& lists) {
    struct Entry {
        float cost[8];
        int8_t how[8];  // i = 1: -1 leaf, -2 node; i > 1: -3 as i - 1, else k slots to the left child
        int8_t k8;      // the node's own split of 8 slots
    };
    std::vector<Entry> dp(nodes.size());
    // areas relative to the root's (boxes of any finite size keep finite costs)
    const double root_area = area2(nodes[0].box), scale = root_area > 0.0 ? 1.0 / root_area : 1.0;
    for (size_t x = nodes.size(); x-- > 0;) {
        const BinNode& nd = nodes[x];
        Entry& e = dp[x];
        const double a = std::min(area2(nd.box) * scale, 1.0);
        if (nd.left == 0) {
            for (int i = 0; i < 8; ++i) {
                e.cost[i] = (float)(a * leaf_price(1u));
                e.how[i] = i == 0 ? -1 : -3;
            }
            e.k8 = 0;
            continue;
        }
        const Entry &l = dp[nd.left], &r = dp[nd.right];
        double d[9];
        int dk[9];
        for (int j = 2; j <= 8; ++j) {
            d[j] = std::numeric_limits<double>::infinity();
            dk[j] = 1;
            for (int k = 1; k < j; ++k) {
                const double s = (double)l.cost[k - 1] + (double)r.cost[j - k - 1];
                if (s < d[j]) {
                    d[j] = s;
                    dk[j] = k;
                }
            }
        }
        const double as_node = a * DP_C_NODE + d[8];
        const bool leaf = nd.count <= leaf_max && a * leaf_price(nd.count) <= as_node;
        e.cost[0] = (float)(leaf ? a * leaf_price(nd.count) : as_node);
        e.how[0] = leaf ? -1 : -2;
        e.k8 = (int8_t)dk[8];
        for (int i = 2; i <= 8; ++i) {
            if (d[i] < (double)e.cost[i - 2]) {
                e.cost[i - 1] = (float)d[i];
                e.how[i - 1] = (int8_t)dk[i];
            } else {
                e.cost[i - 1] = e.cost[i - 2];
                e.how[i - 1] = -3;
            }
        }
    }
    if (dp[0].how[0] == -1) return leaf_tag(nodes[0]);
    // the binary nodes that fill node x's slots: its children's subtrees spread over k8 and 8 - k8
    std::vector<uint32_t> slots;
    auto gather = [&](auto&& self, uint32_t x, int i) -> void {
        while (i > 1 && dp[x].how[i - 1] == -3) --i;
        if (i == 1) {
            slots.push_back(x);
            return;
        }
        const int k = dp[x].how[i - 1];
        self(self, nodes[x].left, k);
        self(self, nodes[x].right, i - k);
    };
    std::vector<uint32_t> todo{0u}, list_of{0u};  // binary node -> its list
    lists.assign(1, {});
    for (size_t q = 0; q < todo.size(); ++q) {
        const uint32_t x = todo[q], li = list_of[q];
        slots.clear();
        gather(gather, nodes[x].left, dp[x].k8);
        gather(gather, nodes[x].right, 8 - dp[x].k8);
        std::vector<Cand> kids;
        for (uint32_t s : slots) {
            Cand c;
            c.box = nodes[s].box;
            if (dp[s].how[0] == -1) {
                c.child = leaf_tag(nodes[s]);
            } else {
                c.child = (uint32_t)lists.size();
                lists.emplace_back();
                todo.push_back(s);
                list_of.push_back(c.child);
            }
            kids.push_back(c);
        }
        lists[li] = std::move(kids);
    }
    return 0u;
}

// Levels of the list tree under list 0.
uint32_t list_levels(const std::vector<std::vector<Cand>>& lists) {
    std::vector<uint32_t> level(lists.size(), 1u);
    uint32_t deepest = 1;
    for (size_t i = 0; i < lists.size(); ++i)  // a list's children come after it
        for (const Cand& c : lists[i])
            if (!(c.child & LEAF_BIT)) {
                level[c.child] = level[i] + 1;
                deepest = std::max(deepest, level[c.child]);
            }
    return deepest;
}

// Levels of the 8-wide tree under `tag` (0 for a leaf tag), counted without
// recursion; stops once past `limit`.
uint32_t w8_levels(const std::vector<uint32_t>& nodes, uint32_t tag, uint32_t limit) {
    if (tag & LEAF_BIT) return 0;
    uint32_t deepest = 0;
    std::vector<std::pair<uint32_t, uint32_t>> todo{{tag, 1u}};
    while (!todo.empty()) {
        const auto [node, level] = todo.back();
        todo.pop_back();
        deepest = std::max(deepest, level);
        if (deepest > limit) break;
        for (size_t s = 0; s < 8; ++s) {
            const uint32_t c = nodes[(size_t)node * W8_WORDS + 16 + s];
            if (!(c & LEAF_BIT)) todo.push_back({c, level + 1});
        }
    }
    return deepest;
}

// The traversal kernel bounds a level's stack by 7 entries a tree level for
// at most this many levels (it does not check pushes against the bound).
constexpr uint32_t W8_MAX_LEVELS = 16;

// Levels of a 4-wide f32 tree (32-word nodes, children in words 24-27) under
// `tag`, counted without recursion; stops once past `limit` (or at a node past
// the array). An 8-wide node spans at most 7 levels of a tree whose nodes have
// two or more children (each opened child adds a slot), so such a source deeper
// than 7 * W8_MAX_LEVELS cannot collapse within the bound. Checked before the
// collapse, which recurses once per 8-wide level: that recursion then stays
// shallow whatever the input.
uint32_t wide_levels(const std::vector<uint32_t>& nodes, uint32_t tag, uint32_t limit) {
    if (tag & LEAF_BIT) return 0;
    uint32_t deepest = 0;
    std::vector<std::pair<uint32_t, uint32_t>> todo{{tag, 1u}};
    while (!todo.empty()) {
        const auto [node, level] = todo.back();
        todo.pop_back();
        deepest = std::max(deepest, level);
        if (deepest > limit || ((size_t)node + 1) * TOP_WORDS > nodes.size()) return limit + 1;
        for (size_t s = 0; s < 4; ++s) {
            const uint32_t c = nodes[(size_t)node * TOP_WORDS + 24 + s];
            if (!(c & LEAF_BIT)) todo.push_back({c, level + 1});
        }
    }
    return deepest;
}

// The CW8 tree of one mesh built from its triangles [tb, tb + n) of `out`, whose
// records it reorders, into `nodes` (indices from 0). False with `err` empty when
// the mesh's 4-wide tree (root tag `pr` in `plain`) does not hold each of those
// triangles exactly once, a vertex is not finite, or the tree would pass
// W8_MAX_LEVELS: the caller collapses that tree instead.
bool rebuild_mesh(const HostLayout& plain, uint32_t pr, uint32_t tb, uint32_t n, HostLayout& out,
                  std::vector<uint32_t>& nodes, uint32_t& root, std::string& err) {
    if (n == 0 || (size_t)tb + n > out.verts.size() / TRI_FLOATS) return false;
    for (size_t k = (size_t)TRI_FLOATS * tb; k < (size_t)TRI_FLOATS * (tb + n); ++k)
        if (!std::isfinite(out.verts[k])) return false;
    std::vector<uint8_t> seen(n, 0);
    std::vector<uint32_t> todo{pr};
    while (!todo.empty()) {
        const uint32_t t = todo.back();
        todo.pop_back();
        if (t & LEAF_BIT) {
            const uint32_t first = t & ((1u << LEAF_COUNT_SHIFT) - 1u), cnt = ((t >> LEAF_COUNT_SHIFT) & 0xfu) + 1u;
            for (uint32_t k = first; k < first + cnt; ++k) {
                if (k < tb || k - tb >= n || seen[k - tb]) return false;
                seen[k - tb] = 1;
            }
            continue;
        }
        if (((size_t)t + 1) * TOP_WORDS > plain.nodes.size()) return false;
        for (int s = 0; s < 4; ++s) {
            const uint32_t c = plain.nodes[(size_t)t * TOP_WORDS + 24 + s];
            if (c != EMPTY) todo.push_back(c);
        }
        if (todo.size() > 4 * (size_t)n + 64) return false;  // not a tree over these triangles
    }
    for (uint8_t s : seen)
        if (!s) return false;
    std::vector<uint32_t> order;
    const std::vector<BinNode> bin = sah_binary(triangle_boxes(&out.verts[(size_t)TRI_FLOATS * tb], n), order);
    std::vector<std::vector<Cand>> lists;
    const uint32_t rt = dp_collapse(
        bin, DP_LEAF_MAX, [](uint32_t cnt) { return DP_C_PAIR * ((cnt + 1) / 2); },
        [&](const BinNode& x) { return LEAF_BIT | ((x.count - 1u) << LEAF_COUNT_SHIFT) | (tb + x.first); }, lists);
    if (rt & LEAF_BIT) {
        root = rt;
    } else {
        if (list_levels(lists) > W8_MAX_LEVELS) return false;
        W8Builder w{nullptr, &nodes, &err};
        w.lists = &lists;
        if (!w.root(rt, root)) return false;
    }
    // triangle position tb + p now holds the triangle that was at tb + order[p]
    auto permute = [&](auto& v, size_t width) {
        std::vector<typename std::decay_t<decltype(v)>::value_type> tmp(v.begin() + width * tb,
                                                                        v.begin() + width * (tb + n));
        for (uint32_t p = 0; p < n; ++p)
            std::copy(tmp.begin() + width * order[p], tmp.begin() + width * (order[p] + 1),
                      v.begin() + width * (tb + p));
    };
    permute(out.tris, TRI_FLOATS);
    permute(out.tri_ids, TRI_ID_WORDS);
    permute(out.verts, TRI_FLOATS);
    return true;
}

// A CW8 top level over instance boxes, one instance a leaf (LEAF | instance):
// binned SAH, then the cost-optimal 8-wide collapse. False with `err` empty when
// the tree would pass W8_MAX_LEVELS.
bool top_from_boxes(const std::vector<Bounds>& boxes, std::vector<uint32_t>& top, uint32_t& root, std::string& err) {
    std::vector<uint32_t> order;
    const std::vector<BinNode> bin = sah_binary(boxes, order);
    std::vector<std::vector<Cand>> lists;
    const uint32_t rt = dp_collapse(
        bin, 1, [](uint32_t) { return DP_C_INST; }, [&](const BinNode& x) { return LEAF_BIT | order[x.first]; },
        lists);
    if (rt & LEAF_BIT) {
        root = rt;
        return true;
    }
    if (list_levels(lists) > W8_MAX_LEVELS) return false;
    W8Builder w{nullptr, &top, &err};
    w.lists = &lists;
    return w.root(rt, root);
}

// The world-to-object rows (3 x 4, record words 0-11) inverted in f64: world =
// inv (object - T). False for a singular or non-finite transform.
bool inverse_rows(const uint32_t* r, double inv[3][3], double T[3]) {
    double R[3][3];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) R[a][b] = bits_f(r[4 * a + b]);
        T[a] = bits_f(r[4 * a + 3]);
    }
    const double c00 = R[1][1] * R[2][2] - R[1][2] * R[2][1], c01 = R[1][2] * R[2][0] - R[1][0] * R[2][2],
                 c02 = R[1][0] * R[2][1] - R[1][1] * R[2][0];
    const double det = R[0][0] * c00 + R[0][1] * c01 + R[0][2] * c02;
    if (!std::isfinite(det) || std::fabs(det) < 1e-30) return false;
    const double m[3][3] = {
        {c00, R[0][2] * R[2][1] - R[0][1] * R[2][2], R[0][1] * R[1][2] - R[0][2] * R[1][1]},
        {c01, R[0][0] * R[2][2] - R[0][2] * R[2][0], R[0][2] * R[1][0] - R[0][0] * R[1][2]},
        {c02, R[0][1] * R[2][0] - R[0][0] * R[2][1], R[0][0] * R[1][1] - R[0][1] * R[1][0]}};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            inv[a][b] = m[a][b] / det;
            if (!std::isfinite(inv[a][b])) return false;
        }
    return true;
}

// Pads a world box of an instance's geometry (f64, lo <= hi) as the record's own
// box is padded (src/gi_bvh.rs instance_words: 1e-5 (1 + max |coordinate|) per
// axis), rounds it outward to f32 and clips it to the record box `rec`; false when
// the result is empty or not finite (the record box is kept then).
bool padded_world_box(const double lo[3], const double hi[3], const Bounds& rec, Bounds& out) {
    for (int a = 0; a < 3; ++a) {
        const double pad = 1e-5 * (1.0 + std::max(std::fabs(lo[a]), std::fabs(hi[a])));
        out.lo[a] = std::max(std::nextafter((float)(lo[a] - pad), -INFINITY), rec.lo[a]);
        out.hi[a] = std::min(std::nextafter((float)(hi[a] + pad), INFINITY), rec.hi[a]);
        if (!std::isfinite(out.lo[a]) || !std::isfinite(out.hi[a]) || out.lo[a] > out.hi[a]) return false;
    }
    return true;
}

// The world box of each instance's own vertices, in place of the box of its
// mesh box's transformed corners (record words 12-17), which for a rotated mesh
// is up to about 1.7 times wider on each axis: a top-level leaf box only culls.
// The vertices (the triangle records `out.verts` of the instance's mesh) go to
// world space in f64 through the inverse of the record's world-to-object rows
// (words 0-11); the box is padded as the record's own box is (src/gi_bvh.rs
// instance_words: 1e-5 (1 + max |coordinate|) per axis), which is far wider
// than the f32 roundings of the object-space ray a hit is found with, then
// rounded outward to f32 and clipped to the record box (the kernel tests that
// box exactly at instance entry anyway). Instances of meshes above
// TIGHT_MAX_TRIS triangles, or with a singular or non-finite transform, keep
// the record box.
constexpr uint32_t TIGHT_MAX_TRIS = 1u << 16;
void tighten_instance_boxes(const HostLayout& plain, const HostLayout& out, std::vector<Bounds>& boxes) {
    const size_t total_tris = out.verts.size() / TRI_FLOATS;
    for (uint32_t i = 0; i < (uint32_t)boxes.size(); ++i) {
        const uint32_t* r = &out.inst[(size_t)i * INSTANCE_WORDS];
        const uint32_t mesh = plain.inst[(size_t)i * INSTANCE_WORDS + 21];
        const size_t tb = plain.inst[(size_t)i * INSTANCE_WORDS + 20];
        const size_t te = (size_t)mesh + 1 < plain.mesh_base.size() / 3 ? plain.mesh_base[3 * (mesh + 1) + 2]
                                                                         : total_tris;
        if (te <= tb || te > total_tris || te - tb > TIGHT_MAX_TRIS) continue;
        double inv[3][3], T[3];
        if (!inverse_rows(r, inv, T)) continue;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = TRI_FLOATS * tb; k < TRI_FLOATS * te; k += 3) {
            const double p[3] = {out.verts[k] - T[0], out.verts[k + 1] - T[1], out.verts[k + 2] - T[2]};
            for (int a = 0; a < 3; ++a) {
                const double w = inv[a][0] * p[0] + inv[a][1] * p[1] + inv[a][2] * p[2];
                lo[a] = std::min(lo[a], w);
                hi[a] = std::max(hi[a], w);
            }
        }
        Bounds t;
        if (padded_world_box(lo, hi, boxes[i], t)) boxes[i] = t;
    }
}

// The reflection records' version: the world box of each instance's triangle
// records (a, e1, e2 at word `tb[i]`, `n[i]` of them; n 0 keeps the record box),
// each vertex a, a + e1, a + e2 taken to world space in f64 and widened by its
// triangle's own box slack (2e-5 of its longest edge and 2^-20 of its largest
// coordinate, the bound rs_trace's barycentric slack needs, as the mesh trees
// use) scaled by the inverse's row sums, then padded and clipped as above.
void tighten_rs_boxes(const std::vector<uint32_t>& words, const std::vector<uint32_t>& tb,
                      const std::vector<uint32_t>& n, std::vector<Bounds>& boxes) {
    for (uint32_t i = 0; i < (uint32_t)boxes.size(); ++i) {
        if (n[i] == 0 || n[i] > TIGHT_MAX_TRIS || (size_t)tb[i] + (size_t)RS_TRI_WORDS * n[i] > words.size()) continue;
        double inv[3][3], T[3];
        if (!inverse_rows(&words[RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS], inv, T)) continue;
        double rows[3];
        for (int a = 0; a < 3; ++a) rows[a] = std::fabs(inv[a][0]) + std::fabs(inv[a][1]) + std::fabs(inv[a][2]);
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool finite = true;
        for (uint32_t k = 0; k < n[i] && finite; ++k) {
            const uint32_t* w = &words[(size_t)tb[i] + (size_t)RS_TRI_WORDS * k];
            double v[3][3], e = 0.0, m = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double p = bits_f(w[a]), e1 = bits_f(w[3 + a]), e2 = bits_f(w[6 + a]);
                v[0][a] = p;
                v[1][a] = p + e1;
                v[2][a] = p + e2;
                e = std::max({e, std::fabs(e1), std::fabs(e2)});
                m = std::max({m, std::fabs(p), std::fabs(p + e1), std::fabs(p + e2)});
            }
            const double slack = 2e-5 * e + std::ldexp(m, -20);
            finite = std::isfinite(slack);
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) {
                    const double x = inv[a][0] * (v[c][0] - T[0]) + inv[a][1] * (v[c][1] - T[1]) +
                                     inv[a][2] * (v[c][2] - T[2]);
                    lo[a] = std::min(lo[a], x - slack * rows[a]);
                    hi[a] = std::max(hi[a], x + slack * rows[a]);
                }
        }
        Bounds t;
        if (finite && padded_world_box(lo, hi, boxes[i], t)) boxes[i] = t;
    }
}

// The CW8 top level built from the instances' world boxes (record words 12-17,
// tightened to the instances' vertices), one instance a leaf. False with `err` empty when the 4-wide top level does not
// reach each instance exactly once, a box is not finite, or the tree would pass
// W8_MAX_LEVELS: the caller collapses the 4-wide top level instead.
// The CW8 top level from the instances' world boxes (record words 12-17), each
// tightened to its mesh's vertices. False (err empty) when a box is not finite.
bool top_from_records(const HostLayout& plain, HostLayout& out, std::string& err) {
    const uint32_t count = plain.instance_count;
    if (count == 0) return false;
    std::vector<Bounds> boxes(count);
    for (uint32_t i = 0; i < count; ++i)
        for (int a = 0; a < 3; ++a) {
            boxes[i].lo[a] = bits_f(out.inst[(size_t)i * INSTANCE_WORDS + 12 + a]);
            boxes[i].hi[a] = bits_f(out.inst[(size_t)i * INSTANCE_WORDS + 15 + a]);
            if (!std::isfinite(boxes[i].lo[a]) || !std::isfinite(boxes[i].hi[a])) return false;
        }
    tighten_instance_boxes(plain, out, boxes);
    return top_from_boxes(boxes, out.top, out.top_root, err);
}

bool rebuild_top(const HostLayout& plain, HostLayout& out, std::string& err) {
    const uint32_t count = plain.instance_count;
    if (count == 0) return false;
    std::vector<uint8_t> seen(count, 0);
    std::vector<uint32_t> todo{plain.top_root};
    size_t visits = 0;
    while (!todo.empty()) {
        const uint32_t t = todo.back();
        todo.pop_back();
        if (t & LEAF_BIT) {
            const uint32_t i = t & ~LEAF_BIT;
            if (i >= count || seen[i]) return false;
            seen[i] = 1;
            continue;
        }
        if (((size_t)t + 1) * TOP_WORDS > plain.top.size() || ++visits > 4 * (size_t)count + 64) return false;
        for (int s = 0; s < 4; ++s) {
            const uint32_t c = plain.top[(size_t)t * TOP_WORDS + 24 + s];
            if (c != EMPTY) todo.push_back(c);
        }
    }
    for (uint8_t s : seen)
        if (!s) return false;
    return top_from_records(plain, out, err);
}

// The CW8 tree of one reflection mesh (node base `base`, root `root`, triangle
// records at word `tb`) built from its triangles into `nodes` (indices from 0),
// its triangle records copied to `tris` in the new leaf order. Each triangle's box holds a, a + e1 and a + e2 computed exactly,
// widened by 2e-5 of its edges for rs_trace's barycentric slack (4e-6) and by
// 2^-20 of its coordinates, rounded outward. False with `err` empty when the
// tree does not hold each of the triangles [0, n) once (n its largest leaf end),
// a value is not finite, or the tree would pass W8_MAX_LEVELS: the caller
// collapses the reflection tree instead.
bool rebuild_rs_mesh(const std::vector<uint32_t>& words, uint32_t base, uint32_t root, uint32_t tb,
                     std::vector<uint32_t>& nodes, std::vector<uint32_t>& tris, uint32_t& cw8_root, std::string& err) {
    if (root & LEAF_BIT) return false;
    std::vector<uint8_t> seen;
    std::vector<uint32_t> todo{root};
    size_t visits = 0;
    while (!todo.empty()) {
        const uint32_t t = todo.back();
        todo.pop_back();
        if (t & LEAF_BIT) {
            const uint32_t first = (t & ~LEAF_BIT) >> 4, cnt = t & 15u;
            if (seen.size() < (size_t)first + cnt) seen.resize((size_t)first + cnt, 0);
            for (uint32_t k = first; k < first + cnt; ++k) {
                if (seen[k]) return false;
                seen[k] = 1;
            }
            continue;
        }
        const size_t at = (size_t)base + (size_t)t * RS_NODE_WORDS;
        if (at + RS_NODE_WORDS > words.size() || ++visits > (size_t)1 << 27) return false;
        for (int s = 0; s < 4; ++s) {
            const uint32_t c = words[at + 6 + s];
            if (c != EMPTY) todo.push_back(c);
        }
    }
    const uint32_t n = (uint32_t)seen.size();
    if (n == 0 || n >= MAX_FIRST_TRI || (size_t)tb + (size_t)RS_TRI_WORDS * n > words.size()) return false;
    for (uint8_t s : seen)
        if (!s) return false;
    std::vector<Bounds> boxes(n);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t* w = &words[(size_t)tb + (size_t)RS_TRI_WORDS * k];
        double lo[3], hi[3], e = 0.0, m = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double p = bits_f(w[a]), e1 = bits_f(w[3 + a]), e2 = bits_f(w[6 + a]);
            if (!std::isfinite(p) || !std::isfinite(e1) || !std::isfinite(e2)) return false;
            lo[a] = std::min({p, p + e1, p + e2});
            hi[a] = std::max({p, p + e1, p + e2});
            e = std::max({e, std::fabs(e1), std::fabs(e2)});
            m = std::max({m, std::fabs(lo[a]), std::fabs(hi[a])});
        }
        const double pad = 2e-5 * e + std::ldexp(m, -20);
        for (int a = 0; a < 3; ++a) {
            float l = (float)(lo[a] - pad), h = (float)(hi[a] + pad);
            if ((double)l > lo[a] - pad) l = std::nextafter(l, -std::numeric_limits<float>::infinity());
            if ((double)h < hi[a] + pad) h = std::nextafter(h, std::numeric_limits<float>::infinity());
            if (!std::isfinite(l) || !std::isfinite(h)) return false;
            boxes[k].lo[a] = l;
            boxes[k].hi[a] = h;
        }
    }
    std::vector<uint32_t> order;
    const std::vector<BinNode> bin = sah_binary(boxes, order);
    std::vector<std::vector<Cand>> lists;
    const uint32_t rt = dp_collapse(
        bin, DP_LEAF_MAX, [](uint32_t cnt) { return DP_C_PAIR * ((cnt + 1) / 2); },
        [&](const BinNode& x) { return LEAF_BIT | (x.first << 4) | x.count; }, lists);
    if (rt & LEAF_BIT) {
        cw8_root = rt;
    } else {
        if (list_levels(lists) > W8_MAX_LEVELS) return false;
        W8Builder w{nullptr, &nodes, &err};
        w.lists = &lists;
        if (!w.root(rt, cw8_root)) return false;
    }
    tris.reserve((size_t)RS_TRI_WORDS * n);
    for (uint32_t p = 0; p < n; ++p) {
        const size_t src = (size_t)tb + (size_t)RS_TRI_WORDS * order[p];
        tris.insert(tris.end(), words.begin() + src, words.begin() + src + RS_TRI_WORDS);
    }
    return true;
}

}  // namespace

bool build_cw8(const std::vector<uint32_t>& words, const HostLayout& plain, HostLayout& out, std::string& err) {
    out = HostLayout();
    out.inst = plain.inst;
    out.tris = plain.tris;
    out.tri_ids = plain.tri_ids;
    out.verts = plain.verts;
    out.instance_count = plain.instance_count;
    // BLAS: every mesh root reachable from an instance record, built once -- from
    // its triangles where the mesh's 4-wide tree holds each of them once, else
    // collapsed from that tree. Meshes are built on threads, each into nodes of
    // its own (each moves only its own triangles), then placed one after another.
    struct MeshJob {
        uint32_t pr, tb, n;
        std::vector<uint32_t> nodes;
        uint32_t root = EMPTY;
        std::string err;
    };
    std::vector<MeshJob> jobs;
    std::unordered_map<uint32_t, uint32_t> job_of;  // plain root -> job
    const size_t total_tris = out.verts.size() / TRI_FLOATS;
    for (uint32_t i = 0; i < plain.instance_count; ++i) {
        const uint32_t pr = plain.inst[(size_t)i * INSTANCE_WORDS + 18];
        if (job_of.count(pr)) continue;
        const uint32_t mesh = plain.inst[(size_t)i * INSTANCE_WORDS + 21];
        const uint32_t tb = plain.inst[(size_t)i * INSTANCE_WORDS + 20];
        const size_t te = (size_t)mesh + 1 < plain.mesh_base.size() / 3 ? plain.mesh_base[3 * (mesh + 1) + 2]
                                                                         : total_tris;
        job_of.emplace(pr, (uint32_t)jobs.size());
        jobs.push_back(MeshJob{pr, tb, (uint32_t)(te - tb), {}, EMPTY, {}});
    }
    auto build_one = [&](MeshJob& j) {
        if (rebuild_mesh(plain, j.pr, j.tb, j.n, out, j.nodes, j.root, j.err) || !j.err.empty()) return;
        if (wide_levels(plain.nodes, j.pr, 7 * W8_MAX_LEVELS) > 7 * W8_MAX_LEVELS) {
            j.err = "CW8 relayout: a mesh tree deeper than 16 levels";
            return;
        }
        W8Builder b{&plain.nodes, &j.nodes, &j.err};
        b.tris = &out;
        if (b.root(j.pr, j.root) && w8_levels(j.nodes, j.root, W8_MAX_LEVELS) > W8_MAX_LEVELS)
            j.err = "CW8 relayout: a mesh tree deeper than 16 levels";
    };
    {
        const uint32_t threads = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), 16));
        std::vector<std::thread> pool;
        std::atomic_uint next{0};
        for (uint32_t t = 0; t < std::min<size_t>(threads, jobs.size()); ++t)
            pool.emplace_back([&]() {
                for (uint32_t k; (k = next.fetch_add(1)) < jobs.size();) {
                    try {
                        build_one(jobs[k]);
                    } catch (const std::bad_alloc&) {
                        jobs[k].err = "CW8 relayout: out of host memory";
                    }
                }
            });
        for (auto& th : pool) th.join();
    }
    std::vector<uint32_t> roots(jobs.size());
    for (size_t k = 0; k < jobs.size(); ++k) {
        MeshJob& j = jobs[k];
        if (!j.err.empty()) {
            err = j.err;
            return false;
        }
        // the mesh's node indices (its children's tags and its root) move by its first node
        const uint32_t base = (uint32_t)(out.nodes.size() / W8_WORDS);
        for (size_t w = 0; w < j.nodes.size(); w += W8_WORDS)
            for (size_t s = 16; s < 24; ++s)
                if (!(j.nodes[w + s] & LEAF_BIT)) j.nodes[w + s] += base;
        out.nodes.insert(out.nodes.end(), j.nodes.begin(), j.nodes.end());
        roots[k] = (j.root & LEAF_BIT) ? j.root : j.root + base;
        std::vector<uint32_t>().swap(j.nodes);
    }
    for (uint32_t i = 0; i < plain.instance_count; ++i)
        out.inst[(size_t)i * INSTANCE_WORDS + 18] = roots[job_of.at(plain.inst[(size_t)i * INSTANCE_WORDS + 18])];
    // Top level: built from the instances' world boxes where the 4-wide top level
    // reaches each instance once, else that top level collapsed the same way as
    // the meshes' trees (instance leaves keep their LEAF | instance tags).
    out.top_on = plain.top_on;
    out.top_root = plain.top_root;
    if (plain.top_on && !rebuild_top(plain, out, err)) {
        if (!err.empty()) return false;
        out.top.clear();
        out.top_root = plain.top_root;
        W8Builder t{&plain.top, &out.top, &err};
        if (wide_levels(plain.top, plain.top_root, 7 * W8_MAX_LEVELS) > 7 * W8_MAX_LEVELS) {
            err = "CW8 relayout: a top level deeper than 16 levels";
            return false;
        }
        if (!t.root(plain.top_root, out.top_root)) return false;
        if (w8_levels(out.top, out.top_root, W8_MAX_LEVELS) > W8_MAX_LEVELS) {
            err = "CW8 relayout: a top level deeper than 16 levels";
            return false;
        }
    }
    // Words without a top level (fewer than 40 instances): CW8 builds one all the same.
    // The CW8 kernel does not test an instance's world box at entry -- the top-level leaf
    // box it came through holds the box of the instance's triangles, quantized outward --
    // so the top level is its instances' only cull.
    if (!plain.top_on && plain.instance_count > 0) {
        if (top_from_records(plain, out, err)) {
            out.top_on = true;
        } else {
            if (!err.empty()) return false;
            out.top.clear();  // not finite: the flat loop, every instance entered
        }
    }
    if (out.top.empty()) out.top.assign(W8_WORDS, 0u);
    if (out.nodes.empty()) out.nodes.assign(W8_WORDS, 0u);
    // Each instance's BLAS root node, copied beside the records: the kernel loads an
    // instance's root together with its record instead of after it.
    out.roots.assign((size_t)std::max(out.instance_count, 1u) * W8_WORDS, 0u);
    for (uint32_t i = 0; i < out.instance_count; ++i) {
        const uint32_t root = out.inst[(size_t)i * INSTANCE_WORDS + 18];
        if (root & LEAF_BIT) continue;
        std::copy(out.nodes.begin() + (size_t)root * W8_WORDS, out.nodes.begin() + ((size_t)root + 1) * W8_WORDS,
                  out.roots.begin() + (size_t)i * W8_WORDS);
    }
    (void)words;
    return true;
}

bool build_rs_cw8(const std::vector<uint32_t>& words, HostLayout& out, std::string& err) {
    out = HostLayout();
    const uint32_t count = words.empty() ? 0u : words[0];
    out.inst = words;
    out.instance_count = count;
    // Each mesh's CW8 tree is built from its triangle records where its tree holds
    // each of them once (the records then follow the words in the new leaf order,
    // and the instance records point there). Otherwise each mesh tree (node base,
    // root) is decoded as rs_trace decodes it (origin + step * code, unfused) into
    // 32-word f32 nodes and collapsed into CW8 nodes; RS leaf tags (LEAF | first
    // triangle << 4 | count) pass through unchanged.
    std::vector<uint32_t> plain;
    W8Builder b{&plain, &out.nodes, &err};
    std::unordered_map<uint64_t, uint32_t> done;  // node base << 32 | root -> CW8 root tag
    // the rebuilds, one per (node base, root, triangle base), on threads
    struct RsJob {
        uint32_t base, root, tb;
        std::vector<uint32_t> nodes, tris;
        uint32_t cw8_root = EMPTY, new_tb = 0, n = 0;
        bool ok = false, placed = false;
        std::string err;
    };
    std::vector<RsJob> jobs;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> job_of;
    for (uint32_t i = 0; i < count; ++i) {
        const size_t r = RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS;
        const auto k3 = std::make_tuple(words[r + 18], words[r + 20], words[r + 19]);
        if (job_of.emplace(k3, (uint32_t)jobs.size()).second)
            jobs.push_back(RsJob{words[r + 18], words[r + 20], words[r + 19], {}, {}, EMPTY, 0, 0, false, false, {}});
    }
    {
        const uint32_t threads = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), 16));
        std::vector<std::thread> pool;
        std::atomic_uint next{0};
        for (uint32_t t = 0; t < std::min<size_t>(threads, jobs.size()); ++t)
            pool.emplace_back([&]() {
                for (uint32_t k; (k = next.fetch_add(1)) < jobs.size();) {
                    RsJob& j = jobs[k];
                    try {
                        j.ok = rebuild_rs_mesh(words, j.base, j.root, j.tb, j.nodes, j.tris, j.cw8_root, j.err);
                    } catch (const std::bad_alloc&) {
                        j.err = "RS CW8 relayout: out of host memory";
                    }
                }
            });
        for (auto& th : pool) th.join();
    }
    std::vector<uint32_t> rec_tb(count, 0u), rec_n(count, 0u);  // rebuilt meshes' triangle records per record
    for (uint32_t i = 0; i < count; ++i) {
        const size_t r = RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS;
        const uint32_t base = words[r + 18], root = words[r + 20];
        RsJob& j = jobs[job_of.at(std::make_tuple(base, root, words[r + 19]))];
        if (!j.err.empty()) {
            err = j.err;
            return false;
        }
        if (j.ok) {
            if (!j.placed) {
                // its node indices move by the first node it lands on; its records follow the words
                const uint32_t nb = (uint32_t)(out.nodes.size() / W8_WORDS);
                for (size_t w = 0; w < j.nodes.size(); w += W8_WORDS)
                    for (size_t s = 16; s < 24; ++s)
                        if (!(j.nodes[w + s] & LEAF_BIT)) j.nodes[w + s] += nb;
                if (!(j.cw8_root & LEAF_BIT)) j.cw8_root += nb;
                if (out.inst.size() + j.tris.size() >= ((size_t)1 << 32)) {
                    err = "RS CW8 relayout: words past 2^32";
                    return false;
                }
                j.new_tb = (uint32_t)out.inst.size();
                j.n = (uint32_t)(j.tris.size() / RS_TRI_WORDS);
                out.nodes.insert(out.nodes.end(), j.nodes.begin(), j.nodes.end());
                out.inst.insert(out.inst.end(), j.tris.begin(), j.tris.end());
                std::vector<uint32_t>().swap(j.nodes);
                std::vector<uint32_t>().swap(j.tris);
                j.placed = true;
            }
            out.inst[r + 19] = j.new_tb;
            out.inst[r + 20] = j.cw8_root;
            rec_tb[i] = j.new_tb;
            rec_n[i] = j.n;
            continue;
        }
        const uint64_t key = (uint64_t)base << 32 | root;
        auto it = done.find(key);
        uint32_t cw8_root = EMPTY;
        if (it != done.end()) {
            cw8_root = it->second;
        } else {
            uint32_t plain_root = root;
            if (!(root & LEAF_BIT)) {
                std::unordered_map<uint32_t, uint32_t> id;  // source node id -> f32 node index
                std::vector<uint32_t> todo{root};
                id[root] = (uint32_t)(plain.size() / TOP_WORDS);
                plain.resize(plain.size() + TOP_WORDS, 0u);
                while (!todo.empty()) {
                    const uint32_t n = todo.back();
                    todo.pop_back();
                    const size_t src = base + (size_t)n * RS_NODE_WORDS;
                    if (src + RS_NODE_WORDS > words.size()) {
                        err = "a node past the end of the words";
                        return false;
                    }
                    const uint32_t* w = &words[src];
                    const size_t o = (size_t)id[n] * TOP_WORDS;
                    for (int sl = 0; sl < 4; ++sl) {
                        const uint32_t c = w[6 + sl];
                        if (c == EMPTY) {
                            plain[o + 24 + sl] = EMPTY;
                            continue;
                        }
                        float lo[3], hi[3];
                        decode_unfused(w, sl, lo, hi);
                        for (int a = 0; a < 3; ++a) {
                            plain[o + a * 4 + sl] = f_bits(lo[a]);
                            plain[o + 12 + a * 4 + sl] = f_bits(hi[a]);
                        }
                        if (c & LEAF_BIT) {
                            plain[o + 24 + sl] = c;
                            continue;
                        }
                        auto [at, fresh] = id.emplace(c, (uint32_t)(plain.size() / TOP_WORDS));
                        if (fresh) {
                            plain.resize(plain.size() + TOP_WORDS, 0u);
                            todo.push_back(c);
                        }
                        plain[o + 24 + sl] = at->second;
                    }
                }
                plain_root = id[root];
            }
            if (wide_levels(plain, plain_root, 7 * W8_MAX_LEVELS) > 7 * W8_MAX_LEVELS ||
                !b.root(plain_root, cw8_root) || w8_levels(out.nodes, cw8_root, W8_MAX_LEVELS) > W8_MAX_LEVELS) {
                if (err.empty()) err = "RS CW8 relayout: a mesh tree deeper than 16 levels";
                return false;
            }
            done.emplace(key, cw8_root);
        }
        out.inst[r + 20] = cw8_root;
    }
    // Top level: built from the instances' world boxes (record words 12-17), each
    // tightened to its mesh's triangle records where the mesh tree was rebuilt
    // from them; rs_trace's closest hit is the least (t, instance, primitive)
    // whatever the order the instances are visited in. Boxes that are not finite,
    // or a tree past W8_MAX_LEVELS: the reference's 4-wide f32 nodes, collapsed
    // the same way as the meshes' trees. Words without a top level get one all the
    // same (the kernel tests no record box at instance entry for the 8-wide layouts:
    // the top-level leaf box holds the instance's triangles); where it cannot be
    // built, the flat loop enters every instance.
    const bool had_top = count && words[1] != 0;
    out.top_on = count > 0;
    out.top_root = had_top ? words[2] : 0u;
    bool rebuilt = false;
    if (out.top_on) {
        std::vector<Bounds> boxes(count);
        bool finite = true;
        for (uint32_t i = 0; i < count && finite; ++i)
            for (int a = 0; a < 3; ++a) {
                const size_t r = RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS;
                boxes[i].lo[a] = bits_f(out.inst[r + 12 + a]);
                boxes[i].hi[a] = bits_f(out.inst[r + 15 + a]);
                finite = finite && std::isfinite(boxes[i].lo[a]) && std::isfinite(boxes[i].hi[a]);
            }
        if (finite) {
            tighten_rs_boxes(out.inst, rec_tb, rec_n, boxes);
            rebuilt = top_from_boxes(boxes, out.top, out.top_root, err);
            if (!rebuilt) {
                if (!err.empty()) return false;
                out.top.clear();
                out.top_root = had_top ? words[2] : 0u;
            }
        }
        if (!rebuilt && !had_top) out.top_on = false;
    }
    if (out.top_on && !rebuilt) {
        const size_t at = words[1];
        const size_t n = std::min<size_t>(top_reserve(count), words.size() - std::min(words.size(), at));
        std::vector<uint32_t> top(words.begin() + at, words.begin() + at + n);
        W8Builder t{&top, &out.top, &err};
        if (wide_levels(top, words[2], 7 * W8_MAX_LEVELS) > 7 * W8_MAX_LEVELS ||
            !t.root(words[2], out.top_root) || w8_levels(out.top, out.top_root, W8_MAX_LEVELS) > W8_MAX_LEVELS) {
            if (err.empty()) err = "RS CW8 relayout: a top level deeper than 16 levels";
            return false;
        }
    }
    if (out.top.empty()) out.top.assign(W8_WORDS, 0u);
    if (out.nodes.empty()) out.nodes.assign(W8_WORDS, 0u);
    // each instance's BLAS root node copied beside the records, as for CW8
    out.roots.assign((size_t)std::max(count, 1u) * W8_WORDS, 0u);
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t root = out.inst[RS_HEADER_WORDS + (size_t)i * RS_RECORD_WORDS + 20];
        if (root & LEAF_BIT) continue;
        std::copy(out.nodes.begin() + (size_t)root * W8_WORDS, out.nodes.begin() + ((size_t)root + 1) * W8_WORDS,
                  out.roots.begin() + (size_t)i * W8_WORDS);
    }
    return true;
}

}  // namespace rq
