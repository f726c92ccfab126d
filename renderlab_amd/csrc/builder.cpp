// Runtime BVH4 builder over one triangle list: binned-SAH or median splits,
// at most eight triangles a leaf, child boxes quantized outward to 8 bits so
// both the fused and the unfused f32 decode enclose them.
//
// Behaviour follows crates/asset/src/near_field.rs:172-541 (build_bvh,
// build_bvh_sah, build, Builder::{subtree, split, split_sah, split_median,
// leaf, child_bounds, node}, refit_bvh, quantized_node); the record layout is
// crates/asset/src/cluster_bvh.rs:21-54. The arithmetic order of every
// quantity that lands in a record (centroids, bins, areas, steps, codes) is
// kept, since the hit results depend on the box bits.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace rq {
namespace {

constexpr uint32_t SAH_BINS = 16;  // near_field.rs:273

float half_area(const Bounds& b) {
    float e[3];
    for (int i = 0; i < 3; ++i) e[i] = mx(b.hi[i] - b.lo[i], 0.0f);
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

struct Builder {
    const float* positions;
    const uint32_t* indices;
    std::vector<float> centroids;  // 3 per triangle
    std::vector<Bounds> boxes;     // per triangle (SAH builds)
    Bvh* bvh;

    Bounds centroid_bounds(const std::vector<uint32_t>& g) const {
        Bounds b = empty_bounds();
        for (uint32_t t : g)
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = mn(b.lo[a], centroids[3 * t + a]);
                b.hi[a] = mx(b.hi[a], centroids[3 * t + a]);
            }
        return b;
    }

    // near_field.rs:317-366
    bool split_sah(const std::vector<uint32_t>& g, std::vector<uint32_t>& left,
                   std::vector<uint32_t>& right) const {
        Bounds c = centroid_bounds(g);
        bool found = false;
        float best_cost = 0.0f;
        int best_axis = 0;
        uint32_t best_plane = 0;
        for (int axis = 0; axis < 3; ++axis) {
            float extent = c.hi[axis] - c.lo[axis];
            if (std::isnan(extent) || extent <= 0.0f) continue;
            uint32_t count[SAH_BINS] = {};
            Bounds bins[SAH_BINS];
            for (auto& b : bins) b = empty_bounds();
            for (uint32_t t : g) {
                float f = (centroids[3 * t + axis] - c.lo[axis]) / extent * (float)SAH_BINS;
                uint32_t b = std::min(sat_index(f, 0xffffffffu), SAH_BINS - 1);
                count[b] += 1;
                bins[b] = merge(bins[b], boxes[t]);
            }
            float right_cost[SAH_BINS] = {};
            Bounds acc = empty_bounds();
            uint32_t n = 0;
            for (uint32_t b = SAH_BINS - 1; b >= 1; --b) {
                acc = merge(acc, bins[b]);
                n += count[b];
                right_cost[b] = half_area(acc) * (float)n;
            }
            acc = empty_bounds();
            n = 0;
            for (uint32_t b = 0; b < SAH_BINS - 1; ++b) {
                acc = merge(acc, bins[b]);
                n += count[b];
                if (n == 0 || n == g.size()) continue;
                float cost = half_area(acc) * (float)n + right_cost[b + 1];
                if (!found || cost < best_cost) {
                    found = true;
                    best_cost = cost;
                    best_axis = axis;
                    best_plane = b;
                }
            }
        }
        if (!found) return false;
        float extent = c.hi[best_axis] - c.lo[best_axis];
        left.clear();
        right.clear();
        for (uint32_t t : g) {
            float f = (centroids[3 * t + best_axis] - c.lo[best_axis]) / extent * (float)SAH_BINS;
            uint32_t b = std::min(sat_index(f, 0xffffffffu), SAH_BINS - 1);
            (b <= best_plane ? left : right).push_back(t);
        }
        return true;
    }

    // near_field.rs:367-379: longest centroid axis (ties to the later axis, as
    // Iterator::max_by), stable order by (centroid, triangle), halves.
    void split_median(std::vector<uint32_t> g, std::vector<uint32_t>& left,
                      std::vector<uint32_t>& right) const {
        Bounds c = centroid_bounds(g);
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (total_key(c.hi[a] - c.lo[a]) >= total_key(c.hi[axis] - c.lo[axis])) axis = a;
        std::sort(g.begin(), g.end(), [&](uint32_t x, uint32_t y) {
            int32_t kx = total_key(centroids[3 * x + axis]), ky = total_key(centroids[3 * y + axis]);
            return kx != ky ? kx < ky : x < y;
        });
        size_t mid = g.size() / 2;
        left.assign(g.begin(), g.begin() + mid);
        right.assign(g.begin() + mid, g.end());
    }

    void split(std::vector<uint32_t>&& g, std::vector<uint32_t>& a, std::vector<uint32_t>& b) const {
        if (!boxes.empty() && split_sah(g, a, b)) return;
        split_median(std::move(g), a, b);
    }

    // near_field.rs:380-401
    uint32_t leaf(const std::vector<uint32_t>& tris) {
        uint32_t offset = (uint32_t)bvh->indices.size();
        for (uint32_t t : tris)
            for (int k = 0; k < 3; ++k) bvh->indices.push_back(indices[3 * t + k]);
        uint32_t id = (uint32_t)bvh->leaf_offset.size();
        bvh->leaf_offset.push_back(offset);
        bvh->leaf_count.push_back((uint32_t)tris.size() * 3u);
        bvh->leaf_first.push_back(tris[0]);
        return id | LEAF_BIT;
    }

    // near_field.rs:402-430
    Bounds child_bounds(uint32_t child) const {
        Bounds b = empty_bounds();
        if (child & LEAF_BIT) {
            uint32_t l = child & ~LEAF_BIT;
            uint32_t off = bvh->leaf_offset[l], cnt = bvh->leaf_count[l];
            for (uint32_t k = off; k < off + cnt; ++k) {
                const float* p = positions + 3ull * bvh->indices[k];
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = mn(b.lo[a], p[a]);
                    b.hi[a] = mx(b.hi[a], p[a]);
                }
            }
            return b;
        }
        const Node& n = bvh->nodes[child];
        for (int slot = 0; slot < 4; ++slot) {
            if (n.children[slot] == EMPTY) continue;
            for (int fused = 0; fused < 2; ++fused) {
                float lo[3], hi[3];
                decode_child(n, slot, fused != 0, lo, hi);
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = mn(b.lo[a], lo[a]);
                    b.hi[a] = mx(b.hi[a], hi[a]);
                }
            }
        }
        return b;
    }

    uint32_t node(const std::vector<uint32_t>& children);

    // near_field.rs:286-306
    uint32_t subtree(std::vector<uint32_t>&& tris) {
        if (tris.size() <= LEAF_TRIANGLES) return leaf(tris);
        std::vector<std::vector<uint32_t>> groups;
        groups.push_back(std::move(tris));
        for (int round = 0; round < 2; ++round) {
            std::vector<std::vector<uint32_t>> next;
            for (auto& g : groups) {
                if (g.size() <= LEAF_TRIANGLES) {
                    next.push_back(std::move(g));
                    continue;
                }
                std::vector<uint32_t> a, b;
                split(std::move(g), a, b);
                next.push_back(std::move(a));
                next.push_back(std::move(b));
            }
            groups = std::move(next);
        }
        std::vector<uint32_t> children;
        for (auto& g : groups) children.push_back(subtree(std::move(g)));
        return node(children);
    }
};

// near_field.rs:481-526
Node quantized_node(const std::vector<uint32_t>& children, const std::vector<Bounds>& boxes) {
    Bounds all = empty_bounds();
    for (const Bounds& b : boxes) all = merge(all, b);
    Node node;
    std::memset(&node, 0, sizeof node);
    for (int a = 0; a < 4; ++a) node.children[a] = EMPTY;
    for (int axis = 0; axis < 3; ++axis) {
        node.origin[axis] = all.lo[axis];
        float extent = all.hi[axis] - all.lo[axis];
        float step = mx(extent / 255.0f, mx(1.17549435e-38f, std::fabs(all.lo[axis]) * 1e-7f));
        const float grow = 1.0f + 1.0f / 1048576.0f;
        while (all.lo[axis] + step * 255.0f < all.hi[axis] ||
               std::fmaf(step, 255.0f, all.lo[axis]) < all.hi[axis])
            step *= grow;
        node.step[axis] = step;
    }
    for (size_t slot = 0; slot < children.size(); ++slot) {
        node.children[slot] = children[slot];
        for (int axis = 0; axis < 3; ++axis) {
            float o = node.origin[axis], s = node.step[axis];
            auto above = [&](uint32_t q, float v) {
                return o + s * (float)q > v || std::fmaf(s, (float)q, o) > v;
            };
            auto below = [&](uint32_t q, float v) {
                return o + s * (float)q < v || std::fmaf(s, (float)q, o) < v;
            };
            float fl = std::floor((boxes[slot].lo[axis] - o) / s);
            fl = mn(mx(fl, 0.0f), 255.0f);
            uint32_t low = (uint32_t)fl;
            while (low > 0 && above(low, boxes[slot].lo[axis])) low -= 1;
            float fh = std::ceil((boxes[slot].hi[axis] - o) / s);
            fh = mn(mx(fh, 0.0f), 255.0f);
            uint32_t high = (uint32_t)fh;
            while (high < 255 && below(high, boxes[slot].hi[axis])) high += 1;
            node.lower[axis][slot] = (uint8_t)low;
            node.upper[axis][slot] = (uint8_t)high;
        }
    }
    return node;
}

uint32_t Builder::node(const std::vector<uint32_t>& children) {
    std::vector<Bounds> b;
    b.reserve(children.size());
    for (uint32_t c : children) b.push_back(child_bounds(c));
    Node n = quantized_node(children, b);
    uint32_t id = (uint32_t)bvh->nodes.size();
    bvh->nodes.push_back(n);
    return id;
}

bool check_input(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                 uint32_t index_count, std::string& err) {
    if (index_count == 0 || index_count % 3 != 0 || !indices) {
        err = "BVH geometry needs a nonempty triangle list";
        return false;
    }
    for (uint32_t k = 0; k < index_count; ++k)
        if (indices[k] >= vertex_count) {
            err = "BVH geometry has an invalid index or position";
            return false;
        }
    for (uint64_t v = 0; v < 3ull * vertex_count; ++v)
        if (!std::isfinite(positions[v])) {
            err = "BVH geometry has an invalid index or position";
            return false;
        }
    return true;
}

}  // namespace

// cluster_bvh.rs:40-53: o + s*q (unfused) or fma(s, q, o) (fused).
void decode_child(const Node& n, int slot, bool fused, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        float o = n.origin[a], s = n.step[a];
        float ql = (float)n.lower[a][slot], qh = (float)n.upper[a][slot];
        lo[a] = fused ? std::fmaf(s, ql, o) : o + s * ql;
        hi[a] = fused ? std::fmaf(s, qh, o) : o + s * qh;
    }
}

uint32_t bvh_depth(const Bvh& bvh, uint32_t child) {
    if (child & LEAF_BIT) return 0;
    // Post-order ids: children precede parents, so one forward pass suffices.
    std::vector<uint32_t> d(bvh.nodes.size(), 0);
    for (size_t i = 0; i < bvh.nodes.size(); ++i) {
        uint32_t m = 0;
        for (uint32_t c : bvh.nodes[i].children)
            if (c != EMPTY && !(c & LEAF_BIT)) m = std::max(m, d[c]);
        d[i] = 1 + m;
    }
    return child < d.size() ? d[child] : 0;
}

// near_field.rs:206-262
bool build_bvh(const float* positions, uint32_t vertex_count, const uint32_t* indices,
               uint32_t index_count, bool sah, Bvh& out, std::string& err) {
    if (!check_input(positions, vertex_count, indices, index_count, err)) return false;
    uint32_t tri_count = index_count / 3;
    Builder b;
    b.positions = positions;
    b.indices = indices;
    b.centroids.resize(3ull * tri_count);
    for (uint32_t t = 0; t < tri_count; ++t)
        for (int a = 0; a < 3; ++a) {
            float s = positions[3ull * indices[3 * t] + a] + positions[3ull * indices[3 * t + 1] + a];
            s = s + positions[3ull * indices[3 * t + 2] + a];
            b.centroids[3ull * t + a] = s / 3.0f;
        }
    if (sah) {
        b.boxes.resize(tri_count);
        for (uint32_t t = 0; t < tri_count; ++t) {
            Bounds bb = empty_bounds();
            for (int k = 0; k < 3; ++k) {
                const float* p = positions + 3ull * indices[3 * t + k];
                for (int a = 0; a < 3; ++a) {
                    bb.lo[a] = mn(bb.lo[a], p[a]);
                    bb.hi[a] = mx(bb.hi[a], p[a]);
                }
            }
            b.boxes[t] = bb;
        }
    }
    out = Bvh();
    out.positions.assign(positions, positions + 3ull * vertex_count);
    out.indices.reserve(index_count);
    b.bvh = &out;
    std::vector<uint32_t> all(tri_count);
    for (uint32_t t = 0; t < tri_count; ++t) all[t] = t;
    out.root = b.subtree(std::move(all));
    return true;
}

// near_field.rs:184-190
bool build_bvh_sah(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                   uint32_t index_count, Bvh& out, std::string& err) {
    if (!build_bvh(positions, vertex_count, indices, index_count, true, out, err)) return false;
    if (bvh_depth(out, out.root) > MAX_SAH_DEPTH)
        return build_bvh(positions, vertex_count, indices, index_count, false, out, err);
    return true;
}

// near_field.rs:445-478
bool refit_bvh(Bvh& bvh, const float* positions, uint32_t vertex_count, std::string& err) {
    if (vertex_count * 3ull != bvh.positions.size()) {
        err = "BVH refit needs finite positions of the same vertices";
        return false;
    }
    for (uint64_t v = 0; v < 3ull * vertex_count; ++v)
        if (!std::isfinite(positions[v])) {
            err = "BVH refit needs finite positions of the same vertices";
            return false;
        }
    bvh.positions.assign(positions, positions + 3ull * vertex_count);
    Builder b;
    b.positions = bvh.positions.data();
    b.indices = nullptr;
    b.bvh = &bvh;
    for (size_t n = 0; n < bvh.nodes.size(); ++n) {
        std::vector<uint32_t> children;
        for (uint32_t c : bvh.nodes[n].children)
            if (c != EMPTY) children.push_back(c);
        std::vector<Bounds> boxes;
        for (uint32_t c : children) boxes.push_back(b.child_bounds(c));
        bvh.nodes[n] = quantized_node(children, boxes);
    }
    return true;
}

}  // namespace rq
