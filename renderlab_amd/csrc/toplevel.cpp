// The top level over instance world boxes: a binned-SAH (or median) BVH2,
// collapsed into 4-wide nodes of float boxes, rebuilt from median splits when
// too deep for the traversal's 48-entry stack; refit and place.
//
// Behaviour follows src/instance_bvh.rs:24-472 (reserve, enabled, place,
// refit, refit_wide, write_child, sah_split, build, build_with, collapse,
// wide_depth, subtree, union). The 32-word wide node: child boxes axis-major
// (lo x[4], lo y[4], lo z[4], hi x[4], hi y[4], hi z[4]), four tagged
// children (node index, LEAF|instance, or EMPTY), four pad words. The
// reference's worker-thread build for large sets is not restated: this
// library builds on the calling thread.
#include <algorithm>

#include "internal.h"

namespace rq {
namespace {

constexpr uint32_t BINS = 16;  // src/instance_bvh.rs:287

float area(const Bounds& b) {
    float e[3];
    for (int a = 0; a < 3; ++a) e[a] = mx(b.hi[a] - b.lo[a], 0.0f);
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

struct Centre {
    const std::vector<Bounds>* boxes;
    float at(uint32_t i, int a) const { return (*boxes)[i].lo[a] + (*boxes)[i].hi[a]; }
};

// src/instance_bvh.rs:298-342
bool sah_split(const std::vector<Bounds>& boxes, const uint32_t* order, size_t n, const Centre& c,
               const Bounds& cb, int& axis_out, uint32_t& last_out, float& base, float& scale_out) {
    bool found = false;
    float best = 0.0f;
    for (int axis = 0; axis < 3; ++axis) {
        float extent = cb.hi[axis] - cb.lo[axis];
        if (!(extent > 0.0f)) continue;
        float scale = (float)BINS / extent;
        uint32_t count[BINS] = {};
        Bounds bins[BINS];
        for (auto& b : bins) b = empty_bounds();
        for (size_t k = 0; k < n; ++k) {
            uint32_t i = order[k];
            uint32_t b = std::min(sat_index((c.at(i, axis) - cb.lo[axis]) * scale, 0xffffffffu), BINS - 1);
            count[b] += 1;
            bins[b] = merge(bins[b], boxes[i]);
        }
        float right_area[BINS] = {};
        uint32_t right_n[BINS] = {};
        Bounds acc = empty_bounds();
        uint32_t m = 0;
        for (uint32_t b = BINS - 1; b >= 1; --b) {
            acc = merge(acc, bins[b]);
            m += count[b];
            right_area[b - 1] = area(acc);
            right_n[b - 1] = m;
        }
        acc = empty_bounds();
        m = 0;
        for (uint32_t b = 0; b < BINS - 1; ++b) {
            acc = merge(acc, bins[b]);
            m += count[b];
            if (m == 0 || right_n[b] == 0) continue;
            float cost = area(acc) * (float)m + right_area[b] * (float)right_n[b];
            if (!found || cost < best) {
                found = true;
                best = cost;
                axis_out = axis;
                last_out = b;
            }
        }
    }
    if (!found) return false;
    base = cb.lo[axis_out];
    scale_out = (float)BINS / (cb.hi[axis_out] - cb.lo[axis_out]);
    return true;
}

struct Tagged {
    uint32_t id;
    Bounds box;
};

// src/instance_bvh.rs:414-460: binary nodes, 16 words (child 0 box, child 1
// box, child tags), parents before children.
Tagged subtree(const std::vector<Bounds>& boxes, uint32_t* order, size_t n, std::vector<uint32_t>& nodes,
               bool sah) {
    if (n == 1) return Tagged{LEAF_BIT | order[0], boxes[order[0]]};
    Centre c{&boxes};
    Bounds cb = empty_bounds();
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            float v = c.at(order[k], a);
            cb.lo[a] = mn(cb.lo[a], v);
            cb.hi[a] = mx(cb.hi[a], v);
        }
    int axis = 0;
    uint32_t last = 0;
    float base = 0.0f, scale = 0.0f;
    size_t mid;
    if (sah && sah_split(boxes, order, n, c, cb, axis, last, base, scale)) {
        size_t left = 0;
        for (size_t k = 0; k < n; ++k) {
            uint32_t b = std::min(sat_index((c.at(order[k], axis) - base) * scale, 0xffffffffu), BINS - 1);
            if (b <= last) {
                std::swap(order[left], order[k]);
                left += 1;
            }
        }
        mid = left;
    } else {
        axis = 0;
        for (int a = 1; a < 3; ++a)
            if (total_key(cb.hi[a] - cb.lo[a]) >= total_key(cb.hi[axis] - cb.lo[axis])) axis = a;
        mid = n / 2;
        std::nth_element(order, order + mid, order + n, [&](uint32_t x, uint32_t y) {
            int32_t kx = total_key(c.at(x, axis)), ky = total_key(c.at(y, axis));
            return kx != ky ? kx < ky : x < y;
        });
    }
    size_t node = nodes.size() / TOP_BIN_WORDS;
    nodes.resize(nodes.size() + TOP_BIN_WORDS, 0u);
    Tagged a = subtree(boxes, order, mid, nodes, sah);
    Tagged b = subtree(boxes, order + mid, n - mid, nodes, sah);
    uint32_t* w = &nodes[node * TOP_BIN_WORDS];
    for (int k = 0; k < 3; ++k) {
        w[k] = f_bits(a.box.lo[k]);
        w[3 + k] = f_bits(a.box.hi[k]);
        w[6 + k] = f_bits(b.box.lo[k]);
        w[9 + k] = f_bits(b.box.hi[k]);
    }
    w[12] = a.id;
    w[13] = b.id;
    return Tagged{(uint32_t)node, merge(a.box, b.box)};
}

void write_child(uint32_t* node, int c, const Bounds& b) {
    for (int a = 0; a < 3; ++a) {
        node[a * 4 + c] = f_bits(b.lo[a]);
        node[12 + a * 4 + c] = f_bits(b.hi[a]);
    }
}

// src/instance_bvh.rs:370-401: preorder; each wide node takes its binary
// node's two children, then repeatedly opens the largest-area inner child
// (ties to the lower position) until it has four.
uint32_t collapse(const std::vector<uint32_t>& binary, uint32_t node, std::vector<uint32_t>& out) {
    auto child = [&](uint32_t n, int c) {
        const uint32_t* w = &binary[(size_t)n * TOP_BIN_WORDS];
        Tagged t;
        t.id = w[12 + c];
        for (int k = 0; k < 3; ++k) {
            t.box.lo[k] = bits_f(w[c * 6 + k]);
            t.box.hi[k] = bits_f(w[c * 6 + 3 + k]);
        }
        return t;
    };
    uint32_t index = (uint32_t)(out.size() / TOP_WORDS);
    out.resize(out.size() + TOP_WORDS, 0u);
    std::vector<Tagged> kids = {child(node, 0), child(node, 1)};
    while (kids.size() < 4) {
        int pick = -1;
        for (int i = 0; i < (int)kids.size(); ++i) {
            if (kids[i].id & LEAF_BIT) continue;
            if (pick < 0 || area(kids[i].box) > area(kids[pick].box)) pick = i;
        }
        if (pick < 0) break;
        uint32_t id = kids[pick].id;
        kids.erase(kids.begin() + pick);
        kids.insert(kids.begin() + pick, child(id, 1));
        kids.insert(kids.begin() + pick, child(id, 0));
    }
    for (int c = 0; c < 4; ++c) {
        size_t at = (size_t)index * TOP_WORDS;
        if (c >= (int)kids.size()) {
            out[at + 24 + c] = EMPTY;
            continue;
        }
        uint32_t id = kids[c].id;
        if (!(id & LEAF_BIT)) id = collapse(binary, id, out);
        write_child(&out[at], c, kids[c].box);
        out[at + 24 + c] = id;
    }
    return index;
}

uint32_t wide_depth(const std::vector<uint32_t>& nodes, uint32_t node) {
    size_t at = (size_t)node * TOP_WORDS;
    uint32_t m = 0;
    for (int c = 0; c < 4; ++c) {
        uint32_t id = nodes[at + 24 + c];
        if (id != EMPTY && !(id & LEAF_BIT)) m = std::max(m, wide_depth(nodes, id));
    }
    return 1 + m;
}

std::vector<uint32_t> build_with(const std::vector<Bounds>& boxes, bool sah) {
    std::vector<uint32_t> binary;
    binary.reserve(boxes.size() * TOP_BIN_WORDS);
    std::vector<uint32_t> order(boxes.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (uint32_t)i;
    subtree(boxes, order.data(), order.size(), binary, sah);
    std::vector<uint32_t> out;
    out.reserve(binary.size());
    collapse(binary, 0, out);
    if (sah && 1 + 3 * wide_depth(out, 0) > TOP_STACK) return build_with(boxes, false);
    return out;
}

Bounds refit_wide(uint32_t* nodes, uint32_t node, const Bounds* boxes) {
    size_t at = (size_t)node * TOP_WORDS;
    Bounds all = empty_bounds();
    for (int c = 0; c < 4; ++c) {
        uint32_t id = nodes[at + 24 + c];
        if (id == EMPTY) continue;
        Bounds b = (id & LEAF_BIT) ? boxes[id & ~LEAF_BIT] : refit_wide(nodes, id, boxes);
        write_child(&nodes[at], c, b);
        all = merge(all, b);
    }
    return all;
}

bool enabled(uint32_t count, int force) {
    bool on = force < 0 ? count >= MIN_INSTANCES : force > 0;
    return on && count > 0;
}

}  // namespace

// src/instance_bvh.rs:36-38
uint32_t top_reserve(uint32_t count) { return std::max<uint32_t>(count > 0 ? count - 1 : 0, 1) * TOP_WORDS; }

// src/instance_bvh.rs:346-366 (at least two boxes)
std::vector<uint32_t> top_build(const std::vector<Bounds>& boxes, bool sah) { return build_with(boxes, sah); }

// src/instance_bvh.rs:224-268 (wide)
void top_refit(uint32_t* nodes, size_t words, const Bounds* boxes, size_t count) {
    if (count < 2 || words < TOP_WORDS) return;
    refit_wide(nodes, 0, boxes);
}

// src/instance_bvh.rs:162-220, synchronous: reuse when the records are
// unchanged, refit up to REFITS times when only boxes moved, else rebuild.
void top_place(std::vector<uint32_t>& words, uint32_t header, uint32_t stride, uint32_t slots,
               const uint32_t* previous, size_t previous_words, int force) {
    size_t base = words.size();
    uint32_t count = words[0];
    bool on = enabled(count, force);
    if (on) {
        words[1] = (uint32_t)base;
        words[2] = count == 1 ? LEAF_BIT : 0u;
    }
    size_t reserve = top_reserve(slots);
    bool same_layout = previous && previous_words == base + reserve && previous[0] == words[0] &&
                       previous[1] == words[1] && previous[2] == words[2];
    if (same_layout && std::equal(words.begin() + 4, words.end(), previous + 4)) {
        words[3] = previous[3];
        words.insert(words.end(), previous + base, previous + previous_words);
        return;
    }
    words.resize(base + reserve, 0u);
    if (!on || count < 2) return;
    std::vector<Bounds> boxes(count);
    for (uint32_t i = 0; i < count; ++i) {
        size_t r = header + (size_t)i * stride;
        for (int k = 0; k < 3; ++k) {
            boxes[i].lo[k] = bits_f(words[r + 12 + k]);
            boxes[i].hi[k] = bits_f(words[r + 15 + k]);
        }
    }
    if (same_layout && previous[3] < REFITS) {
        words[3] = std::min(previous[3] + 1, REFITS);
        std::copy(previous + base, previous + base + reserve, words.begin() + base);
        top_refit(&words[base], reserve, boxes.data(), boxes.size());
        return;
    }
    words[3] = 0;
    std::vector<uint32_t> nodes = build_with(boxes, true);
    std::copy(nodes.begin(), nodes.end(), words.begin() + base);
}

}  // namespace rq
