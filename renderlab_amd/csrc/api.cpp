// Error reporting and the host-only C entry points of the BLAS builder and the
// top level (no device calls here).
#include <algorithm>
#include <cstring>

#include "internal.h"

namespace rq {
namespace {
thread_local std::string last_error;
}
void set_error(const std::string& msg) { last_error = msg; }
}  // namespace rq

using namespace rq;

extern "C" const char* rq_last_error(void) { return last_error.c_str(); }

extern "C" rq_status rq_bvh_build(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                                  uint32_t index_count, uint32_t sah, rq_bvh** out) {
    if (!out || (!positions && vertex_count)) {
        set_error("rq_bvh_build: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    *out = nullptr;
    auto b = new rq_bvh();
    std::string err;
    bool ok = sah ? build_bvh_sah(positions, vertex_count, indices, index_count, b->bvh, err)
                  : build_bvh(positions, vertex_count, indices, index_count, false, b->bvh, err);
    if (!ok) {
        delete b;
        set_error(err);
        return RQ_ERR_INVALID_INPUT;
    }
    *out = b;
    return RQ_OK;
}

extern "C" rq_status rq_bvh_refit(rq_bvh* bvh, const float* positions, uint32_t vertex_count) {
    if (!bvh || !positions) {
        set_error("rq_bvh_refit: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    std::string err;
    if (!refit_bvh(bvh->bvh, positions, vertex_count, err)) {
        set_error(err);
        return RQ_ERR_INVALID_INPUT;
    }
    return RQ_OK;
}

extern "C" rq_status rq_bvh_info(const rq_bvh* bvh, uint32_t info[6]) {
    if (!bvh || !info) {
        set_error("rq_bvh_info: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    const Bvh& b = bvh->bvh;
    info[0] = (uint32_t)b.nodes.size();
    info[1] = (uint32_t)b.leaf_offset.size();
    info[2] = (uint32_t)(b.positions.size() / 3);
    info[3] = (uint32_t)b.indices.size();
    info[4] = b.root;
    info[5] = bvh_depth(b, b.root);
    return RQ_OK;
}

extern "C" rq_status rq_bvh_records(const rq_bvh* bvh, uint32_t* nodes, uint32_t* leaves, float* positions,
                                    uint32_t* indices) {
    if (!bvh) {
        set_error("rq_bvh_records: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    const Bvh& b = bvh->bvh;
    if (nodes && !b.nodes.empty()) std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(Node));
    if (leaves)
        for (size_t l = 0; l < b.leaf_offset.size(); ++l) {
            leaves[2 * l] = b.leaf_offset[l];
            leaves[2 * l + 1] = b.leaf_count[l];
        }
    if (positions && !b.positions.empty())
        std::memcpy(positions, b.positions.data(), b.positions.size() * sizeof(float));
    if (indices && !b.indices.empty()) std::memcpy(indices, b.indices.data(), b.indices.size() * 4);
    return RQ_OK;
}

extern "C" void rq_bvh_destroy(rq_bvh* bvh) { delete bvh; }

extern "C" uint32_t rq_instance_bvh_reserve(uint32_t count) { return top_reserve(count); }

extern "C" rq_status rq_instance_bvh_build(const float* boxes, uint32_t count, uint32_t sah, uint32_t* nodes,
                                           uint32_t* written) {
    if (!boxes || !nodes || count < 2) {
        set_error("rq_instance_bvh_build: needs at least two boxes");
        return RQ_ERR_INVALID_INPUT;
    }
    std::vector<Bounds> b(count);
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            b[i].lo[k] = boxes[6ull * i + k];
            b[i].hi[k] = boxes[6ull * i + 3 + k];
        }
    std::vector<uint32_t> out = top_build(b, sah != 0);
    std::copy(out.begin(), out.end(), nodes);
    if (written) *written = (uint32_t)out.size();
    return RQ_OK;
}

extern "C" rq_status rq_instance_bvh_refit(uint32_t* nodes, uint32_t node_words, const float* boxes,
                                           uint32_t count) {
    if (!nodes || !boxes) {
        set_error("rq_instance_bvh_refit: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    std::vector<Bounds> b(count);
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            b[i].lo[k] = boxes[6ull * i + k];
            b[i].hi[k] = boxes[6ull * i + 3 + k];
        }
    top_refit(nodes, node_words, b.data(), b.size());
    return RQ_OK;
}

extern "C" rq_status rq_instance_bvh_place(const uint32_t* words, uint32_t records_words, uint32_t header,
                                           uint32_t stride, uint32_t slots, const uint32_t* previous,
                                           uint32_t previous_words, int force, uint32_t* out) {
    if (!words || !out || records_words < 4 || header < 4) {
        set_error("rq_instance_bvh_place: bad arguments");
        return RQ_ERR_INVALID_INPUT;
    }
    uint32_t count = words[0];
    if ((uint64_t)header + (uint64_t)count * stride > records_words || stride < 18) {
        set_error("rq_instance_bvh_place: records past the end");
        return RQ_ERR_INVALID_INPUT;
    }
    std::vector<uint32_t> w(words, words + records_words);
    top_place(w, header, stride, slots, previous, previous_words, force);
    std::copy(w.begin(), w.end(), out);
    return RQ_OK;
}

extern "C" rq_status rq_scene_gi_words(const rq_scene* scene, const uint32_t** words, size_t* word_count) {
    if (!scene || !words || !word_count) {
        set_error("rq_scene_gi_words: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    rq_status st = sync_host_words(scene);
    if (st != RQ_OK) return st;
    *words = scene->words.data();
    *word_count = scene->words.size();
    return RQ_OK;
}

extern "C" rq_status rq_scene_stats(const rq_scene* scene, uint64_t stats[8]) {
    if (!scene || !stats) {
        set_error("rq_scene_stats: null argument");
        return RQ_ERR_INVALID_INPUT;
    }
    std::memcpy(stats, scene->stats, sizeof scene->stats);
    return RQ_OK;
}

extern "C" rq_status rq_scene_layout_image(rq_scene* scene, rq_layout layout, uint32_t part, const void** data,
                                           size_t* bytes) {
    const bool rs = layout == RQ_LAYOUT_RS_CW8_Q8;
    if (!scene || !data || !bytes || !((int)layout == 1 || (int)layout == 2 || rs) || part > 7 ||
        scene->kind != (rs ? RQ_SCENE_REFLECTION : RQ_SCENE_GI)) {
        set_error("rq_scene_layout_image: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    HostLayout& h = rs ? scene->rs_cw8 : layout == RQ_LAYOUT_PLAIN_F32 ? scene->plain : scene->cw8;
    std::string err;
    rq_status st = sync_host_words(scene);
    if (st != RQ_OK) return st;
    if (rs) {
        if (h.inst.empty() && !build_rs_cw8(scene->words, h, err)) {
            set_error("rs cw8 relayout: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
    } else if (h.inst.empty() && scene->words[0] > 0) {
        if (scene->plain.inst.empty() && !build_plain(scene->words, scene->plain, err)) {
            set_error("plain relayout: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
        if (layout == RQ_LAYOUT_CW8_Q8 && !build_cw8(scene->words, scene->plain, scene->cw8, err)) {
            set_error("cw8 relayout: " + err);
            return RQ_ERR_INVALID_INPUT;
        }
    }
    switch (part) {
        case 0: *data = h.inst.data(); *bytes = h.inst.size() * 4; break;
        case 1: *data = h.top.data(); *bytes = h.top.size() * 4; break;
        case 2: *data = h.nodes.data(); *bytes = h.nodes.size() * 4; break;
        case 3: *data = h.tris.data(); *bytes = h.tris.size() * 4; break;
        case 4: *data = h.tri_ids.data(); *bytes = h.tri_ids.size() * 4; break;
        case 6: *data = h.verts.data(); *bytes = h.verts.size() * 4; break;
        case 7: *data = h.roots.data(); *bytes = h.roots.size() * 4; break;
        default: {
            static thread_local uint32_t meta[4];
            meta[0] = h.top_on;
            meta[1] = h.top_root;
            meta[2] = h.instance_count;
            meta[3] = 0;
            *data = meta;
            *bytes = sizeof meta;
        }
    }
    return RQ_OK;
}
