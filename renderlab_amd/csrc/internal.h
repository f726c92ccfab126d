// Host-side data structures shared by the builder, the packers, the layouts and
// the device launchers. Names follow the reference's record vocabulary (nodes,
// leaves, positions, indices, instance records, top level).
#pragma once
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rq.h"

namespace rq {

constexpr uint32_t LEAF_BIT = 0x80000000u;  // crates/asset/src/cluster_bvh.rs:12
constexpr uint32_t EMPTY = 0xffffffffu;     // crates/asset/src/cluster_bvh.rs:13
constexpr uint32_t LEAF_TRIANGLES = 8;      // crates/asset/src/near_field.rs:23
constexpr uint32_t MAX_SAH_DEPTH = 15;      // crates/asset/src/near_field.rs:178
constexpr uint32_t HEADER_WORDS = 16;       // src/gi_bvh.rs:19
constexpr uint32_t INSTANCE_WORDS = 32;     // src/gi_bvh.rs:25
constexpr uint32_t NODE_WORDS = 16;         // src/gi_bvh.rs:26
constexpr uint32_t TOP_WORDS = 32;          // src/instance_bvh.rs:29 (wide node)
constexpr uint32_t TOP_BIN_WORDS = 16;      // src/instance_bvh.rs:28 (binary node)
constexpr uint32_t TOP_STACK = 48;          // src/instance_bvh.rs:32
constexpr uint32_t MIN_INSTANCES = 40;      // src/instance_bvh.rs:72
constexpr uint32_t REFITS = 64;             // src/instance_bvh.rs:94
// PLAIN/CW8 triangle records: vertex a and the f32 edges b - a, c - a (36 B);
// the side record read once per hit: position indices, BVH slot, primitive id.
constexpr uint32_t TRI_FLOATS = 9;
constexpr uint32_t TRI_ID_WORDS = 6;
// Reflection geometry (src/reflection_software.rs:26-44).
constexpr int LAYOUT_COUNT = 5;  // rq_layout values
constexpr uint32_t RS_HEADER_WORDS = 8;
constexpr uint32_t RS_RECORD_WORDS = 24;
constexpr uint32_t RS_NODE_WORDS = 16;
constexpr uint32_t RS_TRI_WORDS = 12;
constexpr uint32_t RS_TOP_NODE_WORDS = 32;
// RS leaf tags are LEAF | first triangle << 4 | count: a mesh's triangles must stay below 2^27.
constexpr uint32_t RS_MAX_MESH_TRIS = 1u << 27;
// Scene kinds: the GI rays' gi_geometry words, or the reflection rays' rl_geometry words.
constexpr int RQ_SCENE_GI = 0;
constexpr int RQ_SCENE_REFLECTION = 1;

inline float bits_f(uint32_t w) { float f; std::memcpy(&f, &w, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t w; std::memcpy(&w, &f, 4); return w; }
// Keep-the-first-on-ties min/max (Rust's f32::min/max agree on every non-zero,
// non-NaN pair; the builder never sees NaN).
inline float mn(float a, float b) { return b < a ? b : a; }
inline float mx(float a, float b) { return b > a ? b : a; }
// f32::total_cmp as an integer key.
inline int32_t total_key(float f) {
    int32_t k = (int32_t)f_bits(f);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}
// `as usize` of an f32: saturating, NaN to 0.
inline uint32_t sat_index(float f, uint32_t cap) {
    if (!(f > 0.0f)) return 0;
    if (f >= (float)cap) return cap;
    return (uint32_t)f;
}

// The runtime BVH4 node record: exactly the 16 words of src/gi_bvh.rs:191-202
// (origin, step, four tagged children, axis-major byte planes; child c is byte c).
struct Node {
    float origin[3];
    float step[3];
    uint32_t children[4];
    uint8_t lower[3][4];
    uint8_t upper[3][4];
};
static_assert(sizeof(Node) == 64, "node record is 16 words");

struct Bvh {
    std::vector<Node> nodes;
    std::vector<uint32_t> leaf_offset;  // index units
    std::vector<uint32_t> leaf_count;   // index units
    std::vector<uint32_t> leaf_first;   // each leaf's first source triangle (a cooked leaf's source offset / 3)
    std::vector<float> positions;       // 3 per vertex
    std::vector<uint32_t> indices;
    uint32_t root = 0;
};

struct Bounds {
    float lo[3];
    float hi[3];
};
inline Bounds empty_bounds() {
    const float inf = __builtin_inff();
    return Bounds{{inf, inf, inf}, {-inf, -inf, -inf}};
}
inline Bounds merge(const Bounds& a, const Bounds& b) {
    Bounds r;
    for (int i = 0; i < 3; ++i) {
        r.lo[i] = mn(a.lo[i], b.lo[i]);
        r.hi[i] = mx(a.hi[i], b.hi[i]);
    }
    return r;
}

// builder.cpp
bool build_bvh(const float* positions, uint32_t vertex_count, const uint32_t* indices,
               uint32_t index_count, bool sah, Bvh& out, std::string& err);
bool build_bvh_sah(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                   uint32_t index_count, Bvh& out, std::string& err);
bool refit_bvh(Bvh& bvh, const float* positions, uint32_t vertex_count, std::string& err);
uint32_t bvh_depth(const Bvh& bvh, uint32_t child);
void decode_child(const Node& n, int slot, bool fused, float lo[3], float hi[3]);

// toplevel.cpp
uint32_t top_reserve(uint32_t count);
std::vector<uint32_t> top_build(const std::vector<Bounds>& boxes, bool sah);
void top_refit(uint32_t* nodes, size_t words, const Bounds* boxes, size_t count);
void top_place(std::vector<uint32_t>& words, uint32_t header, uint32_t stride, uint32_t slots,
               const uint32_t* previous, size_t previous_words, int force);

// pack.cpp
// One cooked placement's near-field masks and the exact model bits they were
// cooked under (src/gi_bvh.rs:38-43).
struct NearMasks {
    float model[16];
    std::vector<uint32_t> masks;
};
// One mesh's runtime BVH as packed, its bounds over all positions, its primitive
// table (empty unless kept) and its cooked placements' masks.
struct Blas {
    Bvh bvh;
    float lo[3], hi[3];
    std::vector<uint32_t> primitives;
    std::vector<NearMasks> near;
    size_t words() const {
        return bvh.nodes.size() * NODE_WORDS + bvh.leaf_offset.size() * 2 + bvh.positions.size() +
               bvh.indices.size() + primitives.size();
    }
};
rq_status pack_scene(std::vector<Blas>& blas, const rq_instance* instances, uint32_t instance_count,
                     int top_level, uint32_t near_bins, std::chrono::steady_clock::time_point began,
                     rq_scene** out);
void blas_bounds(Blas& b);
bool map_primitives(const std::vector<uint32_t>& bvh_indices, const uint32_t* source, uint32_t n,
                    std::vector<uint32_t>& out);
bool mat4_inverse(const float m[16], float inv[16]);
void scene_bounds(rq_scene* scene);
// The packed words' length, on the host or (GPU-built, not yet fetched) the device.
inline size_t scene_words(const rq_scene* s);
// Fetches the stale ranges of the host words from the device copy (REF, or RS
// for a reflection scene).
rq_status sync_host_words(const rq_scene* scene);
// Fetches the stale host words inside [lo, hi) only.
rq_status sync_host_range(const rq_scene* scene, uint64_t lo, uint64_t hi);
// Instance record words 12-17: the padded world box of a mesh box under a model
// (src/gi_bvh.rs instance_words).
void instance_box(const float model[16], const float lo[3], const float hi[3], uint32_t out[6]);
// reflect.cpp: words of the reflection header, records, top level and scene bounds
uint32_t rs_records_words(uint32_t instances);
// world_bounds (src/reflection_software.rs:372-388): an instance's padded world box.
void rs_world_bounds(const float* M, const float lo_in[3], const float hi_in[3], float lo[3], float hi[3]);

// Per-mesh facts recovered from an instance record of packed words.
struct MeshView {
    uint32_t nodes, leaves, positions, indices;  // word offsets (record words 18-21)
    uint32_t node_count, leaf_count, vertex_count, index_count;
};
bool mesh_view(const std::vector<uint32_t>& words, uint32_t record, MeshView& out, std::string& err);

// Device-side layouts (device.hip); opaque here.
struct DeviceLayout {
    bool ready = false;
    void* base = nullptr;  // one allocation
    size_t bytes = 0;
    // word/element offsets into `base`, layout specific
    uint64_t off[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t count[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // PLAIN and CW8: per distinct mesh (node word offset in the reference
    // words, first node, first triangle), for in-place refreshes after a refit.
    std::vector<uint32_t> mesh_base;
};

// One mesh's GPU refit plan: its node ids grouped by height (device), cached.
struct RefitPlan {
    uint32_t nodes = 0;                  // the mesh's node word offset (its key)
    void* order = nullptr;               // device: node ids, lowest height first
    void* boxes = nullptr;               // device: the mesh's leaf boxes (6 floats a leaf)
    std::vector<uint32_t> level_start;   // slice of `order` per height
};

struct HostLayout {
    // Plain and CW8 host images (uploaded verbatim).
    std::vector<uint32_t> inst;    // 32 words per instance
    std::vector<uint32_t> top;     // top-level wide nodes (32 words)
    std::vector<uint32_t> nodes;   // BLAS nodes (plain: 32 words, cw8: 20 words)
    std::vector<uint32_t> mesh_base;  // (node word offset, first node, first triangle) per mesh
    std::vector<float> tris;       // TRI_FLOATS per triangle: a, b - a, c - a
    std::vector<uint32_t> tri_ids; // TRI_ID_WORDS per triangle: ia, ib, ic, slot, primitive, 0
    std::vector<float> verts;      // TRI_FLOATS per triangle: the vertices a, b, c (watertight tests)
    std::vector<uint32_t> roots;   // CW8: per instance a copy of its BLAS root node (24 words; zero for a leaf root)
    uint32_t top_root = 0;
    uint32_t top_on = 0;
    uint32_t instance_count = 0;
};

bool build_plain(const std::vector<uint32_t>& words, HostLayout& out, std::string& err);
bool build_cw8(const std::vector<uint32_t>& words, const HostLayout& plain, HostLayout& out,
               std::string& err);
// RQ_LAYOUT_RS_CW8_Q8 from rl_geometry words: `inst` holds the words with each
// record's root (word 20) replaced by its mesh's CW8 root, `nodes` the CW8 BLAS
// nodes, `top` the CW8 top level.
bool build_rs_cw8(const std::vector<uint32_t>& words, HostLayout& out, std::string& err);

void set_error(const std::string& msg);

// The traversal's per-lane stack spill area: words per lane of a persistent grid.
constexpr uint32_t TRACE_SPILL_WORDS = 224;
// The persistent kernels' ray cursor: the launch's 64-ray chunks are dealt
// round robin to CURSOR_PARTS partitions, each with its own claim counter,
// CURSOR_STRIDE words (256 bytes) apart so that no two share a cache line.
constexpr uint32_t CURSOR_PARTS = 32;
constexpr uint32_t CURSOR_STRIDE = 64;
constexpr size_t CURSOR_BYTES = (size_t)CURSOR_PARTS * CURSOR_STRIDE * sizeof(uint32_t);
// Device memory a trace launch uses in place of its stream's own (no allocation):
// the ray cursor (CURSOR_BYTES, cleared by the launch) and the spill area of `lanes` lanes.
struct FixedScratch {
    uint32_t* counter;
    uint32_t* spill;
    size_t lanes;
};
// rq_trace_batches with the CW8 node-step speculation forced on (1) or off (0),
// or as the batch tuning says (-1).
rq_status trace_batches_impl(rq_scene* const* scenes, uint32_t scene_count, rq_layout layout, const rq_batch* batches,
                             uint32_t batch_count, void* stream, int speculate);
// rq_trace / rq_trace_indirect / rq_trace_count_bytes, and rq_trace_noalloc with `fixed`.
rq_status launch_trace(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count, uint32_t mode,
                       rq_hit* hits, uint32_t* vis_bits, void* stream, unsigned long long* bytes,
                       const uint32_t* count_ptr = nullptr, const FixedScratch* fixed = nullptr);

}  // namespace rq

struct rq_bvh {
    rq::Bvh bvh;
};

struct rq_scene {
    int kind = rq::RQ_SCENE_GI;
    std::vector<uint32_t> words;  // the reference gi_geometry (or rl_geometry) words
    float bounds[6] = {0, 0, 0, 1, 1, 1};  // union of the instance world boxes (ray reordering keys)
    uint64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    rq::HostLayout plain, cw8, rs_cw8;
    rq::DeviceLayout dev[rq::LAYOUT_COUNT];  // by rq_layout
    std::vector<rq::RefitPlan> refit_plans;
    // Kept by rq_scene_build for refits: each instance's model (16 floats,
    // column-major), each mesh's node word offset, the top-level switch.
    std::vector<float> models;
    std::vector<uint32_t> mesh_nodes;
    // GPU-built scenes: per mesh, 9 words -- node, leaf, position, index and
    // primitive-table word offsets (0: none), node count, leaf count, triangles,
    // root tag.
    std::vector<uint32_t> mesh_map;
    int top_force = -1;
    uint32_t near_bins = 0;  // octahedral bins per side of the near-field masks (0: none)
    void* refit_scratch = nullptr;  // device: per-block bounds partials of a refit
    // Word ranges (offset, count) where the device REF words are newer than
    // `words` (a refit wrote them on the device); sync_host_words fetches them.
    mutable std::vector<std::pair<uint64_t, uint64_t>> stale;
    // A GPU-built scene keeps only its header, records and top level on the host
    // until a host reader asks: `words` then grows to total_words (0: words.size()).
    mutable size_t total_words = 0;
    int device = -1;
};

inline size_t rq::scene_words(const rq_scene* s) { return s->total_words > s->words.size() ? s->total_words : s->words.size(); }
