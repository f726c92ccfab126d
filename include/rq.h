/*
 * rq: MI355X software BVH ray query (renderlab non-RT path), C ABI.
 *
 * The reference has no FFI: its host packs one u32 storage buffer
 * (src/gi_bvh.rs:151-274, src/instance_bvh.rs:162-220) and uploads it with
 * wgpu's queue.write_buffer (src/gi_cache.rs:4351-4361, 4481-4485); its device
 * side is the WGSL seam gi_closest / gi_visible / gi_visible_within
 * (shaders/gi_rays_software.wgsl:280-291). Each entry point below replaces one
 * of those pieces and cites it. Plain pointers and sizes only.
 *
 * Conventions
 *  - Every call returns rq_status; nothing aborts. rq_last_error() gives the
 *    calling thread's last message.
 *  - rq_scene owns all device memory it uploads; ray, hit and bit buffers are
 *    caller-owned DEVICE pointers unless a parameter says host.
 *  - Build functions run on the host and are re-entrant. One scene may be
 *    traced from several streams at once (read-only device data).
 *  - A ray with a non-finite origin or direction, a NaN or negative tmin, a
 *    NaN tmax or tmax > 1e30 is a miss with flag RQ_HIT_INVALID.
 */
#ifndef RQ_H
#define RQ_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rq_status {
    RQ_OK = 0,
    RQ_ERR_INVALID_INPUT = 1, /* reference: Error::InvalidInput (src/gi_bvh.rs:33-36, 236-241) */
    RQ_ERR_OUT_OF_MEMORY = 2,
    RQ_ERR_DEVICE = 3,        /* HIP error, or no GPU present */
    RQ_ERR_UNSUPPORTED = 4,
    RQ_ERR_BUDGET = 5,        /* serving: admission refused by the HBM budget */
    RQ_ERR_NOT_READY = 6,
    RQ_ERR_DEADLINE = 7       /* serving: admission refused by the latency budget */
} rq_status;

/* Device layouts of one scene. */
typedef enum rq_layout {
    RQ_LAYOUT_REF_BVH4Q8 = 0, /* the reference gi_geometry words, verbatim */
    RQ_LAYOUT_PLAIN_F32 = 1,  /* BVH4, child boxes pre-decoded to f32, triangles pre-gathered */
    RQ_LAYOUT_CW8_Q8 = 2,     /* 8-wide nodes, 8-bit outward-quantized child boxes, f32 triangles */
    RQ_LAYOUT_RS_BVH4Q8 = 3,  /* the reflection rays' rl_geometry words, verbatim (rs_trace) */
    RQ_LAYOUT_RS_CW8_Q8 = 4   /* rl_geometry's records and triangles over CW8 nodes: rs_trace's
                                 results (its closest hit does not depend on the visit order) */
} rq_layout;

/* gi_traverse modes (shaders/gi_rays_software.wgsl:214-216). */
typedef enum rq_mode {
    RQ_MODE_CLOSEST = 0,    /* gi_closest */
    RQ_MODE_ANY = 1,        /* gi_visible: any hit, sky visibility */
    RQ_MODE_ANY_OPAQUE = 2  /* gi_visible_within: opaque casters only */
} rq_mode;
/* Or'ed into the mode of rq_trace / rq_trace_indirect: trace a launch of at
   least 65536 rays in (direction octant, origin Morton code) order, for
   incoherent secondary rays. Results keep each ray's index and bits. */
#define RQ_REORDER 0x100u
/* Or'ed into the mode of rq_trace / rq_trace_indirect / rq_trace_noalloc
   (REF, PLAIN and CW8 layouts): the watertight leaf test (Woop, Benthin and
   Wald, JCGT 2013) in place of the reference's Moller-Trumbore
   (shaders/gi_rays_software.wgsl:41-58). A ray through an edge or vertex shared
   by two triangles hits at least one of them, where Moller-Trumbore can miss
   both; hits elsewhere are the same triangles, with t and (u, v) from the
   watertight arithmetic (within a few f32 roundings of Moller-Trumbore's). No
   counterpart in the reference: an opt-in mode, off by default. */
#define RQ_WATERTIGHT 0x200u

/* One ray, 32 bytes: origin, tmin, direction (need not be unit), tmax. */
typedef struct rq_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} rq_ray;

#define RQ_HIT_HIT 1u
#define RQ_HIT_INVALID 2u
/* GiHit (shaders/gi_rays_software.wgsl:19) as a buffer record, 48 bytes.
   t is tmax and inst 0xffffffff on a miss. tri holds the hit triangle's
   position indices, or (primitive, 0, 0) for closest hits on meshes built
   with keep_primitives (:147-148, 161-162). slot is the triangle's index in
   its mesh's BVH order (the FP64 verification reads it). */
typedef struct rq_hit {
    float t;
    uint32_t inst;
    uint32_t tri[3];
    float u, v;
    float normal[3];
    uint32_t flags;
    uint32_t slot;
} rq_hit;

/* One mesh: a triangle list. keep_primitives asks for the primitive map of
   Blas::with_primitives (src/gi_bvh.rs:82-87, 108-136). */
typedef struct rq_mesh {
    const float* positions; /* host, 3 floats per vertex */
    uint32_t vertex_count;
    const uint32_t* indices; /* host, 3 per triangle */
    uint32_t index_count;
    uint32_t keep_primitives;
} rq_mesh;

/* One placement (src/gi_cache.rs:4400-4431): object-to-world model matrix,
   column-major as glam's Mat4::to_cols_array. flags bit 0: casts no shadow. */
typedef struct rq_instance {
    float model[16];
    uint32_t mesh;
    uint32_t flags;
    uint32_t identity;
    uint32_t placement;
} rq_instance;

typedef struct rq_scene rq_scene;
typedef struct rq_bvh rq_bvh;

const char* rq_last_error(void);
/* Number of visible GPUs (0 on a host without one). */
int rq_device_count(void);

/* ---- BLAS build: crates/asset/src/near_field.rs:172-190 (build_bvh, build_bvh_sah) */
rq_status rq_bvh_build(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                       uint32_t index_count, uint32_t sah, rq_bvh** out);
/* near_field.rs:445-456 (refit_bvh): same topology, new positions. */
rq_status rq_bvh_refit(rq_bvh* bvh, const float* positions, uint32_t vertex_count);
/* Counts: [0] nodes, [1] leaves, [2] vertices, [3] indices, [4] root tag, [5] depth. */
rq_status rq_bvh_info(const rq_bvh* bvh, uint32_t info[6]);
/* Runtime records (src/gi_bvh.rs:191-216): 16 words a node, 2 a leaf, 3 a
   position, then the indices. Each output may be NULL. */
rq_status rq_bvh_records(const rq_bvh* bvh, uint32_t* nodes, uint32_t* leaves, float* positions,
                         uint32_t* indices);
void rq_bvh_destroy(rq_bvh* bvh);

/* ---- Top level: src/instance_bvh.rs:346-412 (build, build_with) over
   instance boxes (lo xyz, hi xyz per box). Writes up to reserve(count)
   words and returns the count written. sah=0 takes median splits. */
uint32_t rq_instance_bvh_reserve(uint32_t count);
rq_status rq_instance_bvh_build(const float* boxes, uint32_t count, uint32_t sah, uint32_t* nodes,
                                uint32_t* written);
/* src/instance_bvh.rs:224-275 (refit, refit_wide) */
rq_status rq_instance_bvh_refit(uint32_t* nodes, uint32_t node_words, const float* boxes,
                                uint32_t count);
/* src/instance_bvh.rs:162-220 (place) without the worker thread: appends the
   top level over `words`' records after `records_words` words (header
   `header`, `stride` words a record), sets header words 1-3, reusing or
   refitting `previous` (may be NULL) as the reference does. `out` holds
   records_words + reserve(slots) words. `force`: -1 automatic (>= 40
   instances), 0 never, 1 always. */
rq_status rq_instance_bvh_place(const uint32_t* words, uint32_t records_words, uint32_t header,
                                uint32_t stride, uint32_t slots, const uint32_t* previous,
                                uint32_t previous_words, int force, uint32_t* out);

/* ---- Scene: the whole gi_geometry buffer.
   rq_scene_build replaces set_indirect_geometry + prepare_gi_geometry
   (src/gi_cache.rs:3188-3196, 4246-4487): SAH BLAS per mesh, layout,
   mesh_words, instance_words, extra words 24-31 and place. */
rq_status rq_scene_build(const rq_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                         uint32_t instance_count, int top_level, rq_scene** out);
/* Adopts a reference-packed buffer verbatim (the bytes queue.write_buffer
   would upload, src/gi_cache.rs:4361). RQ_ERR_INVALID_INPUT when a mesh tree
   or the top level reaches a node twice, or an id, leaf or index points
   outside its array: traversal of such words would not end or would read
   outside the buffer. */
rq_status rq_scene_from_gi_words(const uint32_t* words, size_t word_count, rq_scene** out);
/* ---- Cooked geometry read from disk (crates/asset). The cluster BVH sidecar
   "RLB\0" (crates/asset/src/cluster_bvh/codec.rs:4-188, checked as
   Bvh::validate, crates/asset/src/cluster_bvh.rs:170-333) and the near field
   "RNF\0" (crates/asset/src/near_field.rs:540-680, NearField::validate
   :106-135). max_bytes 0 takes the reference's default limits (150,000,000 and
   160,000,000). Malformed bytes give RQ_ERR_INVALID_INPUT with the reference's
   message. Encoders write nothing and report the length when `capacity` is
   smaller. */
typedef struct rq_cluster_bvh rq_cluster_bvh;
typedef struct rq_near_field rq_near_field;
rq_status rq_cluster_bvh_decode(const uint8_t* bytes, size_t length, size_t max_bytes, rq_cluster_bvh** out);
rq_status rq_cluster_bvh_encode(const rq_cluster_bvh* bvh, size_t max_bytes, uint8_t* out, size_t capacity,
                                size_t* length);
/* near_field::build_bvh / build_bvh_sah (near_field.rs:172-262): the
   one-level, zero-error cluster BVH of a triangle list. */
rq_status rq_cluster_bvh_from_mesh(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                                   uint32_t index_count, int sah, rq_cluster_bvh** out);
/* nodes, leaves, levels, level leaves, positions, indices */
rq_status rq_cluster_bvh_counts(const rq_cluster_bvh* bvh, uint32_t counts[6]);
void rq_cluster_bvh_destroy(rq_cluster_bvh* bvh);
rq_status rq_near_field_decode(const uint8_t* bytes, size_t length, size_t max_bytes, rq_near_field** out);
rq_status rq_near_field_encode(const rq_near_field* field, size_t max_bytes, uint8_t* out, size_t capacity,
                               size_t* length);
/* counts: meshes, placements; settings: lod_error, reach, skip */
rq_status rq_near_field_info(const rq_near_field* field, uint32_t counts[2], uint32_t* bins, float settings[3],
                             uint8_t source[32]);
void rq_near_field_destroy(rq_near_field* field);

/* One mesh of a scene from exactly one source: a triangle list built at load
   (set_indirect_geometry), a cooked cluster BVH traced at level `index`
   (set_indirect_bvh, src/gi_cache.rs:3209-3217) or mesh `index` of a near
   field (set_indirect_near_field, :3218-3270). */
typedef struct rq_cooked_mesh {
    const rq_mesh* mesh;
    const rq_cluster_bvh* bvh;
    const rq_near_field* field;
    uint32_t index;
    uint32_t reserved;
} rq_cooked_mesh;
/* rq_scene_build over such meshes: near-field placements' masks packed after
   the meshes (src/gi_bvh.rs:151-226), their base in instance word 23 for every
   instance whose model has the cooked transform's exact bits (:95-104). All
   near fields with masks must share their bins. Cooked leaves may hold more
   than 16 triangles: such meshes trace with the REF layout only. */
rq_status rq_scene_build_cooked(const rq_cooked_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                                uint32_t instance_count, int top_level, rq_scene** out);
/* gi_near_blocked (shaders/gi_rays_software.wgsl:330-354) over a batch: whether
   the near-field masks of instance `inst` block direction d at the point
   (1 - u - v, u, v) of the triangle with position indices tri (every vertex
   with weight blocks it; `back`: the point is seen from behind). Bit i of
   blocked_bits (device, zeroed here) is set for a blocked query; an instance
   without masks blocks nothing. Needs the REF layout uploaded. */
typedef struct rq_near_query {
    uint32_t inst;
    uint32_t tri[3];
    float bary[2];
    uint32_t back;
    float d[3];
} rq_near_query;
rq_status rq_near_blocked(const rq_scene* scene, const rq_near_query* queries, uint32_t count,
                          uint32_t* blocked_bits, void* stream);

/* ---- The GI radiance cache's ray stages around the traversal, reading the
   cache's table words (`gi_table`, shaders/gi_cells.wgsl: 64-word header,
   24-word hot records, max(24, header[23])-word cold records) in device memory.
   A slot or item outside the table gives a ray with a NaN origin (a flagged
   miss: visible, no hit).

   gi_ray (shaders/gi_cache.wgsl:1563-1576): the visibility ray of each item
   (slot << 7 | k): origin q + n * gi_origin_offset(q, edge), direction k of
   the cell's hemisphere (gi_direction over gi_basis(n)), t in [0, 1e20]. */
rq_status rq_gi_item_rays(const uint32_t* table, size_t table_words, const uint32_t* items, uint32_t count,
                          rq_ray* rays, void* stream);
/* The gi_rays pass: the item rays into `workspace` (count rays), then
   gi_visible over them, the bit of item i set in staged_bits when its
   direction sees the sky (the reference's atomicOr into the staged words). */
rq_status rq_gi_rays(rq_scene* scene, rq_layout layout, const uint32_t* table, size_t table_words,
                     const uint32_t* items, uint32_t count, rq_ray* workspace, uint32_t* staged_bits,
                     void* stream);
/* gi_gather's fixed directions (shaders/gi_gather_software.wgsl:82-110) for
   the listed cells: directions phase + j * stride (stride = max(header[21], 1),
   phase = key & (stride - 1)) whose visibility bit is clear, as closest-hit
   rays with their items (slot << 7 | k), cell by cell in list order and
   directions ascending. *count (device memory) receives the total; rays past
   max_rays are not written. `workspace`: device memory of
   rq_gi_gather_workspace(slot_count) bytes. Trace them with rq_trace_indirect
   (count_dev = count). */
rq_status rq_gi_gather_rays(const uint32_t* table, size_t table_words, const uint32_t* slots, uint32_t slot_count,
                            rq_ray* rays, uint32_t* items, uint32_t max_rays, uint32_t* count, void* workspace,
                            size_t workspace_bytes, void* stream);
size_t rq_gi_gather_workspace(uint32_t slot_count);

/* rq_scene_build with every mesh's BVH built on the GPU: the same packed words
   (the SAH BVH4 of near_field::build_bvh_sah, crates/asset/src/near_field.rs:
   172-430, with its median fallback past depth 15; layout, instance records
   and top level as rq_scene_build), built level by level for all meshes at
   once, and the REF layout left uploaded on the current device. Same inputs
   and errors as rq_scene_build. Synchronizes `stream`. */
rq_status rq_scene_build_gpu(const rq_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                             uint32_t instance_count, int top_level, rq_scene** out, void* stream);
/* rq_scene_build_reflection with the BVHs built on the GPU (no reference
   counterpart, as rq_scene_build_gpu): the same rl_geometry words, from the
   GPU build of the meshes (trees, triangle order and primitive ids) with the
   reflection leaf tags and 12-word triangles written on the device; records,
   top level and scene bounds on the host. The words are left uploaded for
   RQ_LAYOUT_RS_BVH4Q8; the mesh words reach the host when a host reader asks. */
rq_status rq_scene_build_reflection_gpu(const rq_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                                        uint32_t instance_count, int top_level, rq_scene** out, void* stream);
/* GPU refit of one mesh of a built scene (rq_scene_build) whose REF layout is
   uploaded, for a re-skinned pose: replaces the skin_refit path of
   src/gi_cache.rs:2843-2863 -- near_field::refit_bvh (crates/asset/src/
   near_field.rs:445-479) on the mesh, its instances' world boxes recomputed
   (src/gi_bvh.rs instance_words) and instance_bvh::place over the previous
   top level (src/instance_bvh.rs:162-220: refit while only boxes moved,
   rebuild every 64 refits). `positions`: device memory, vertex_count * 3
   floats in the mesh's vertex order. The BVH is refit on the device (one
   launch per node height); the scene's host words follow, so the packed words
   equal what the CPU refit and repack give. PLAIN and CW8 copies are dropped
   (upload them again to trace with them). Same vertex count and finite
   positions, else RQ_ERR_INVALID_INPUT with nothing changed. Synchronizes
   `stream`. A mesh no instance draws is left alone. */
rq_status rq_scene_refit_mesh(rq_scene* scene, uint32_t mesh, const float* positions, uint32_t vertex_count,
                              void* stream);

/* ---- Reflection geometry: the rl_geometry buffer the non-RT reflection rays
   traverse (shaders/reflections_software_bvh4.wgsl rs_trace, binding 36).
   rq_scene_build_reflection replaces Blas::build + pack + instance_words +
   instance_bvh::place + finish_scene_bounds (src/reflection_software.rs:92-388):
   the same SAH BVH4 per mesh with its own leaf tags, 12-word triangles with
   source primitive ids, 24-word instance records (instances with a singular
   model are skipped), the top level and the scene bounds. Instance flags bit 0
   = casts no shadow; `placement` is the record's slot. Such a scene traces only
   with RQ_LAYOUT_RS_BVH4Q8, modes RQ_MODE_CLOSEST (rs_closest) and
   RQ_MODE_ANY (rs_blocked). Its hit records: inst = record index, tri =
   (primitive, 0, 0), u, v, t, slot = placement slot, normal zero. */
rq_status rq_scene_build_reflection(const rq_mesh* meshes, uint32_t mesh_count, const rq_instance* instances,
                                    uint32_t instance_count, int top_level, rq_scene** out);
/* Adopts reference-packed rl_geometry words verbatim (src/gi_cache.rs binding 36 upload).
   RQ_ERR_INVALID_INPUT when a tree reaches a node twice or a node, leaf or
   top-level leaf points outside the buffer or the records. */
rq_status rq_scene_from_rl_words(const uint32_t* words, size_t word_count, rq_scene** out);
/* The packed reference words, gi_geometry or (reflection scene) rl_geometry
   (host memory owned by the scene). */
rq_status rq_scene_gi_words(const rq_scene* scene, const uint32_t** words, size_t* word_count);
/* [0] instances, [1] meshes, [2] triangles (over instances), [3] unique triangles,
   [4] BLAS nodes, [5] top-level root tag, [6] max BLAS depth, [7] microseconds the
   build took (rq_scene_build / rq_scene_build_gpu; 0 for adopted words) */
rq_status rq_scene_stats(const rq_scene* scene, uint64_t stats[8]);
/* Uploads the requested layouts (bit 1<<rq_layout) to the current device. */
rq_status rq_scene_upload(rq_scene* scene, uint32_t layout_mask);
rq_status rq_scene_device_bytes(const rq_scene* scene, rq_layout layout, size_t* bytes);
/* Copies one part of an uploaded layout back to the host (no reference
   counterpart: a diagnostic for checking device images). PLAIN / CW8 parts as
   rq_scene_layout_image numbers them (0-4 and 6, 256-byte padded); RS_CW8 parts 0-2
   likewise, and its part 3 the root copies (layout_image part 7); REF / RS part 0 is the whole word buffer. *bytes = the part's size; nothing is copied when
   `capacity` is smaller. */
rq_status rq_scene_download_part(const rq_scene* scene, rq_layout layout, uint32_t part, void* dst,
                                 size_t capacity, size_t* bytes);
/* Host image of a relayout (PLAIN or CW8), built on demand without a device:
   part 0 instance records, 1 top-level nodes, 2 BLAS nodes, 3 triangles,
   4 position indices, 5 {top_on, top_root, instances, 0}, 6 the triangles'
   vertices (a, b, c: the watertight test's records), 7 (CW8) a copy of each
   instance's BLAS root node (24 words, zero for a leaf root). For inspection.
   RS_CW8 (a reflection scene): part 0 the words with each record's root
   replaced by its CW8 root, 1 the CW8 top level, 2 the CW8 BLAS nodes, 5 and 7
   as above; 3 and 4 are empty. */
rq_status rq_scene_layout_image(rq_scene* scene, rq_layout layout, uint32_t part, const void** data,
                                size_t* bytes);
void rq_scene_destroy(rq_scene* scene);

/* ---- The buffer form of a hit (shaders/primary_rays.wgsl:19; the reflection
   rays' RsHit.record, shaders/reflections_software_bvh4.wgsl:42, 90):
   vec4<u32>(placement + 1, primitive, bits(u), bits(v)), all zero on a miss.
   For gi_geometry scenes, placement is gi_hit_placement (instance word 29,
   shaders/gi_rays_software.wgsl:314-316) and primitive is gi_hit_primitive
   (:310-313: tri[0] when the instance has a primitive table, word 28 nonzero,
   else 0xffffffff); for rl_geometry scenes they are the hit's placement slot
   and primitive, as rs_closest returns them (:201). */
typedef struct rq_hit_record {
    uint32_t placement_plus1;
    uint32_t primitive;
    float u, v;
} rq_hit_record;
/* One hit on the host: its record and gi_hit_identity (:297-309: instance
   word 24, or for a merged world the identity + 1 of its placement from the
   position table at words 25-26; 0 for reflection scenes, misses and
   see-through records). identity may be NULL. */
rq_status rq_gi_hit_decode(const rq_scene* scene, const rq_hit* hit, uint32_t record[4], uint32_t* identity);
/* The same for `count` device hits into device arrays (records and identities
   may each be NULL): one kernel reading the scene's REF words (gi_geometry
   scenes need RQ_LAYOUT_REF_BVH4Q8 uploaded). */
rq_status rq_hit_records(const rq_scene* scene, const rq_hit* hits, uint32_t count, rq_hit_record* records,
                         uint32_t* identities, void* stream);

/* ---- Trace: gi_closest / gi_visible / gi_visible_within over a batch.
   rays, hits and vis_bits are device pointers; hits and vis_bits may each be
   NULL. vis_bits gets bit i set when ray i is unoccluded (gi_visible's
   result); the caller zeroes nothing, the call writes every word it covers.
   stream: a hipStream_t, or NULL for the default stream. */
rq_status rq_trace(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count,
                   uint32_t mode, rq_hit* hits, uint32_t* vis_bits, void* stream);

/* rq_trace without allocation. No counterpart in the reference: it is for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc). The ray cursor and
   the stack spill area come from one static device array of
   rq_trace_noalloc_scratch_bytes() = 112 MiB + 8 KiB per GPU (held until the
   process ends) instead of the stream's own memory, and the persistent grid
   is capped to the 131,072 lanes that array holds; the ray count is not
   limited. RQ_REORDER is refused (RQ_ERR_UNSUPPORTED): its sort space grows
   with the ray count. Results equal rq_trace's. Calls that overlap on one GPU
   (two streams or host threads) share the array and race: order them. */
rq_status rq_trace_noalloc(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count, uint32_t mode,
                           rq_hit* hits, uint32_t* vis_bits, void* stream);
size_t rq_trace_noalloc_scratch_bytes(void);

/* rq_trace with the ray count read on the device at launch time:
   min(*count_dev, max_count) rays (count_dev: a device word written by an
   earlier launch on the same stream, such as rq_gen_secondary's counts), so a
   frame's secondary traces are enqueued without a host round trip. */
rq_status rq_trace_indirect(rq_scene* scene, rq_layout layout, const rq_ray* rays,
                            const uint32_t* count_dev, uint32_t max_count, uint32_t mode, rq_hit* hits,
                            uint32_t* vis_bits, void* stream);

/* One ray batch of a multi-scene launch. */
typedef struct rq_batch {
    uint32_t scene;  /* index into the scenes array */
    uint32_t mode;   /* rq_mode */
    uint32_t count;
    uint32_t pad;
    const rq_ray* rays;  /* device */
    rq_hit* hits;        /* device, may be NULL */
    uint32_t* vis_bits;  /* device, may be NULL */
} rq_batch;
/* Every batch of every scene in ONE persistent launch: lanes take rays from
   the concatenated batches and traverse their batch's scene in its mode. All
   scenes must have `layout` uploaded. No counterpart in the reference (its
   callers dispatch per pass); results equal per-batch rq_trace calls. */
rq_status rq_trace_batches(rq_scene* const* scenes, uint32_t scene_count, rq_layout layout,
                           const rq_batch* batches, uint32_t batch_count, void* stream);

/* Up to RQ_MAX_RANGES ray ranges of ONE scene, each with its own mode and
   outputs, in ONE persistent launch: the passes of a frame that do not
   depend on each other (sun shadow and GI bounces) share one pool of rays,
   so one launch's drain overlaps the other's work. Uses the closest-hit
   tuning of `layout`. No counterpart in the reference (its passes dispatch
   separately); results equal one rq_trace call per range. */
#define RQ_MAX_RANGES 4u
typedef struct rq_range {
    const rq_ray* rays;  /* device */
    uint32_t count;
    uint32_t mode;       /* rq_mode (RQ_REORDER is not accepted here) */
    rq_hit* hits;        /* device, may be NULL */
    uint32_t* vis_bits;  /* device, may be NULL; cleared first */
} rq_range;
rq_status rq_trace_ranges(rq_scene* scene, rq_layout layout, const rq_range* ranges, uint32_t range_count,
                          void* stream);

/* The same traversal with a tally of the global bytes it moves (rays,
   records, nodes, triangles, stack spills, hit records): the algorithmic
   bytes of one launch for the roofline. Synchronous; not for timing. */
rq_status rq_trace_count_bytes(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count,
                               uint32_t mode, rq_hit* hits, uint32_t* vis_bits, uint64_t* bytes);
/* The same launch with its step statistics: [0] bytes, [1] wave iterations,
   [2] node steps, [3] triangle-pair steps, [4] instance entries, [5] refills;
   summed over wave iterations: [6] lanes holding a live ray, [7] [8] [9]
   lanes wanting a node / triangle-pair / instance step; [10] [11] wave
   iterations running node / triangle-pair steps, [12] node iterations in which
   some lane entered an instance, [13] speculative node steps, [14] drain
   iterations, [15] top-level node steps. */
rq_status rq_trace_count_stats(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count,
                               uint32_t mode, rq_hit* hits, uint32_t* vis_bits, uint64_t stats[16]);

/* Launch parameters of the traversal kernels (the autotuner's knobs). */
typedef struct rq_tuning {
    uint32_t block;        /* threads per workgroup: 64, 128, 256 */
    uint32_t stack;        /* LDS stack entries per lane: 8 .. 48 */
    uint32_t persistent;   /* 0: one lane per ray; 1: persistent, per-lane refill */
    uint32_t waves_per_eu; /* occupancy hint 0 (none), 1 .. 8 */
    uint32_t grid_per_cu;  /* persistent workgroups per CU (0: auto) */
    uint32_t refill;       /* idle lanes a wave gathers before it takes new rays: 1 .. 64 */
    uint32_t speculate;    /* CW8: lanes with triangles pending join the wave's node steps */
    uint32_t drain;        /* once the rays run out, every live lane steps each iteration */
} rq_tuning;
rq_status rq_get_tuning(rq_layout layout, uint32_t mode, rq_tuning* out);
rq_status rq_set_tuning(rq_layout layout, uint32_t mode, const rq_tuning* tuning);
/* The same knobs for rq_trace_batches launches. */
rq_status rq_get_batch_tuning(rq_layout layout, rq_tuning* out);
rq_status rq_set_batch_tuning(rq_layout layout, const rq_tuning* tuning);

/* ---- FP64 verification: the only FP64 on the GPU. For each hit ray,
   re-intersects its reported triangle in f64 (object-space ray computed in
   f64 from the f32 world ray) against the original positions (needs the REF
   layout uploaded). With S the largest magnitude the f32 computation rounds
   (the world-to-object terms |r_k . o| + |t_k|, the vertex coordinates) plus
   |t64|*|d|, out (host):
   [0] max |t-t64|/|t64|, [1] max |u-u64|, |v-v64|, [2] hits checked,
   [3] max |t-t64|*|d| / S: the distance error against the size of the
       numbers it came from, which bounds f32 error for short hits far from
       the origin where [0] cannot,
   [4] hits with [0]'s ratio above 1e-5,
   [5] over hits away from grazing (cosine to the normal > 0.01) and slivers
       (edge-angle sine > 1e-3): max of [3]'s ratio * cosine / 2^-24, the
       error in f32 roundings of S,
   [6] hits counted in [5], [7] zero. */
rq_status rq_verify_fp64(rq_scene* scene, const rq_ray* rays, const rq_hit* hits, uint32_t count,
                         double out[8], void* stream);

/* ---- Ray generation (device). Formulas: shaders/primary_rays.wgsl:50-66
   (primary, reversed Z), shaders/reference_path_tracer.wgsl:149-167,
   237-256, 466-470 (offset, sun shadow, cosine bounce, pcg). */
/* inverse: column-major inverse view-projection (host). */
rq_status rq_gen_primary(const float inverse[16], uint32_t width, uint32_t height, rq_ray* rays,
                         void* stream);
/* The same rays in screen-tile order (no reference counterpart: a coherence
   option for callers that keep their own pixel mapping): ray i is pixel
   i % 64 of 8x8 tile i / 64, tiles and the pixels in a tile row-major; width
   and height must be multiples of 8. 8% faster primary traversal on the 1080p
   frame (a wave's rays leave through one tile). */
rq_status rq_gen_primary_tiled(const float inverse[16], uint32_t width, uint32_t height, rq_ray* rays,
                               void* stream);
/* Secondary rays of `count` primary rays and their closest hits: compacted
   in pixel order into shadow (one per lit hit) and gi (n_gi per hit, one
   array per bounce of `count` rays each, each compacted); counts (device,
   n_gi+1 words) receive the lengths. */
rq_status rq_gen_secondary(const rq_ray* primary, const rq_hit* hits, uint32_t count,
                           const float sun[3], uint32_t seed, uint32_t n_gi, rq_ray* shadow,
                           rq_ray* gi, uint32_t* counts, void* stream);

/* ---- Serving: render-query jobs admitted by an HBM budget and a latency
   budget, traced with continuous batching. No counterpart in the reference
   (its renderer traces once a frame on one queue; its admission is a bounded
   wait, src/engine/mod.rs:1247-1262); this is the request loop around the ray
   query. */
typedef struct rq_job {
    uint32_t scene;       /* index into the scenes array */
    uint32_t scene_kind;  /* camera orbit of rq_gen_camera (0 or 1) */
    uint32_t width, height;
    uint32_t n_gi;        /* 0: primary + shadow; 1: primary + shadow + GI */
    uint32_t pad;
    uint64_t seed;        /* camera pose and GI sampling seed */
    double arrival_s;     /* arrival time, seconds after the run starts */
} rq_job;
typedef struct rq_serve_config {
    rq_layout layout;
    uint32_t slots;       /* jobs admitted and not finished at once, 1..256 */
    uint64_t hbm_budget;  /* bytes of job buffers in flight at once */
    float sun[3];         /* unit direction toward the sun */
    uint32_t latency_budget_us; /* 0: none; else a job whose predicted completion is later than its
                                   arrival plus this is refused (RQ_ERR_DEADLINE) */
} rq_serve_config;
typedef struct rq_job_result {
    double arrival_s, start_s, end_s; /* host clock, seconds after the run starts */
    double gpu_ms;                    /* first to last launch of the job, HIP events */
    uint64_t digest;                  /* order-free hash of the job's hits and visibility bits */
    uint64_t primary, secondary;      /* rays traced */
    uint32_t status;                  /* RQ_OK; RQ_ERR_BUDGET: the job alone exceeds the HBM budget;
                                         RQ_ERR_DEADLINE: refused by the latency budget */
    uint32_t pad;
} rq_job_result;
/* Runs `jobs` in real time. A job is admitted once it has arrived, a slot is
   free and its buffers fit in the budget (first come, first served; first fit
   in one arena of hbm_budget bytes), unless its predicted completion (the
   running launch, the next launch with its primary rays, the one after with
   its secondary rays; each a fixed drain cost plus rays over the measured
   throughput) misses arrival + latency_budget_us: then it is refused, on
   arrival or while it waits. One persistent launch (rq_trace_batches, CW8
   speculation off) runs at a time, carrying the primary rays of the jobs
   admitted since the previous launch and the shadow and GI rays of the jobs
   whose primary rays that launch traced. Returns when every job is done. */
rq_status rq_serve(rq_scene* const* scenes, uint32_t scene_count, const rq_serve_config* config,
                   const rq_job* jobs, uint32_t job_count, rq_job_result* results);

/* ---- Device helpers (HIP runtime of this library; no torch types). */
rq_status rq_malloc(size_t bytes, void** ptr);
rq_status rq_free(void* ptr);
rq_status rq_memcpy_h2d(void* dst, const void* src, size_t bytes);
rq_status rq_memcpy_d2h(void* dst, const void* src, size_t bytes);
rq_status rq_memset(void* dst, int value, size_t bytes);
/* A plain streaming copy kernel (bytes a multiple of 16): the known-size
   reference the roofline's HBM byte counters are scaled by. */
rq_status rq_device_copy(void* dst, const void* src, size_t bytes, void* stream);
rq_status rq_synchronize(void);
rq_status rq_stream_create(void** stream);
rq_status rq_stream_destroy(void* stream); /* also frees the stream's scratch memory */
/* Sizes a stream's scratch memory (ray cursor, stack spill area, secondary
   ray counts) for launches of up to max_rays rays, so later launches on it
   allocate nothing. */
rq_status rq_stream_reserve(void* stream, uint32_t max_rays);
rq_status rq_stream_synchronize(void* stream);
rq_status rq_event_create(void** event);
rq_status rq_event_destroy(void* event);
rq_status rq_event_record(void* event, void* stream);
/* Work queued on `stream` after this call waits until `event` completes
   (hipStreamWaitEvent): how two streams order their launches. */
rq_status rq_stream_wait_event(void* stream, void* event);
rq_status rq_event_synchronize(void* event);
rq_status rq_event_elapsed_ms(void* start, void* stop, float* ms);
rq_status rq_event_query(void* event); /* RQ_OK when complete, RQ_ERR_NOT_READY otherwise */

/* ---- Seeded synthetic scenes (the benchmark configurations).
   kind 0: boxes + UV spheres + ground quad (2,000 boxes, 8 spheres);
   kind 1: lognormal-edge triangle clusters in a 200 m cube plus ground,
           `meshes` meshes and about `triangles` triangles in all. */
rq_status rq_gen_scene(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles,
                       rq_scene** out);
/* rq_gen_scene with the BVHs built by rq_scene_build_gpu (the same words). */
rq_status rq_gen_scene_gpu(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles, rq_scene** out,
                           void* stream);
/* rq_gen_reflection_scene with the BVHs built on the GPU (the same words). */
rq_status rq_gen_reflection_scene_gpu(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles,
                                      rq_scene** out, void* stream);
/* The same meshes and placements packed as reflection geometry (rl_geometry). */
rq_status rq_gen_reflection_scene(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles,
                                  rq_scene** out);
/* Camera on a seeded orbit around the scene: column-major inverse of the
   reversed-Z view-projection (60 degree vertical field of view). */
rq_status rq_gen_camera(uint32_t kind, uint64_t seed, uint32_t width, uint32_t height,
                        float inverse[16]);

#ifdef __cplusplus
}
#endif
#endif
