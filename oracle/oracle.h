/*
 * CPU oracle for the software BVH ray query. TEST INFRASTRUCTURE ONLY: the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load this
 * library as the checker. The product library (renderlab_amd/librq.so) never
 * links, loads or calls it.
 *
 * What it restates (reference files, read as text):
 *   orc_gi_traverse      shaders/gi_rays_software.wgsl:217-279 (gi_traverse),
 *                        :98-207 (gi_instance), :27-70 (helpers), with
 *                        shaders/scene.wgsl:86-90 (unit_normal).
 *   wt_tri_test          (mode ORC_WATERTIGHT) the watertight test of Woop,
 *                        Benthin and Wald (JCGT 2013) in place of gi_triangle:
 *                        the engine's opt-in mode, pinned by watertightness
 *                        and closed forms rather than by the reference.
 *   orc_flat_f32         the flat instance loop of the same file (:225-228)
 *                        with every triangle of every instance tested in leaf
 *                        order: the tree-free answer gi_traverse must equal.
 *   orc_brute_f64        crates/asset/src/cluster_bvh/reference.rs:121-158,
 *                        184-194 (f64 Moller-Trumbore, exact det==0 rejection,
 *                        lexicographic tie rule) over every instance.
 *   orc_reintersect_f64  the same f64 test applied to one reported hit.
 *   orc_rs_trace         shaders/reflections_software_bvh4.wgsl:163-203 (rs_trace),
 *                        :50-141 (rs_instance), :19-47 (helpers): the reflection
 *                        rays over rl_geometry (src/reflection_software.rs).
 *   orc_rs_brute_f64     the same geometry and slack in f64 over every triangle.
 *   orc_validate_blas    crates/asset/src/cluster_bvh.rs:170-333 (Bvh::validate)
 *                        on one mesh's runtime records.
 *   orc_primary_ray      shaders/primary_rays.wgsl:50-66 (primary_ray).
 *   orc_secondary_rays   shaders/reference_path_tracer.wgsl:149-167, 237-256,
 *                        466-470 (pcg, rand, basis, offset, cosine bounce),
 *                        built from orc_pcg, orc_basis, orc_offset and
 *                        orc_cosine_dir, each exported so it can be pinned alone.
 *   orc_hit_record       shaders/primary_rays.wgsl:19 (the 16-byte hit record)
 *                        and gi_hit_identity / gi_hit_primitive /
 *                        gi_hit_placement (shaders/gi_rays_software.wgsl:297-316).
 *
 * Arithmetic convention (fixed here because WGSL leaves it to the driver):
 * no contraction (built with -ffp-contract=off), IEEE division and square
 * root, min/max are IEEE 754-2019 minimum/maximum (-0 < +0), dot products
 * sum x, y, then z.
 */
#ifndef RQ_ORACLE_H
#define RQ_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} orc_ray;

/* Same 48-byte shape as rq_hit in include/rq.h. */
typedef struct orc_hit {
    float t;
    uint32_t inst;
    uint32_t tri[3];
    float u, v;
    float n[3];
    uint32_t flags; /* bit 0 hit, bit 1 invalid input */
    uint32_t slot;  /* BVH-order triangle index inside the hit mesh */
} orc_hit;

/* Visit and load counters of one traversal (words = u32 loads). */
typedef struct orc_counts {
    uint64_t words;
    uint64_t top_nodes;
    uint64_t inst_tests;
    uint64_t inst_entries;
    uint64_t blas_nodes;
    uint64_t leaves;
    uint64_t tri_steps;
    uint64_t tri_tests;
} orc_counts;

/* Or'ed into `mode` of orc_gi_traverse(_batch) and orc_flat_f32: the
   watertight leaf test (wt_tri_test in oracle.c) in place of Moller-Trumbore. */
#define ORC_WATERTIGHT 0x200u

/* One ray through gi_traverse. Returns 0, or -1 if a read left the buffer. */
int orc_gi_traverse(const uint32_t* g, size_t n, const orc_ray* ray, uint32_t mode,
                    orc_hit* hit, orc_counts* counts);
/* A batch on `threads` host threads (0: all). hits and vis_bits may be NULL;
   vis_bits gets bit i set when ray i is unoccluded (gi_visible's sense).
   counts (may be NULL) receives the sum over rays. Returns failures. */
int orc_gi_traverse_batch(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count,
                          uint32_t mode, orc_hit* hits, uint32_t* vis_bits,
                          orc_counts* counts, uint32_t threads);
int orc_flat_f32(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count,
                 uint32_t mode, orc_hit* hits);
/* The reflection rays' rs_trace over rl_geometry words (record: inst = record
   index, tri[0] = primitive, u, v, t, slot = placement slot, normal zero). */
int orc_rs_trace(const uint32_t* g, size_t n, const orc_ray* ray, uint32_t mode, orc_hit* hit,
                 orc_counts* counts);
int orc_rs_trace_batch(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count,
                       uint32_t mode, orc_hit* hits, uint32_t* vis_bits, orc_counts* counts,
                       uint32_t threads);

typedef struct orc_hit64 {
    double t, u, v;
    uint32_t inst;
    uint32_t slot;
    uint32_t hit;
    uint32_t pad;
} orc_hit64;
int orc_brute_f64(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count,
                  uint32_t mode, orc_hit64* hits, uint32_t threads);
/* f64 re-intersection of each reported hit (inst, slot); misses get t=-1. */
int orc_reintersect_f64(const uint32_t* g, size_t n, const orc_ray* rays, const orc_hit* hits,
                        uint32_t count, orc_hit64* out);
/* f64 brute force over rl_geometry words with edge slack `slack`; slot = primitive. */
int orc_rs_brute_f64(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                     double slack, orc_hit64* hits, uint32_t threads);

/* Bvh::validate on one mesh of the packed buffer (node/leaf/position/index word
   offsets, root tag, counts). Returns 0 when valid, else a negative code. */
int orc_validate_blas(const uint32_t* g, size_t n, uint32_t nodes_at, uint32_t node_count,
                      uint32_t leaves_at, uint32_t leaf_count, uint32_t positions_at,
                      uint32_t vertex_count, uint32_t indices_at, uint32_t index_count,
                      uint32_t root, uint32_t* depth_out);

void orc_primary_ray(const float inverse[16], uint32_t px, uint32_t py, uint32_t w, uint32_t h,
                     orc_ray* out);
/* Secondary rays of one primary hit: out[0] shadow toward `sun`, out[1..n_gi]
   cosine bounces. valid bit k of the return value says out[k] is a ray. */
uint32_t orc_secondary_rays(const orc_ray* primary, const orc_hit* hit, const float sun[3],
                            uint32_t pixel, uint32_t seed, uint32_t n_gi, orc_ray* out);
uint32_t orc_pcg(uint32_t v);
void orc_pcg_batch(const uint32_t* in, uint32_t count, uint32_t* out);
/* rand() `count` times from `state`: out[i] in [0, 1), a multiple of 2^-24. */
void orc_rand_sequence(uint32_t state, uint32_t count, float* out);
/* basis(n): columns c0, c1, c2 = n at out[0..2], out[3..5], out[6..8]. */
void orc_basis(const float n[3], float out[9]);
/* The cosine-weighted bounce direction about n for the random pair (u1, u2). */
void orc_cosine_dir(const float n[3], float u1, float u2, float out[3]);
/* offset(p, ng): the shadow and bounce origin. */
void orc_offset(const float p[3], const float ng[3], float out[3]);

/* Hit records (4 words per hit, may be NULL) and gi_hit_identity (may be NULL)
   of `count` hits over gi_geometry words, or rl_geometry hits when
   `reflection` is nonzero. */
void orc_hit_records(const uint32_t* g, size_t n, int reflection, const orc_hit* hits, uint32_t count,
                     uint32_t* records, uint32_t* identities);

/* The GI radiance cache's rays from its table words (shaders/gi_cells.wgsl
   layout; tests/ara_backend/gi_cache_reference.rs restates it on the CPU):
   orc_gi_item_rays: gi_ray's ray for each item (slot << 7 | k), NaN origin
   for an item outside the table (shaders/gi_cache.wgsl:1563-1576, 1048-1068);
   orc_gi_gather_rays: gi_gather's fixed directions of each listed cell whose
   visibility bit is clear, cell by cell, k ascending; returns the total and
   writes at most max_rays (shaders/gi_gather_software.wgsl:82-110). */
/* Near-field masks: the octahedral bin of a direction (gi_near_bin,
   shaders/gi_rays_software.wgsl:330-337; near_field::bin,
   crates/asset/src/near_field.rs:139-145), the direction through a bin's
   centre (near_field::bin_direction, :147-158), and gi_near_blocked
   (shaders/gi_rays_software.wgsl:338-354) read from packed gi words (instance
   word 23: the mask base; 0 blocks nothing). */
uint32_t orc_near_bin(const float d[3], uint32_t bins);
void orc_near_bin_direction(uint32_t bin, uint32_t bins, float d[3]);
int orc_near_blocked(const uint32_t* g, size_t n, uint32_t bins, uint32_t inst, const uint32_t tri[3],
                     const float bary[2], int back, const float d[3]);
void orc_gi_item_rays(const uint32_t* table, size_t words, const uint32_t* items, uint32_t count,
                      orc_ray* rays);
uint32_t orc_gi_gather_rays(const uint32_t* table, size_t words, const uint32_t* slots, uint32_t n,
                            orc_ray* rays, uint32_t* items, uint32_t max_rays);

#ifdef __cplusplus
}
#endif
#endif
