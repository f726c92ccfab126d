/*
 * CPU oracle for the software BVH ray query -- TEST INFRASTRUCTURE, not product
 * code. Only the tests, the smoke check and the benchmark's CPU-baseline leg
 * load it, and only as the checker. Reference citations are per function; the
 * arithmetic convention is stated in oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>
#include <unistd.h>

#define LEAF 0x80000000u
#define EMPTY 0xffffffffu
#define STACK 48u
#define IW 32u
#define FMAXV 3.4028235e38f

/* ---------------------------------------------------------------- scalars */
static inline float bf(uint32_t w) { float f; memcpy(&f, &w, 4); return f; }
static inline uint32_t fb(float f) { uint32_t w; memcpy(&w, &f, 4); return w; }
/* IEEE 754-2019 minimum / maximum: NaN propagates, -0 orders below +0. */
static inline float mn(float a, float b) {
    if (a != a || b != b) return a + b;
    if (a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}
static inline float mx(float a, float b) {
    if (a != a || b != b) return a + b;
    if (a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}

typedef struct v3 { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 sub3(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* --------------------------------------------------------- counted loads */
typedef struct ctx {
    const uint32_t* g;
    size_t n;
    orc_counts c;
    int bad;
    int wt;  /* the watertight leaf test (mode bit ORC_WATERTIGHT) */
} ctx;
static inline uint32_t W(ctx* x, uint64_t i) {
    x->c.words++;
    if (i >= x->n) { x->bad = 1; return 0; }
    return x->g[i];
}
static inline float F(ctx* x, uint64_t i) { return bf(W(x, i)); }
static inline v3 GV(ctx* x, uint64_t i) { return mk(F(x, i), F(x, i + 1), F(x, i + 2)); }

/* gi_rays_software.wgsl:27-30 */
static inline float safe_inv1(float d) {
    float tiny = d >= 0.0f ? 1e-30f : -1e-30f;
    return 1.0f / (fabsf(d) < 1e-30f ? tiny : d);
}
static inline v3 safe_inverse(v3 d) { return mk(safe_inv1(d.x), safe_inv1(d.y), safe_inv1(d.z)); }

/* The watertight mode's culling (Ize, "Robust BVH Ray Traversal", JCGT 2(2),
   2013): a box's exit distance scaled by 1 + 4 * 2^-23 >= 1 + 2 gamma_3 before
   near <= far, conservative against the roundings of (b - o) * inv. */
#define ROBUST_FAR 1.00000048f

/* gi_rays_software.wgsl:35-40; robust: the exit distance scaled by ROBUST_FAR. */
static inline float slab_near_r(v3 o, v3 inv, v3 lo, v3 hi, float tmin, float tmax, int robust) {
    v3 a = mul3(sub3(lo, o), inv);
    v3 b = mul3(sub3(hi, o), inv);
    float near = mx(mx(mx(mn(a.x, b.x), mn(a.y, b.y)), mn(a.z, b.z)), tmin);
    float exit = mn(mn(mx(a.x, b.x), mx(a.y, b.y)), mx(a.z, b.z));
    if (robust) exit = exit * ROBUST_FAR;
    float far = mn(exit, tmax);
    return near <= far ? near : FMAXV;
}
static inline float slab_near(v3 o, v3 inv, v3 lo, v3 hi, float tmin, float tmax) {
    return slab_near_r(o, inv, lo, hi, tmin, tmax, 0);
}

/* gi_rays_software.wgsl:42-58: Moller-Trumbore, both faces; t<0 is a miss. */
static inline v3 tri_test(v3 o, v3 d, v3 a, v3 b, v3 c, float tmin, float tmax) {
    v3 none = mk(-1.0f, 0.0f, 0.0f);
    v3 e1 = sub3(b, a), e2 = sub3(c, a);
    v3 p = cross3(d, e2);
    float det = dot3(e1, p);
    if (fabsf(det) < 1e-30f) return none;
    float inv = 1.0f / det;
    v3 s = sub3(o, a);
    float u = dot3(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return none;
    v3 q = cross3(s, e1);
    float v = dot3(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return none;
    float t = dot3(e2, q) * inv;
    if (t > tmin && t < tmax) return mk(t, u, v);
    return none;
}

/* Kahan's difference of products a*b - c*d: within 2 ulps of the exact value
   (Jeannerod, Louvet and Muller, Math. Comp. 82, 2013), so its sign is exact
   and it is zero exactly when a*b == c*d. */
static inline float diff_of_products(float a, float b, float c, float d) {
    float w = c * d;
    float e = fmaf(-c, d, w);
    float f = fmaf(a, b, -w);
    return f + e;
}
static inline float comp(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

/* Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight
   Ray/Triangle Intersection", JCGT 2(1), 2013), both faces, in f32: translate
   to the ray origin, shear along the dominant axis (kz the largest |d|, the
   lowest axis on ties; kx, ky the next two, swapped when d[kz] < 0), the three
   edge functions by diff_of_products in place of the paper's f64 retry, reject
   mixed signs and det = U + V + W == 0, t = T / det with T = U*Az + V*Bz + W*Cz
   summed in that order, and (u, v) = (V, W) / det, the weights of b and c. Not
   the reference's test (gi_rays_software.wgsl:41-58 is Moller-Trumbore): the
   opt-in watertight mode of this engine. t < 0 is a miss. */
static inline v3 wt_tri_test(v3 o, v3 d, v3 a, v3 b, v3 c, float tmin, float tmax) {
    v3 none = mk(-1.0f, 0.0f, 0.0f);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) { int k = kx; kx = ky; ky = k; }
    float sx = comp(d, kx) / dz, sy = comp(d, ky) / dz, sz = 1.0f / dz;
    v3 A = sub3(a, o), B = sub3(b, o), C = sub3(c, o);
    float az_ = comp(A, kz), bz_ = comp(B, kz), cz_ = comp(C, kz);
    float Ax = comp(A, kx) - sx * az_, Ay = comp(A, ky) - sy * az_;
    float Bx = comp(B, kx) - sx * bz_, By = comp(B, ky) - sy * bz_;
    float Cx = comp(C, kx) - sx * cz_, Cy = comp(C, ky) - sy * cz_;
    float U = diff_of_products(Cx, By, Cy, Bx);
    float V = diff_of_products(Ax, Cy, Ay, Cx);
    float W = diff_of_products(Bx, Ay, By, Ax);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return none;
    float det = U + V + W;
    if (det == 0.0f) return none;
    float T = U * (sz * az_) + V * (sz * bz_) + W * (sz * cz_);
    float inv = 1.0f / det;
    float t = T * inv;
    if (!(t > tmin && t < tmax)) return none;
    return mk(t, V * inv, W * inv);
}

/* shaders/scene.wgsl:86-90 */
static inline v3 unit_normal(v3 n) {
    float scale = mx(mx(fabsf(n.x), fabsf(n.y)), fabsf(n.z));
    float s = mx(scale, 1e-30f);
    v3 sc = mk(n.x / s, n.y / s, n.z / s);
    float k = 1.0f / sqrtf(mx(dot3(sc, sc), 1e-30f));
    return mk(sc.x * k, sc.y * k, sc.z * k);
}

typedef struct best_t {
    int hit;
    float t;
    v3 n;
    uint32_t inst;
    uint32_t tri[3];
    float u, v;
    uint32_t slot;
} best_t;

/* gi_rays_software.wgsl:68-70 */
static inline float limit_of(const best_t* b, uint32_t i) {
    return (b->hit && i < b->inst) ? bf(fb(b->t) + 1u) : b->t;
}

/* Hit normal: mat3x3(r0,r1,r2) * cross(b-a, c-a) (gi_rays_software.wgsl:142). */
static inline v3 hit_normal(v3 r0, v3 r1, v3 r2, v3 a, v3 b, v3 c) {
    v3 x = cross3(sub3(b, a), sub3(c, a));
    v3 n = mk(r0.x * x.x + r1.x * x.y + r2.x * x.z, r0.y * x.x + r1.y * x.y + r2.y * x.z,
              r0.z * x.x + r1.z * x.y + r2.z * x.z);
    return unit_normal(n);
}

/* gi_rays_software.wgsl:98-207 (opaque instances; see-through off). */
static int gi_instance(ctx* x, uint32_t i, v3 origin, v3 dir, v3 world_inv, float tmin, int any_hit,
                       best_t* best, uint32_t* stack) {
    uint64_t r = 16u + (uint64_t)i * IW;
    x->c.inst_tests++;
    if (any_hit && (W(x, r + 27) & 1u) != 0u) return 0;
    if (!(slab_near_r(origin, world_inv, GV(x, r + 12), GV(x, r + 15), tmin, limit_of(best, i), x->wt) <=
          limit_of(best, i)))
        return 0;
    x->c.inst_entries++;
    v3 r0 = GV(x, r), r1 = GV(x, r + 4), r2 = GV(x, r + 8);
    float w0 = F(x, r + 3), w1 = F(x, r + 7), w2 = F(x, r + 11);
    v3 o = mk(dot3(r0, origin) + w0 * 1.0f, dot3(r1, origin) + w1 * 1.0f, dot3(r2, origin) + w2 * 1.0f);
    v3 d = mk(dot3(r0, dir), dot3(r1, dir), dot3(r2, dir));
    v3 inv = safe_inverse(d);
    uint32_t nodes = W(x, r + 18), leaves = W(x, r + 19);
    uint32_t positions = W(x, r + 20), indices = W(x, r + 21);
    uint32_t prim_base = W(x, r + 28);
    stack[0] = W(x, r + 22);
    uint32_t sp = 1;
    while (sp > 0 && !x->bad) {
        sp -= 1;
        uint32_t child = stack[sp];
        if (child & LEAF) {
            x->c.leaves++;
            uint64_t l = leaves + (uint64_t)(child & ~LEAF) * 2u;
            uint32_t first = W(x, l), count = W(x, l + 1);
            for (uint32_t k = 0; k < count && !x->bad; k += 6) {
                x->c.tri_steps++;
                uint64_t xi = (uint64_t)indices + first + k;
                int two = k + 3 < count;
                uint64_t yi = two ? xi + 3 : xi;
                uint32_t ia = W(x, xi), ib = W(x, xi + 1), ic = W(x, xi + 2);
                uint32_t ja = W(x, yi), jb = W(x, yi + 1), jc = W(x, yi + 2);
                v3 a = GV(x, positions + (uint64_t)ia * 3u), b = GV(x, positions + (uint64_t)ib * 3u),
                   c = GV(x, positions + (uint64_t)ic * 3u);
                v3 a2 = GV(x, positions + (uint64_t)ja * 3u), b2 = GV(x, positions + (uint64_t)jb * 3u),
                   c2 = GV(x, positions + (uint64_t)jc * 3u);
                x->c.tri_tests++;
                v3 h = x->wt ? wt_tri_test(o, d, a, b, c, tmin, limit_of(best, i))
                             : tri_test(o, d, a, b, c, tmin, limit_of(best, i));
                if (h.x >= 0.0f) {
                    best->hit = 1; best->t = h.x; best->u = h.y; best->v = h.z; best->inst = i;
                    best->n = hit_normal(r0, r1, r2, a, b, c);
                    best->tri[0] = ia; best->tri[1] = ib; best->tri[2] = ic;
                    best->slot = (first + k) / 3u;
                    if (!any_hit && prim_base != 0u) {
                        best->tri[0] = W(x, prim_base + (uint64_t)(first + k) / 3u);
                        best->tri[1] = 0; best->tri[2] = 0;
                    }
                    if (any_hit) return 1;
                }
                if (two) {
                    x->c.tri_tests++;
                    v3 h2 = x->wt ? wt_tri_test(o, d, a2, b2, c2, tmin, limit_of(best, i))
                                  : tri_test(o, d, a2, b2, c2, tmin, limit_of(best, i));
                    if (h2.x >= 0.0f) {
                        best->hit = 1; best->t = h2.x; best->u = h2.y; best->v = h2.z; best->inst = i;
                        best->n = hit_normal(r0, r1, r2, a2, b2, c2);
                        best->tri[0] = ja; best->tri[1] = jb; best->tri[2] = jc;
                        best->slot = (first + k) / 3u + 1u;
                        if (!any_hit && prim_base != 0u) {
                            best->tri[0] = W(x, prim_base + (uint64_t)(first + k) / 3u + 1u);
                            best->tri[1] = 0; best->tri[2] = 0;
                        }
                        if (any_hit) return 1;
                    }
                }
            }
            continue;
        }
        x->c.blas_nodes++;
        uint64_t nb = nodes + (uint64_t)child * 16u;
        v3 oq = GV(x, nb), st = GV(x, nb + 3);
        float t[4] = {FMAXV, FMAXV, FMAXV, FMAXV};
        uint32_t c[4] = {W(x, nb + 6), W(x, nb + 7), W(x, nb + 8), W(x, nb + 9)};
        uint32_t lx = W(x, nb + 10), ly = W(x, nb + 11), lz = W(x, nb + 12);
        uint32_t hx = W(x, nb + 13), hy = W(x, nb + 14), hz = W(x, nb + 15);
        for (uint32_t s = 0; s < 4; ++s) {
            if (c[s] == EMPTY) continue;
            uint32_t sh = 8u * s;
            v3 lq = mk((float)((lx >> sh) & 0xffu), (float)((ly >> sh) & 0xffu), (float)((lz >> sh) & 0xffu));
            v3 hq = mk((float)((hx >> sh) & 0xffu), (float)((hy >> sh) & 0xffu), (float)((hz >> sh) & 0xffu));
            v3 lo = mk(oq.x + st.x * lq.x, oq.y + st.y * lq.y, oq.z + st.z * lq.z);
            v3 hi = mk(oq.x + st.x * hq.x, oq.y + st.y * hq.y, oq.z + st.z * hq.z);
            t[s] = slab_near_r(o, inv, lo, hi, tmin, limit_of(best, i), x->wt);
        }
        /* Sorting network on (bits(t), child), descending (:189-197). */
        uint32_t k0[2] = {fb(t[0]), c[0]}, k1[2] = {fb(t[1]), c[1]}, k2[2] = {fb(t[2]), c[2]},
                 k3[2] = {fb(t[3]), c[3]}, tmp[2];
#define SWP(A, B) do { memcpy(tmp, A, 8); memcpy(A, B, 8); memcpy(B, tmp, 8); } while (0)
        if (k0[0] < k1[0]) SWP(k0, k1);
        if (k2[0] < k3[0]) SWP(k2, k3);
        if (k0[0] < k2[0]) SWP(k0, k2);
        if (k1[0] < k3[0]) SWP(k1, k3);
        if (k1[0] < k2[0]) SWP(k1, k2);
#undef SWP
        uint32_t* e[4] = {k0, k1, k2, k3};
        for (int s = 0; s < 4; ++s) {
            if (bf(e[s][0]) <= limit_of(best, i) && sp < STACK) stack[sp++] = e[s][1];
        }
    }
    return 0;
}

static inline int ray_invalid(const orc_ray* r) {
    for (int a = 0; a < 3; ++a)
        if (!isfinite(r->o[a]) || !isfinite(r->d[a])) return 1;
    if (!isfinite(r->tmin) || r->tmin < 0.0f || r->tmax != r->tmax || r->tmax > 1e30f) return 1;
    return 0;
}

static void write_hit(const best_t* b, orc_hit* h) {
    h->t = b->t;
    h->inst = b->inst;
    memcpy(h->tri, b->tri, sizeof h->tri);
    h->u = b->u; h->v = b->v;
    h->n[0] = b->n.x; h->n[1] = b->n.y; h->n[2] = b->n.z;
    h->flags = b->hit ? 1u : 0u;
    h->slot = b->hit ? b->slot : 0u;
}

static void init_best(best_t* b, float tmax) {
    memset(b, 0, sizeof *b);
    b->t = tmax;
    b->inst = 0xffffffffu;
}

/* gi_rays_software.wgsl:217-279 (GI_TOP_WIDE, see-through off). */
int orc_gi_traverse(const uint32_t* g, size_t n, const orc_ray* ray, uint32_t mode, orc_hit* hit,
                    orc_counts* counts) {
    ctx x;
    memset(&x, 0, sizeof x);
    x.g = g; x.n = n;
    x.wt = (mode & ORC_WATERTIGHT) != 0u;
    mode &= 0xffu;
    best_t best;
    init_best(&best, ray->tmax);
    if (ray_invalid(ray)) {
        write_hit(&best, hit);
        hit->flags = 2u;
        if (counts) *counts = x.c;
        return 0;
    }
    int any_hit = mode != 0u;
    uint32_t stack[STACK];
    v3 origin = mk(ray->o[0], ray->o[1], ray->o[2]);
    v3 dir = mk(ray->d[0], ray->d[1], ray->d[2]);
    float tmin = ray->tmin;
    v3 world_inv = safe_inverse(dir);
    uint32_t instances = W(&x, 0), top = W(&x, 1);
    if (top == 0u) {
        for (uint32_t i = 0; i < instances && !x.bad; ++i)
            if (gi_instance(&x, i, origin, dir, world_inv, tmin, any_hit, &best, stack)) break;
    } else {
        uint32_t nodes[STACK];
        nodes[0] = W(&x, 2);
        uint32_t sp = 1;
        while (sp > 0 && !x.bad) {
            sp -= 1;
            uint32_t item = nodes[sp];
            if (item & LEAF) {
                if (gi_instance(&x, item & ~LEAF, origin, dir, world_inv, tmin, any_hit, &best, stack))
                    break;
                continue;
            }
            x.c.top_nodes++;
            uint64_t nb = top + (uint64_t)item * 32u;
            float bound = best.hit ? bf(fb(best.t) + 1u) : best.t;
            float dist[4];
            uint32_t ids[4];
            for (int l = 0; l < 4; ++l) {
                float ax = (F(&x, nb + l) - origin.x) * world_inv.x;
                float bx = (F(&x, nb + 12 + l) - origin.x) * world_inv.x;
                float ay = (F(&x, nb + 4 + l) - origin.y) * world_inv.y;
                float by = (F(&x, nb + 16 + l) - origin.y) * world_inv.y;
                float az = (F(&x, nb + 8 + l) - origin.z) * world_inv.z;
                float bz = (F(&x, nb + 20 + l) - origin.z) * world_inv.z;
                float near = mx(mx(mx(mn(ax, bx), mn(ay, by)), mn(az, bz)), tmin);
                float exit = mn(mn(mx(ax, bx), mx(ay, by)), mx(az, bz));
                if (x.wt) exit = exit * ROBUST_FAR;
                float far = mn(exit, bound);
                ids[l] = W(&x, nb + 24 + l);
                dist[l] = (near <= far && ids[l] != EMPTY) ? near : FMAXV;
            }
#define SW(A, B) do { float td = dist[A]; dist[A] = dist[B]; dist[B] = td; \
                      uint32_t ti = ids[A]; ids[A] = ids[B]; ids[B] = ti; } while (0)
            if (dist[1] < dist[0]) SW(0, 1);
            if (dist[3] < dist[2]) SW(2, 3);
            if (dist[2] < dist[0]) SW(0, 2);
            if (dist[3] < dist[1]) SW(1, 3);
            if (dist[2] < dist[1]) SW(1, 2);
#undef SW
            for (int l = 3; l >= 0; --l)
                if (dist[l] <= bound && sp < STACK) nodes[sp++] = ids[l];
        }
    }
    write_hit(&best, hit);
    if (x.bad) hit->flags |= 4u;
    if (counts) *counts = x.c;
    return x.bad ? -1 : 0;
}

/* ------------------------------------------- reflection rays (rs_trace) */
/* shaders/reflections_software_bvh4.wgsl, over the rl_geometry words packed by
   src/reflection_software.rs: 8-word header, 24-word instance records, the GI
   BVH4 node records with leaf tags LEAF | first << 4 | count, 12-word
   triangles (a, e1, e2, primitive). */
#define RS_HEADER 8u
#define RS_REC 24u
#define RS_TRI 12u
static const float RS_EDGE = 4e-6f; /* :19 */

typedef struct rs_best {
    int hit;
    float t;
    uint32_t prim, slot;
    float u, v;
} rs_best;

/* :45-47 rs_limit */
static inline float rs_limit(const rs_best* b, uint32_t best_i, uint32_t i) {
    return (b->hit && i <= best_i) ? bf(fb(b->t) + 1u) : b->t;
}

/* :50-141 rs_instance; 1 when an any-hit ray found a blocker. */
static int rs_instance(ctx* x, uint32_t i, v3 origin, v3 dir, v3 world_inv, float tmin, int any_hit,
                       rs_best* best, uint32_t* best_i, uint32_t* stack) {
    uint64_t r = RS_HEADER + (uint64_t)i * RS_REC;
    x->c.inst_tests++;
    if (any_hit && W(x, r + 22) != 0u) return 0;
    float bound = rs_limit(best, *best_i, i);
    if (slab_near(origin, world_inv, GV(x, r + 12), GV(x, r + 15), tmin, bound) > bound) return 0;
    x->c.inst_entries++;
    v3 r0 = GV(x, r), r1 = GV(x, r + 4), r2 = GV(x, r + 8);
    /* dot(row, vec4(origin, 1)): x, y, z, then w * 1 */
    v3 o = mk(dot3(r0, origin) + F(x, r + 3) * 1.0f, dot3(r1, origin) + F(x, r + 7) * 1.0f,
              dot3(r2, origin) + F(x, r + 11) * 1.0f);
    v3 d = mk(dot3(r0, dir), dot3(r1, dir), dot3(r2, dir));
    v3 inv = safe_inverse(d);
    uint64_t nodes = W(x, r + 18), tris = W(x, r + 19);
    uint32_t slot = W(x, r + 21);
    stack[0] = W(x, r + 20);
    uint32_t sp = 1;
    while (sp > 0 && !x->bad) {
        sp -= 1;
        uint32_t child = stack[sp];
        for (;;) {
            if (child & LEAF) {
                x->c.leaves++;
                uint32_t start = (child & ~LEAF) >> 4, cnt = child & 15u;
                for (uint32_t k = 0; k < cnt && !x->bad; ++k) {
                    uint64_t t = tris + (uint64_t)(start + k) * RS_TRI;
                    x->c.tri_tests++;
                    v3 a = GV(x, t), e1 = GV(x, t + 3), e2 = GV(x, t + 6);
                    uint32_t prim = W(x, t + 9);
                    v3 p = cross3(d, e2);
                    float det = dot3(e1, p);
                    if (fabsf(det) < 1e-30f) continue;
                    float inv_det = 1.0f / det;
                    v3 s = sub3(o, a);
                    float u = dot3(s, p) * inv_det;
                    if (u < -RS_EDGE || u > 1.0f + RS_EDGE) continue;
                    v3 q = cross3(s, e1);
                    float v = dot3(d, q) * inv_det;
                    if (v < -RS_EDGE || u + v > 1.0f + RS_EDGE) continue;
                    float tt = dot3(e2, q) * inv_det;
                    int tied = best->hit && tt == best->t && (i < *best_i || (i == *best_i && prim < best->prim));
                    if (tt > tmin && (tt < best->t || tied)) {
                        best->hit = 1; best->t = tt; best->prim = prim; best->slot = slot;
                        best->u = u; best->v = v;
                        *best_i = i;
                        if (any_hit) return 1;
                    }
                }
                break;
            }
            x->c.blas_nodes++;
            uint64_t nb = nodes + (uint64_t)child * 16u;
            v3 base = GV(x, nb), step = GV(x, nb + 3);
            uint32_t ids[4] = {W(x, nb + 6), W(x, nb + 7), W(x, nb + 8), W(x, nb + 9)};
            uint32_t lower[3] = {W(x, nb + 10), W(x, nb + 11), W(x, nb + 12)};
            uint32_t upper[3] = {W(x, nb + 13), W(x, nb + 14), W(x, nb + 15)};
            float lim = rs_limit(best, *best_i, i);
            float dist[4];
            for (int c = 0; c < 4; ++c) {
                if (ids[c] == EMPTY) { dist[c] = FMAXV; continue; }
                uint32_t sh = 8u * (uint32_t)c;
                /* :36-42 rs_child: f32(byte) * step + base */
                v3 lo = mk((float)((lower[0] >> sh) & 255u) * step.x + base.x,
                           (float)((lower[1] >> sh) & 255u) * step.y + base.y,
                           (float)((lower[2] >> sh) & 255u) * step.z + base.z);
                v3 hi = mk((float)((upper[0] >> sh) & 255u) * step.x + base.x,
                           (float)((upper[1] >> sh) & 255u) * step.y + base.y,
                           (float)((upper[2] >> sh) & 255u) * step.z + base.z);
                dist[c] = slab_near(o, inv, lo, hi, tmin, lim);
            }
#define SW(A, B) do { float td = dist[A]; dist[A] = dist[B]; dist[B] = td; \
                      uint32_t ti = ids[A]; ids[A] = ids[B]; ids[B] = ti; } while (0)
            if (dist[1] < dist[0]) SW(0, 1);
            if (dist[3] < dist[2]) SW(2, 3);
            if (dist[2] < dist[0]) SW(0, 2);
            if (dist[3] < dist[1]) SW(1, 3);
            if (dist[2] < dist[1]) SW(1, 2);
#undef SW
            if (dist[3] <= lim && sp < STACK) stack[sp++] = ids[3];
            if (dist[2] <= lim && sp < STACK) stack[sp++] = ids[2];
            if (dist[1] <= lim && sp < STACK) stack[sp++] = ids[1];
            if (dist[0] <= lim) { child = ids[0]; continue; }
            break;
        }
    }
    return 0;
}

static void rs_write(const rs_best* b, uint32_t best_i, orc_hit* h) {
    memset(h, 0, sizeof *h);
    h->t = b->t;
    h->inst = b->hit ? best_i : 0xffffffffu;
    h->tri[0] = b->hit ? b->prim : 0u;
    h->u = b->hit ? b->u : 0.0f;
    h->v = b->hit ? b->v : 0.0f;
    h->flags = b->hit ? 1u : 0u;
    h->slot = b->hit ? b->slot : 0u;
}

/* :163-203 rs_trace (RS_TOP_WIDE); mode 0 closest, else any hit. */
int orc_rs_trace(const uint32_t* g, size_t n, const orc_ray* ray, uint32_t mode, orc_hit* hit,
                 orc_counts* counts) {
    ctx x;
    memset(&x, 0, sizeof x);
    x.g = g; x.n = n;
    rs_best best;
    memset(&best, 0, sizeof best);
    best.t = ray->tmax;
    uint32_t best_i = 0xffffffffu;
    if (ray_invalid(ray)) {
        rs_write(&best, best_i, hit);
        hit->flags = 2u;
        if (counts) *counts = x.c;
        return 0;
    }
    int any_hit = mode != 0u;
    uint32_t stack[STACK];
    v3 origin = mk(ray->o[0], ray->o[1], ray->o[2]);
    v3 dir = mk(ray->d[0], ray->d[1], ray->d[2]);
    float tmin = ray->tmin;
    v3 world_inv = safe_inverse(dir);
    uint32_t instances = W(&x, 0), top = W(&x, 1), root = W(&x, 2);
    if (instances != 0u) {
        /* the CPU-padded union of every instance (:175-187) */
        uint64_t b = W(&x, 4);
        v3 lo = GV(&x, b), hi = GV(&x, b + 4);
        v3 a = mul3(sub3(lo, origin), world_inv), c = mul3(sub3(hi, origin), world_inv);
        float near = mx(mx(mx(mn(a.x, c.x), mn(a.y, c.y)), mn(a.z, c.z)), tmin);
        float far = mn(mn(mn(mx(a.x, c.x), mx(a.y, c.y)), mx(a.z, c.z)), ray->tmax);
        if (near <= far) {
            best.t = mn(ray->tmax, bf(fb(far) + 2u));
            if (top == 0u) {
                for (uint32_t i = 0; i < instances && !x.bad; ++i)
                    if (rs_instance(&x, i, origin, dir, world_inv, tmin, any_hit, &best, &best_i, stack)) break;
            } else {
                uint32_t nodes[STACK];
                nodes[0] = root;
                uint32_t sp = 1;
                while (sp > 0 && !x.bad) {
                    sp -= 1;
                    uint32_t item = nodes[sp];
                    if (item & LEAF) {
                        if (rs_instance(&x, item & ~LEAF, origin, dir, world_inv, tmin, any_hit, &best, &best_i,
                                        stack))
                            break;
                        continue;
                    }
                    x.c.top_nodes++;
                    uint64_t nb = top + (uint64_t)item * 32u;
                    float bound = best.hit ? bf(fb(best.t) + 1u) : best.t;
                    float dist[4];
                    uint32_t ids[4];
                    for (int l = 0; l < 4; ++l) {
                        float ax = (F(&x, nb + l) - origin.x) * world_inv.x;
                        float bx = (F(&x, nb + 12 + l) - origin.x) * world_inv.x;
                        float ay = (F(&x, nb + 4 + l) - origin.y) * world_inv.y;
                        float by = (F(&x, nb + 16 + l) - origin.y) * world_inv.y;
                        float az = (F(&x, nb + 8 + l) - origin.z) * world_inv.z;
                        float bz = (F(&x, nb + 20 + l) - origin.z) * world_inv.z;
                        float nr = mx(mx(mx(mn(ax, bx), mn(ay, by)), mn(az, bz)), tmin);
                        float fr = mn(mn(mn(mx(ax, bx), mx(ay, by)), mx(az, bz)), bound);
                        ids[l] = W(&x, nb + 24 + l);
                        dist[l] = (nr <= fr && ids[l] != EMPTY) ? nr : FMAXV;
                    }
#define SW(A, B) do { float td = dist[A]; dist[A] = dist[B]; dist[B] = td; \
                      uint32_t ti = ids[A]; ids[A] = ids[B]; ids[B] = ti; } while (0)
                    if (dist[1] < dist[0]) SW(0, 1);
                    if (dist[3] < dist[2]) SW(2, 3);
                    if (dist[2] < dist[0]) SW(0, 2);
                    if (dist[3] < dist[1]) SW(1, 3);
                    if (dist[2] < dist[1]) SW(1, 2);
#undef SW
                    for (int l = 3; l >= 0; --l)
                        if (dist[l] <= bound && sp < STACK) nodes[sp++] = ids[l];
                }
            }
        }
    }
    rs_write(&best, best_i, hit);
    if (x.bad) hit->flags |= 4u;
    if (counts) *counts = x.c;
    return x.bad ? -1 : 0;
}

/* ------------------------------------------------------------- batching */
typedef struct job {
    const uint32_t* g; size_t n; const orc_ray* rays; uint32_t count; uint32_t mode;
    orc_hit* hits; uint32_t* vis; orc_counts sum; int fails; int rs;
    uint32_t next; /* shared cursor */
    pthread_mutex_t lock;
} job;
typedef struct worker { job* j; orc_counts c; int fails; } worker;

static void* run_worker(void* arg) {
    worker* w = (worker*)arg;
    job* j = w->j;
    for (;;) {
        uint32_t start = __atomic_fetch_add(&j->next, 256u, __ATOMIC_RELAXED);
        if (start >= j->count) break;
        uint32_t end = start + 256u < j->count ? start + 256u : j->count;
        for (uint32_t i = start; i < end; ++i) {
            orc_hit h;
            orc_counts c;
            int rc = j->rs ? orc_rs_trace(j->g, j->n, &j->rays[i], j->mode, &h, &c)
                           : orc_gi_traverse(j->g, j->n, &j->rays[i], j->mode, &h, &c);
            if (rc) w->fails++;
            if (j->hits) j->hits[i] = h;
            if (j->vis && !(h.flags & 1u)) __atomic_fetch_or(&j->vis[i / 32u], 1u << (i % 32u), __ATOMIC_RELAXED);
            w->c.words += c.words; w->c.top_nodes += c.top_nodes; w->c.inst_tests += c.inst_tests;
            w->c.inst_entries += c.inst_entries; w->c.blas_nodes += c.blas_nodes; w->c.leaves += c.leaves;
            w->c.tri_steps += c.tri_steps; w->c.tri_tests += c.tri_tests;
        }
    }
    return NULL;
}

static int traverse_batch(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                          orc_hit* hits, uint32_t* vis_bits, orc_counts* counts, uint32_t threads, int rs) {
    if (threads == 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 0 ? (uint32_t)c : 1u;
    }
    if (threads > 256) threads = 256;
    job j;
    memset(&j, 0, sizeof j);
    j.g = g; j.n = n; j.rays = rays; j.count = count; j.mode = mode; j.hits = hits; j.vis = vis_bits; j.rs = rs;
    if (vis_bits) memset(vis_bits, 0, ((size_t)count + 31u) / 32u * 4u);
    worker ws[256];
    pthread_t th[256];
    for (uint32_t t = 0; t < threads; ++t) {
        memset(&ws[t], 0, sizeof ws[t]);
        ws[t].j = &j;
        pthread_create(&th[t], NULL, run_worker, &ws[t]);
    }
    orc_counts sum;
    memset(&sum, 0, sizeof sum);
    int fails = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        fails += ws[t].fails;
        sum.words += ws[t].c.words; sum.top_nodes += ws[t].c.top_nodes; sum.inst_tests += ws[t].c.inst_tests;
        sum.inst_entries += ws[t].c.inst_entries; sum.blas_nodes += ws[t].c.blas_nodes;
        sum.leaves += ws[t].c.leaves; sum.tri_steps += ws[t].c.tri_steps; sum.tri_tests += ws[t].c.tri_tests;
    }
    if (counts) *counts = sum;
    return fails;
}

int orc_gi_traverse_batch(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                          orc_hit* hits, uint32_t* vis_bits, orc_counts* counts, uint32_t threads) {
    return traverse_batch(g, n, rays, count, mode, hits, vis_bits, counts, threads, 0);
}

int orc_rs_trace_batch(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                       orc_hit* hits, uint32_t* vis_bits, orc_counts* counts, uint32_t threads) {
    return traverse_batch(g, n, rays, count, mode, hits, vis_bits, counts, threads, 1);
}

/* -------------------------------------------------- flat f32 loop (no tree) */
int orc_flat_f32(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                 orc_hit* hits) {
    int fails = 0;
    for (uint32_t q = 0; q < count; ++q) {
        ctx x;
        memset(&x, 0, sizeof x);
        x.g = g; x.n = n;
        x.wt = (mode & ORC_WATERTIGHT) != 0u;
        const orc_ray* ray = &rays[q];
        best_t best;
        init_best(&best, ray->tmax);
        if (ray_invalid(ray)) { write_hit(&best, &hits[q]); hits[q].flags = 2u; continue; }
        int any_hit = (mode & 0xffu) != 0u;
        v3 origin = mk(ray->o[0], ray->o[1], ray->o[2]);
        v3 dir = mk(ray->d[0], ray->d[1], ray->d[2]);
        uint32_t instances = W(&x, 0);
        int done = 0;
        for (uint32_t i = 0; i < instances && !done && !x.bad; ++i) {
            uint64_t r = 16u + (uint64_t)i * IW;
            if (any_hit && (W(&x, r + 27) & 1u)) continue;
            v3 r0 = GV(&x, r), r1 = GV(&x, r + 4), r2 = GV(&x, r + 8);
            v3 o = mk(dot3(r0, origin) + F(&x, r + 3) * 1.0f, dot3(r1, origin) + F(&x, r + 7) * 1.0f,
                      dot3(r2, origin) + F(&x, r + 11) * 1.0f);
            v3 d = mk(dot3(r0, dir), dot3(r1, dir), dot3(r2, dir));
            uint32_t leaves = W(&x, r + 19), positions = W(&x, r + 20), indices = W(&x, r + 21);
            uint32_t prim_base = W(&x, r + 28), nodes = W(&x, r + 18);
            /* The mesh's index count: leaves run contiguously up to the last leaf's end. */
            uint32_t leaf_count = (positions - leaves) / 2u;
            (void)nodes;
            uint32_t total = leaf_count ? W(&x, leaves + 2u * (leaf_count - 1u)) +
                                              W(&x, leaves + 2u * (leaf_count - 1u) + 1u)
                                        : 0u;
            for (uint32_t k = 0; k < total && !x.bad; k += 3) {
                uint32_t ia = W(&x, indices + k), ib = W(&x, indices + k + 1), ic = W(&x, indices + k + 2);
                v3 a = GV(&x, positions + (uint64_t)ia * 3u), b = GV(&x, positions + (uint64_t)ib * 3u),
                   c = GV(&x, positions + (uint64_t)ic * 3u);
                v3 h = x.wt ? wt_tri_test(o, d, a, b, c, ray->tmin, limit_of(&best, i))
                            : tri_test(o, d, a, b, c, ray->tmin, limit_of(&best, i));
                if (h.x >= 0.0f) {
                    best.hit = 1; best.t = h.x; best.u = h.y; best.v = h.z; best.inst = i;
                    best.n = hit_normal(r0, r1, r2, a, b, c);
                    best.tri[0] = ia; best.tri[1] = ib; best.tri[2] = ic;
                    best.slot = k / 3u;
                    if (!any_hit && prim_base != 0u) {
                        best.tri[0] = W(&x, prim_base + k / 3u);
                        best.tri[1] = 0; best.tri[2] = 0;
                    }
                    if (any_hit) { done = 1; break; }
                }
            }
        }
        write_hit(&best, &hits[q]);
        if (x.bad) { hits[q].flags |= 4u; fails++; }
    }
    return fails;
}

/* ------------------------------------------------------------ f64 checks */
typedef struct d3 { double x, y, z; } d3;
static inline d3 dk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 dsub(d3 a, d3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b) {
    return dk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline d3 dv(const uint32_t* g, uint64_t i) { return dk(bf(g[i]), bf(g[i + 1]), bf(g[i + 2])); }

/* Object-space f64 ray of instance record r (world-to-object rows, widened). */
static void obj_ray64(const uint32_t* g, uint64_t r, const orc_ray* ray, d3* o, d3* d) {
    double ro[3], rd[3];
    for (int k = 0; k < 3; ++k) {
        uint64_t w = r + 4u * (uint64_t)k;
        ro[k] = (double)bf(g[w]) * ray->o[0] + (double)bf(g[w + 1]) * ray->o[1] +
                (double)bf(g[w + 2]) * ray->o[2] + (double)bf(g[w + 3]);
        rd[k] = (double)bf(g[w]) * ray->d[0] + (double)bf(g[w + 1]) * ray->d[1] +
                (double)bf(g[w + 2]) * ray->d[2];
    }
    *o = dk(ro[0], ro[1], ro[2]);
    *d = dk(rd[0], rd[1], rd[2]);
}

/* reference.rs:121-158 arithmetic; returns 1 with (t,u,v) on a hit. */
static int tri64(d3 o, d3 d, d3 a, d3 b, d3 c, double* t, double* u, double* v) {
    d3 ab = dsub(b, a), ac = dsub(c, a);
    d3 p = dcross(d, ac);
    double det = ddot(ab, p);
    if (det == 0.0) return 0;
    d3 s = dsub(o, a);
    double uu = ddot(s, p) / det;
    if (!(uu >= 0.0 && uu <= 1.0)) return 0;
    d3 q = dcross(s, ab);
    double vv = ddot(d, q) / det;
    if (vv < 0.0 || uu + vv > 1.0) return 0;
    *t = ddot(ac, q) / det;
    *u = uu;
    *v = vv;
    return 1;
}

static int mesh_index_count(const uint32_t* g, size_t n, uint64_t r, uint32_t* total) {
    uint32_t leaves = g[r + 19], positions = g[r + 20];
    uint32_t leaf_count = (positions - leaves) / 2u;
    if (leaf_count == 0) { *total = 0; return 0; }
    uint64_t l = leaves + 2u * (uint64_t)(leaf_count - 1u);
    if (l + 1 >= n) return -1;
    *total = g[l] + g[l + 1];
    return 0;
}

typedef struct bjob {
    const uint32_t* g; size_t n; const orc_ray* rays; uint32_t count; uint32_t mode;
    orc_hit64* hits; uint32_t next;
} bjob;

static void brute_one(const bjob* j, uint32_t q) {
    const uint32_t* g = j->g;
    const orc_ray* ray = &j->rays[q];
    orc_hit64 best;
    memset(&best, 0, sizeof best);
    best.t = -1.0;
    best.inst = 0xffffffffu;
    if (ray_invalid(ray)) { j->hits[q] = best; return; }
    int any_hit = j->mode != 0u;
    uint32_t instances = g[0];
    for (uint32_t i = 0; i < instances; ++i) {
        uint64_t r = 16u + (uint64_t)i * IW;
        if (r + IW > j->n) break;
        if (any_hit && (g[r + 27] & 1u)) continue;
        d3 o, d;
        obj_ray64(g, r, ray, &o, &d);
        uint32_t positions = g[r + 20], indices = g[r + 21], total;
        if (mesh_index_count(g, j->n, r, &total)) continue;
        for (uint32_t k = 0; k + 2 < total; k += 3) {
            double t, u, v;
            d3 a = dv(g, positions + 3u * (uint64_t)g[indices + k]);
            d3 b = dv(g, positions + 3u * (uint64_t)g[indices + k + 1]);
            d3 c = dv(g, positions + 3u * (uint64_t)g[indices + k + 2]);
            if (!tri64(o, d, a, b, c, &t, &u, &v)) continue;
            if (!(t > (double)ray->tmin && t < (double)ray->tmax)) continue;
            if (!best.hit || t < best.t) {
                best.hit = 1; best.t = t; best.u = u; best.v = v; best.inst = i; best.slot = k / 3u;
            }
            if (any_hit) { j->hits[q] = best; return; }
        }
    }
    j->hits[q] = best;
}

static void* brute_worker(void* arg) {
    bjob* j = (bjob*)arg;
    for (;;) {
        uint32_t q = __atomic_fetch_add(&j->next, 1u, __ATOMIC_RELAXED);
        if (q >= j->count) break;
        brute_one(j, q);
    }
    return NULL;
}

int orc_brute_f64(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                  orc_hit64* hits, uint32_t threads) {
    if (threads == 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 0 ? (uint32_t)c : 1u;
    }
    if (threads > 256) threads = 256;
    bjob j = {g, n, rays, count, mode, hits, 0};
    pthread_t th[256];
    for (uint32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, brute_worker, &j);
    for (uint32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* f64 brute force over rl_geometry words: every triangle (a, e1, e2 as
   stored) of every instance, the same edge slack as rs_trace, lexicographic
   (t, instance, primitive) tie rule; hit64.slot holds the primitive. A mesh's
   triangles run from its triangle base to the next base word packed after it. */
typedef struct rjob {
    const uint32_t* g; size_t n; const orc_ray* rays; uint32_t count; uint32_t mode; double slack;
    orc_hit64* hits; uint32_t next;
} rjob;

static uint64_t rs_tri_end(const uint32_t* g, size_t n, uint64_t tris) {
    uint64_t end = n;
    for (uint32_t i = 0; i < g[0]; ++i) {
        uint64_t r = RS_HEADER + (uint64_t)i * RS_REC;
        if (r + RS_REC > n) break;
        if (g[r + 18] > tris && g[r + 18] < end) end = g[r + 18];
        if (g[r + 19] > tris && g[r + 19] < end) end = g[r + 19];
    }
    return end;
}

static void rs_brute_one(const rjob* j, uint32_t q) {
    const uint32_t* g = j->g;
    const orc_ray* ray = &j->rays[q];
    orc_hit64 best;
    memset(&best, 0, sizeof best);
    best.t = -1.0;
    best.inst = 0xffffffffu;
    if (ray_invalid(ray)) { j->hits[q] = best; return; }
    int any_hit = j->mode != 0u;
    for (uint32_t i = 0; i < g[0]; ++i) {
        uint64_t r = RS_HEADER + (uint64_t)i * RS_REC;
        if (r + RS_REC > j->n) break;
        if (any_hit && g[r + 22] != 0u) continue;
        d3 o, d;
        obj_ray64(g, r, ray, &o, &d);
        uint64_t tris = g[r + 19], end = rs_tri_end(g, j->n, tris);
        for (uint64_t t = tris; t + RS_TRI <= end; t += RS_TRI) {
            d3 a = dv(g, t), e1 = dv(g, t + 3), e2 = dv(g, t + 6);
            uint32_t prim = g[t + 9];
            d3 p = dcross(d, e2);
            double det = ddot(e1, p);
            if (det == 0.0) continue;
            d3 s = dsub(o, a);
            double u = ddot(s, p) / det;
            if (u < -j->slack || u > 1.0 + j->slack) continue;
            d3 qv = dcross(s, e1);
            double v = ddot(d, qv) / det;
            if (v < -j->slack || u + v > 1.0 + j->slack) continue;
            double tt = ddot(e2, qv) / det;
            if (!(tt > (double)ray->tmin && tt < (double)ray->tmax)) continue;
            if (!best.hit || tt < best.t || (tt == best.t && (i < best.inst || (i == best.inst && prim < best.slot)))) {
                best.hit = 1; best.t = tt; best.u = u; best.v = v; best.inst = i; best.slot = prim;
            }
            if (any_hit) { j->hits[q] = best; return; }
        }
    }
    j->hits[q] = best;
}

static void* rs_brute_worker(void* arg) {
    rjob* j = (rjob*)arg;
    for (;;) {
        uint32_t q = __atomic_fetch_add(&j->next, 1u, __ATOMIC_RELAXED);
        if (q >= j->count) break;
        rs_brute_one(j, q);
    }
    return NULL;
}

int orc_rs_brute_f64(const uint32_t* g, size_t n, const orc_ray* rays, uint32_t count, uint32_t mode,
                     double slack, orc_hit64* hits, uint32_t threads) {
    if (threads == 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 0 ? (uint32_t)c : 1u;
    }
    if (threads > 256) threads = 256;
    rjob j = {g, n, rays, count, mode, slack, hits, 0};
    pthread_t th[256];
    for (uint32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, rs_brute_worker, &j);
    for (uint32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    return 0;
}

int orc_reintersect_f64(const uint32_t* g, size_t n, const orc_ray* rays, const orc_hit* hits,
                        uint32_t count, orc_hit64* out) {
    int fails = 0;
    for (uint32_t q = 0; q < count; ++q) {
        orc_hit64 h;
        memset(&h, 0, sizeof h);
        h.t = -1.0;
        h.inst = hits[q].inst;
        h.slot = hits[q].slot;
        if (hits[q].flags & 1u) {
            uint64_t r = 16u + (uint64_t)hits[q].inst * IW;
            if (r + IW > n) { fails++; out[q] = h; continue; }
            d3 o, d;
            obj_ray64(g, r, &rays[q], &o, &d);
            uint32_t positions = g[r + 20], indices = g[r + 21];
            uint64_t k = (uint64_t)indices + 3u * (uint64_t)hits[q].slot;
            if (k + 2 >= n) { fails++; out[q] = h; continue; }
            d3 a = dv(g, positions + 3u * (uint64_t)g[k]);
            d3 b = dv(g, positions + 3u * (uint64_t)g[k + 1]);
            d3 c = dv(g, positions + 3u * (uint64_t)g[k + 2]);
            /* The plane solve without the inside test: the reported triangle is given. */
            d3 ab = dsub(b, a), ac = dsub(c, a), p = dcross(d, ac);
            double det = ddot(ab, p);
            if (det != 0.0) {
                d3 s = dsub(o, a), qq = dcross(s, ab);
                h.u = ddot(s, p) / det;
                h.v = ddot(d, qq) / det;
                h.t = ddot(ac, qq) / det;
                h.hit = 1;
            }
        }
        out[q] = h;
    }
    return fails;
}

/* ------------------------------------------------------ Bvh::validate */
static inline float dec(float o, float s, uint32_t q, int fused) {
    return fused ? fmaf(s, (float)q, o) : o + s * (float)q;
}

int orc_validate_blas(const uint32_t* g, size_t n, uint32_t nodes_at, uint32_t node_count,
                      uint32_t leaves_at, uint32_t leaf_count, uint32_t positions_at,
                      uint32_t vertex_count, uint32_t indices_at, uint32_t index_count, uint32_t root,
                      uint32_t* depth_out) {
    if ((uint64_t)nodes_at + 16ull * node_count > n || (uint64_t)leaves_at + 2ull * leaf_count > n ||
        (uint64_t)positions_at + 3ull * vertex_count > n || (uint64_t)indices_at + index_count > n)
        return -1;
    if (node_count >= LEAF || leaf_count == 0) return -2;
    for (uint32_t v = 0; v < 3u * vertex_count; ++v)
        if (!isfinite(bf(g[positions_at + v]))) return -3;
    for (uint32_t k = 0; k < index_count; ++k)
        if (g[indices_at + k] >= vertex_count) return -4;
    uint32_t end = 0;
    for (uint32_t l = 0; l < leaf_count; ++l) {
        uint32_t off = g[leaves_at + 2 * l], cnt = g[leaves_at + 2 * l + 1];
        if (off != end || cnt == 0 || cnt % 3 != 0 || (uint64_t)off + cnt > index_count) return -5;
        end = off + cnt;
    }
    if (end != index_count) return -6;
    /* Per node: its f64 enclosure box, built children-first (post-order ids). */
    double* boxes = (double*)calloc((size_t)node_count * 6u + 6u, sizeof(double));
    uint32_t* depth = (uint32_t*)calloc((size_t)node_count + 1u, sizeof(uint32_t));
    if (!boxes || !depth) { free(boxes); free(depth); return -7; }
    int rc = 0;
    for (uint32_t i = 0; i < node_count && !rc; ++i) {
        const uint32_t* w = g + nodes_at + 16ull * i;
        float o[3] = {bf(w[0]), bf(w[1]), bf(w[2])}, s[3] = {bf(w[3]), bf(w[4]), bf(w[5])};
        int nonempty = 0;
        for (int a = 0; a < 3; ++a)
            if (!isfinite(o[a]) || !isfinite(s[a]) || !(s[a] > 0.0f)) rc = -8;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t dmax = 0;
        for (uint32_t slot = 0; slot < 4 && !rc; ++slot) {
            uint32_t child = w[6 + slot];
            uint32_t ql[3], qh[3];
            for (int a = 0; a < 3; ++a) {
                ql[a] = (w[10 + a] >> (8 * slot)) & 0xffu;
                qh[a] = (w[13 + a] >> (8 * slot)) & 0xffu;
            }
            if (child == EMPTY) {
                for (int a = 0; a < 3; ++a)
                    if (ql[a] || qh[a]) rc = -9;
                continue;
            }
            nonempty = 1;
            double clo[3], chi[3];
            if (child & LEAF) {
                uint32_t id = child & ~LEAF;
                if (id >= leaf_count) { rc = -10; break; }
                uint32_t off = g[leaves_at + 2 * id], cnt = g[leaves_at + 2 * id + 1];
                for (int a = 0; a < 3; ++a) { clo[a] = INFINITY; chi[a] = -INFINITY; }
                for (uint32_t k = off; k < off + cnt; ++k) {
                    uint32_t vi = g[indices_at + k];
                    for (int a = 0; a < 3; ++a) {
                        double p = bf(g[positions_at + 3 * vi + a]);
                        if (p < clo[a]) clo[a] = p;
                        if (p > chi[a]) chi[a] = p;
                    }
                }
            } else {
                if (child >= i) { rc = -11; break; }
                for (int a = 0; a < 3; ++a) { clo[a] = boxes[6 * child + a]; chi[a] = boxes[6 * child + 3 + a]; }
                if (depth[child] > dmax) dmax = depth[child];
            }
            for (int a = 0; a < 3; ++a) {
                double la = dec(o[a], s[a], ql[a], 0), ha = dec(o[a], s[a], qh[a], 0);
                double lf = dec(o[a], s[a], ql[a], 1), hf = dec(o[a], s[a], qh[a], 1);
                if (!isfinite(la) || !isfinite(ha) || !isfinite(lf) || !isfinite(hf) || la > clo[a] ||
                    ha < chi[a] || lf > clo[a] || hf < chi[a]) {
                    rc = -12;
                    break;
                }
                double l = la < lf ? la : lf, h = ha > hf ? ha : hf;
                if (l < lo[a]) lo[a] = l;
                if (h > hi[a]) hi[a] = h;
            }
        }
        if (!rc && !nonempty) rc = -13;
        for (int a = 0; a < 3; ++a) { boxes[6 * i + a] = lo[a]; boxes[6 * i + 3 + a] = hi[a]; }
        depth[i] = dmax + 1u;
    }
    /* Reachability: every node and leaf exactly once from the root. */
    if (!rc) {
        uint8_t* seen_n = (uint8_t*)calloc(node_count + 1u, 1);
        uint8_t* seen_l = (uint8_t*)calloc(leaf_count, 1);
        uint32_t* st = (uint32_t*)malloc(((size_t)node_count * 4u + 8u) * sizeof(uint32_t));
        size_t sp = 0;
        st[sp++] = root;
        size_t visits = 0;
        while (sp && !rc) {
            uint32_t c = st[--sp];
            if (++visits > (size_t)node_count + leaf_count) { rc = -14; break; }
            if (c == EMPTY) { rc = -15; break; }
            if (c & LEAF) {
                uint32_t id = c & ~LEAF;
                if (id >= leaf_count || seen_l[id]) { rc = -16; break; }
                seen_l[id] = 1;
            } else {
                if (c >= node_count || seen_n[c]) { rc = -17; break; }
                seen_n[c] = 1;
                for (int s = 0; s < 4; ++s) {
                    uint32_t ch = g[nodes_at + 16ull * c + 6 + s];
                    if (ch != EMPTY) st[sp++] = ch;
                }
            }
        }
        for (uint32_t i = 0; i < node_count && !rc; ++i)
            if (!seen_n[i]) rc = -18;
        for (uint32_t l = 0; l < leaf_count && !rc; ++l)
            if (!seen_l[l]) rc = -18;
        free(seen_n); free(seen_l); free(st);
    }
    if (depth_out) *depth_out = (root & LEAF) ? 0u : (root < node_count ? depth[root] : 0u);
    free(boxes);
    free(depth);
    return rc;
}

/* ------------------------------------------------------ ray generation */
/* shaders/primary_rays.wgsl:50-66, reversed Z, explicit fma; `inverse` is
   column-major (column c at inverse[4c..4c+3]). Direction = far - near. */
void orc_primary_ray(const float inverse[16], uint32_t px, uint32_t py, uint32_t w, uint32_t h,
                     orc_ray* out) {
    float nx = fmaf(((float)px + 0.5f) / (float)w, 2.0f, -1.0f);
    float ny = fmaf(((float)py + 0.5f) / (float)h, -2.0f, 1.0f);
    float xy[4], a[4], b[4];
    for (int k = 0; k < 4; ++k) {
        xy[k] = fmaf(inverse[k], nx, inverse[4 + k] * ny);
        a[k] = (xy[k] + inverse[8 + k]) + inverse[12 + k];
        b[k] = xy[k] + inverse[12 + k];
    }
    for (int k = 0; k < 3; ++k) {
        float near = a[k] / a[3], far = b[k] / b[3];
        out->o[k] = near;
        out->d[k] = far - near;
    }
    out->tmin = 0.0f;
    out->tmax = 1e20f;
}

/* reference_path_tracer.wgsl:149-152 (Jarzynski and Olano 2020). */
uint32_t orc_pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
void orc_pcg_batch(const uint32_t* in, uint32_t count, uint32_t* out) {
    for (uint32_t i = 0; i < count; ++i) out[i] = orc_pcg(in[i]);
}
/* reference_path_tracer.wgsl:154-158 (rand): the state advances by one hash and
   its top 24 bits become a float in [0, 1). */
static inline float rnd(uint32_t* s) {
    *s = orc_pcg(*s);
    return (float)(*s >> 8u) * (1.0f / 16777216.0f);
}
void orc_rand_sequence(uint32_t state, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; ++i) out[i] = rnd(&state);
}

/* reference_path_tracer.wgsl:160-167 (basis, Duff et al. 2017): columns c0, c1,
   c2 = n, stored out[3c + row]. */
void orc_basis(const float n[3], float out[9]) {
    float sg = n[2] >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sg + n[2]);
    float b = n[0] * n[1] * a;
    float c[9] = {1.0f + sg * n[0] * n[0] * a, sg * b, -sg * n[0],
                  b, sg + n[1] * n[1] * a, -n[1],
                  n[0], n[1], n[2]};
    memcpy(out, c, sizeof c);
}

/* reference_path_tracer.wgsl:466-468: the cosine-weighted direction of the
   random pair (u1, u2) about n: r = sqrt(u1), phi = 2 pi u2, basis(n) applied to
   (r cos phi, r sin phi, sqrt(max(0, 1 - r^2))), the matrix-vector product
   summed column by column. */
void orc_cosine_dir(const float n[3], float u1, float u2, float out[3]) {
    float c[9];
    orc_basis(n, c);
    float rr = sqrtf(u1);
    float phi = 6.2831855f * u2;
    float l[3] = {rr * cosf(phi), rr * sinf(phi), sqrtf(mx(0.0f, 1.0f - rr * rr))};
    for (int k = 0; k < 3; ++k) out[k] = c[k] * l[0] + c[3 + k] * l[1] + c[6 + k] * l[2];
}

/* reference_path_tracer.wgsl:237-240 (offset): p pushed along ng by
   1e-4 + 4e-7 max(|p.x|, |p.y|, |p.z|). */
void orc_offset(const float p[3], const float ng[3], float out[3]) {
    float m = mx(mx(fabsf(p[0]), fabsf(p[1])), fabsf(p[2]));
    float k = 1e-4f + 4e-7f * m;
    for (int i = 0; i < 3; ++i) out[i] = p[i] + ng[i] * k;
}

/* The rays leaving one primary hit. The hit point is origin + direction * t;
   the geometric normal is turned against the incoming ray
   (reference_path_tracer.wgsl:224); the shadow ray toward the sun starts at the
   offset point and is valid when the sun lies on the normal's side (:249-256);
   each GI bounce draws two rand() values for its cosine direction (:466-470),
   valid when it leaves on the normal's side. The random state of pixel p
   starts at p ^ (seed * 0x9E3779B9). */
uint32_t orc_secondary_rays(const orc_ray* primary, const orc_hit* hit, const float sun[3], uint32_t pixel,
                            uint32_t seed, uint32_t n_gi, orc_ray* out) {
    if (!(hit->flags & 1u)) return 0u;
    float t = hit->t;
    v3 o = mk(primary->o[0], primary->o[1], primary->o[2]);
    v3 d = mk(primary->d[0], primary->d[1], primary->d[2]);
    float p[3] = {o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    v3 ngv = mk(hit->n[0], hit->n[1], hit->n[2]);
    if (dot3(ngv, d) > 0.0f) ngv = mk(-ngv.x, -ngv.y, -ngv.z);
    float ng[3] = {ngv.x, ngv.y, ngv.z}, start[3];
    orc_offset(p, ng, start);
    uint32_t valid = 0;
    v3 s = mk(sun[0], sun[1], sun[2]);
    orc_ray* r = &out[0];
    memcpy(r->o, start, sizeof start);
    r->d[0] = s.x; r->d[1] = s.y; r->d[2] = s.z;
    r->tmin = 0.0f; r->tmax = 1e20f;
    if (dot3(ngv, s) > 0.0f) valid |= 1u;
    uint32_t state = pixel ^ (seed * 0x9E3779B9u);
    for (uint32_t gi = 0; gi < n_gi; ++gi) {
        float u1 = rnd(&state);
        float u2 = rnd(&state);
        float wi[3];
        orc_cosine_dir(ng, u1, u2, wi);
        orc_ray* q = &out[1 + gi];
        memcpy(q->o, start, sizeof start);
        memcpy(q->d, wi, sizeof wi);
        q->tmin = 0.0f; q->tmax = 1e20f;
        if (dot3(ngv, mk(wi[0], wi[1], wi[2])) > 0.0f) valid |= 1u << (1u + gi);
    }
    return valid;
}

/* ------------------------------------------------------ hit records */
/* The 16-byte hit record vec4<u32>(placement + 1, primitive, bits(u), bits(v)),
   zero on a miss (shaders/primary_rays.wgsl:19; RsHit.record,
   shaders/reflections_software_bvh4.wgsl:42, 90), with gi_hit_identity,
   gi_hit_primitive and gi_hit_placement as shaders/gi_rays_software.wgsl:297-316
   write them: the identity table is scanned in full and its last pair whose
   first position is <= tri.x wins. A hit on a record past header word 4 (none
   exists) gives zeros. Reflection hits carry slot and primitive themselves. */
void orc_hit_records(const uint32_t* g, size_t n, int reflection, const orc_hit* hits, uint32_t count,
                     uint32_t* records, uint32_t* identities) {
    for (uint32_t k = 0; k < count; ++k) {
        const orc_hit* h = &hits[k];
        uint32_t rec[4] = {0, 0, 0, 0}, identity = 0;
        if (h->flags & 1u) {
            if (reflection) {
                rec[0] = h->slot + 1u;
                rec[1] = h->tri[0];
                rec[2] = fb(h->u);
                rec[3] = fb(h->v);
            } else if (n >= 16 && h->inst < g[4] && 16 + (uint64_t)h->inst * IW + IW <= n) {
                uint64_t r = 16 + (uint64_t)h->inst * IW;
                int opaque = h->inst < g[0];
                rec[0] = g[r + 29] + 1u;                                          /* gi_hit_placement */
                rec[1] = (!opaque || g[r + 28] == 0u) ? 0xffffffffu : h->tri[0];  /* gi_hit_primitive */
                rec[2] = fb(h->u);
                rec[3] = fb(h->v);
                if (opaque) {                                                      /* gi_hit_identity */
                    identity = g[r + 24];
                    if (identity == 0u) {
                        uint64_t map = g[r + 25];
                        uint32_t pairs = g[r + 26];
                        for (uint32_t i = 0; i < pairs; ++i)
                            if (map + 2ull * i + 1 < n && g[map + 2ull * i] <= h->tri[0]) identity = g[map + 2ull * i + 1];
                    }
                }
            }
        }
        if (records) memcpy(&records[4ull * k], rec, sizeof rec);
        if (identities) identities[k] = identity;
    }
}

/* ------------------------------------------------------ GI cache rays */
/* Table words: shaders/gi_cells.wgsl:18-185 (header 64 words; 24-word hot
   records; cold records of max(24, header[23]) words), as
   tests/ara_backend/gi_cache_reference.rs:50-118 decodes them. */
typedef struct {
    int ok;
    float n[3], q[3], edge;
    uint32_t key_x, key_y;
    uint64_t bits;
} gi_cell_t;

static float gi_bits_f(uint32_t w) {
    float f;
    memcpy(&f, &w, 4);
    return f;
}

/* gi_unpack_normal (shaders/gi_cells.wgsl:481-487): unpack2x16snorm, the
   octahedral fold, then v / |v|. */
static void gi_unpack_normal(uint32_t w, float n[3]) {
    float qx = (float)(int16_t)(uint16_t)(w & 0xffffu) / 32767.0f;
    float qy = (float)(int16_t)(uint16_t)(w >> 16) / 32767.0f;
    qx = fminf(fmaxf(qx, -1.0f), 1.0f);
    qy = fminf(fmaxf(qy, -1.0f), 1.0f);
    float x = qx, y = qy, z = 1.0f - fabsf(qx) - fabsf(qy);
    float t = fmaxf(-z, 0.0f);
    x = x + (x >= 0.0f ? -t : t);
    y = y + (y >= 0.0f ? -t : t);
    float len = sqrtf(x * x + y * y + z * z);
    n[0] = x / len;
    n[1] = y / len;
    n[2] = z / len;
}

static gi_cell_t gi_load_cell(const uint32_t* t, size_t words, uint32_t slot) {
    gi_cell_t c;
    memset(&c, 0, sizeof c);
    if (words <= 64) return c;
    uint64_t slots = (uint64_t)t[1] + 1, cold_words = t[23] > 24 ? t[23] : 24;
    if (slot >= slots) return c;
    uint64_t hot = 64 + (uint64_t)slot * 24, cold = 64 + slots * 24 + (uint64_t)slot * cold_words;
    if (cold + 24 > words || hot + 24 > words) return c;
    c.key_x = t[hot] & 0x7fffffffu;
    c.key_y = t[cold];
    c.edge = gi_bits_f(t[5]) * exp2f((float)(t[cold + 3] & 0xffu));  /* gi_edge, gi_cells.wgsl:240 */
    gi_unpack_normal(t[cold + 7], c.n);
    for (int i = 0; i < 3; ++i) c.q[i] = gi_bits_f(t[cold + 8 + i]);
    c.bits = (uint64_t)t[cold + 12] | ((uint64_t)t[cold + 13] << 32);
    c.ok = 1;
    return c;
}

/* gi_origin_offset (shaders/gi_cache.wgsl:1064-1068). */
static float gi_origin_offset(const float q[3], float edge) {
    float m = fmaxf(fmaxf(fabsf(q[0]), fabsf(q[1])), fabsf(q[2]));
    float mm = fmaxf(m, 1e-30f);
    uint32_t b;
    memcpy(&b, &mm, 4);
    b += 1;
    float next;
    memcpy(&next, &b, 4);
    float ulp = next - mm;
    return fmaxf(fmaxf(1e-5f, 32.0f * ulp) + 1e-4f * edge, 1e-4f + 4e-7f * m);
}

/* gi_direction (shaders/gi_cache.wgsl:1048-1054) through gi_basis
   (shaders/gi_cells.wgsl:463-469); frame * v summed column by column. */
static void gi_direction(const gi_cell_t* c, int rotate, uint32_t k, float d[3]) {
    float z = 1.0f - ((float)k + 0.5f) / 64.0f;
    float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    float rotation = rotate ? (float)(c->key_y >> 8) * (float)(6.283185307179586 / 16777216.0) : 0.0f;
    float phi = (float)k * 2.399963229728653f + rotation;
    float lx = r * cosf(phi), ly = r * sinf(phi), lz = z;
    const float* n = c->n;
    float s = n[2] >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n[2]);
    float b = n[0] * n[1] * a;
    float c0[3] = {1.0f + s * n[0] * n[0] * a, s * b, -s * n[0]};
    float c1[3] = {b, s + n[1] * n[1] * a, -n[1]};
    for (int i = 0; i < 3; ++i) d[i] = c0[i] * lx + c1[i] * ly + n[i] * lz;
}

static orc_ray gi_cell_ray(const gi_cell_t* c, int rotate, uint32_t k) {
    orc_ray r;
    float off = gi_origin_offset(c->q, c->edge), d[3];
    gi_direction(c, rotate, k, d);
    for (int i = 0; i < 3; ++i) {
        r.o[i] = c->q[i] + c->n[i] * off;
        r.d[i] = d[i];
    }
    r.tmin = 0.0f;
    r.tmax = 1e20f;
    return r;
}

void orc_gi_item_rays(const uint32_t* table, size_t words, const uint32_t* items, uint32_t count, orc_ray* rays) {
    int rotate = words > 2 && (table[2] & 4u) != 0;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t k = items[i] & 127u;
        gi_cell_t c = gi_load_cell(table, words, items[i] >> 7);
        if (c.ok && k < 64) {
            rays[i] = gi_cell_ray(&c, rotate, k);
        } else {
            orc_ray r;
            for (int a = 0; a < 3; ++a) {
                r.o[a] = NAN;
                r.d[a] = 0.0f;
            }
            r.tmin = 0.0f;
            r.tmax = 1e20f;
            rays[i] = r;
        }
    }
}

uint32_t orc_gi_gather_rays(const uint32_t* table, size_t words, const uint32_t* slots, uint32_t n, orc_ray* rays,
                            uint32_t* items, uint32_t max_rays) {
    int rotate = words > 2 && (table[2] & 4u) != 0;
    uint32_t stride = words > 21 && table[21] > 1 ? table[21] : 1, at = 0;
    for (uint32_t i = 0; i < n; ++i) {
        gi_cell_t c = gi_load_cell(table, words, slots[i]);
        if (!c.ok) continue;
        for (uint32_t k = c.key_x & (stride - 1); k < 64; k += stride) {
            if ((c.bits >> k) & 1u) continue;  /* the sky is known visible there */
            if (at < max_rays) {
                rays[at] = gi_cell_ray(&c, rotate, k);
                items[at] = (slots[i] << 7) | k;
            }
            ++at;
        }
    }
    return at;
}

/* ------------------------------------------------------ near-field masks */
uint32_t orc_near_bin(const float d[3], uint32_t bins) {
    float s = fabsf(d[0]) + fabsf(d[1]) + fabsf(d[2]);
    float u = d[0] / s, v = d[1] / s;
    if (!(d[2] >= 0.0f)) {  /* near_field.rs:160-169 (octahedral) */
        float fu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        float fv = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = fu;
        v = fv;
    }
    float n = (float)bins;
    float cu = floorf((u * 0.5f + 0.5f) * n), cv = floorf((v * 0.5f + 0.5f) * n);
    cu = cu < 0.0f ? 0.0f : (cu > n - 1.0f ? n - 1.0f : cu);
    cv = cv < 0.0f ? 0.0f : (cv > n - 1.0f ? n - 1.0f : cv);
    return (uint32_t)cv * bins + (uint32_t)cu;
}

void orc_near_bin_direction(uint32_t bin, uint32_t bins, float d[3]) {
    float n = (float)bins;
    float u = ((float)(bin % bins) + 0.5f) / n * 2.0f - 1.0f;
    float v = ((float)(bin / bins) + 0.5f) / n * 2.0f - 1.0f;
    float z = 1.0f - fabsf(u) - fabsf(v);
    float t = fmaxf(-z, 0.0f);
    float x = u + (u >= 0.0f ? -t : t);
    float y = v + (v >= 0.0f ? -t : t);
    float l = sqrtf(x * x + y * y + z * z);
    d[0] = x / l;
    d[1] = y / l;
    d[2] = z / l;
}

int orc_near_blocked(const uint32_t* g, size_t n, uint32_t bins, uint32_t inst, const uint32_t tri[3],
                     const float bary[2], int back, const float d[3]) {
    if (!bins || n < 16 || inst >= g[0] || 16 + (size_t)inst * 32 + 24 > n) return 0;
    uint32_t base = g[16 + (size_t)inst * 32 + 23];
    if (!base) return 0;
    uint32_t words = bins * bins / 32u, b = orc_near_bin(d, bins);
    uint64_t at = (uint64_t)base + (back ? words : 0u) + b / 32u, stride = 2ull * words;
    uint32_t bit = 1u << (b % 32u);
    float w[3] = {1.0f - bary[0] - bary[1], bary[0], bary[1]};
    float s = 0.0f;
    for (int v = 0; v < 3; ++v) {
        uint64_t k = at + (uint64_t)tri[v] * stride;
        if (k < n && (g[k] & bit)) s += w[v];
    }
    return s >= 0.999f;
}
