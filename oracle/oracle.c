/*
 * oracle.c -- CPU oracle for the ray-query hot path.  TEST INFRASTRUCTURE
 * ONLY: it is loaded by tests/, by __graft_entry__.smoke() and by the
 * CPU-baseline leg of bench.py, as the checker and the timed CPU baseline.
 * The library under liar_amd/ never calls it.
 *
 * It restates, in plain C, LiAR's ray-query semantics (paths relative to the
 * reference repository root):
 *   - BoundedRay::inRange: t > near*(1+eps) && t < far*(1-eps), eps = 1e-6f
 *     (src/liar/kernel/bounded_ray.h:69-77, src/liar/kernel/kernel_common.cpp:29-33)
 *   - TriangleMesh::doIntersect: nearest triangle t > near, then inRange
 *     (src/liar/scenery/triangle_mesh.cpp:291-317); doIsIntersecting: any
 *     triangle with near < t < far (triangle_mesh.cpp:321-330)
 *   - Sphere::doIntersect / doIsIntersecting: nearest root t > near, then
 *     inRange (src/liar/scenery/sphere.cpp:104-127)
 *   - List::doIntersect: minimum t over children; doIsIntersecting: any child
 *     (src/liar/scenery/list.cpp:106-150).  Equal t: smaller primitive id
 *     (LiAR keeps the first found, object_trees.h:243-244, which depends on
 *     traversal order; the id rule makes the answer order-independent).
 *
 *   - Transformation::doIntersect / doIsIntersecting: the ray in local space
 *     (origin W o, direction W d / |W d|, bounds times |W d|), world t =
 *     t_local / |W d| (src/liar/scenery/transformation.cpp:117-131); the
 *     instances combine as a List.  FP32, in the library's operation order.
 *
 * Lass (lass 1.11.0-250+g078d9309, conanfile.py:39), which holds LiAR's
 * triangle/sphere/box arithmetic and its QBVH, is not vendored in the
 * reference, so the FP32 expressions below are this project's statement of
 * Moller-Trumbore and of the stable sphere quadratic; they are pinned by the
 * FP64 versions (or_*_f64), by brute force on small inputs and by closed
 * forms in tests/test_oracle.py.  Lass's exact op order and tie winners are
 * parity unpinned (no reference vector exists for them).
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off; explicit fmaf order).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define OR_MISS 0xFFFFFFFFu
#define OR_EPS 1e-6f
#define OR_EMPTY 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* primitive tests (FP32 and FP64)                                          */
/* ------------------------------------------------------------------------ */

/* Moller-Trumbore, two-sided, FP32 with a fixed fmaf order.  e1 = v1 - v0,
 * e2 = v2 - v0.  Returns 1 with (t,u,v) when the line crosses the triangle. */
int or_tri_f32(const float o[3], const float d[3], const float v0[3], const float e1[3],
               const float e2[3], float* t, float* u, float* v) {
    float p0 = fmaf(d[1], e2[2], -(d[2] * e2[1]));
    float p1 = fmaf(d[2], e2[0], -(d[0] * e2[2]));
    float p2 = fmaf(d[0], e2[1], -(d[1] * e2[0]));
    float det = fmaf(e1[0], p0, fmaf(e1[1], p1, e1[2] * p2));
    if (det == 0.0f) return 0;
    float inv = 1.0f / det;
    float s0 = o[0] - v0[0], s1 = o[1] - v0[1], s2 = o[2] - v0[2];
    float uu = fmaf(s0, p0, fmaf(s1, p1, s2 * p2)) * inv;
    if (!(uu >= 0.0f && uu <= 1.0f)) return 0;
    float q0 = fmaf(s1, e1[2], -(s2 * e1[1]));
    float q1 = fmaf(s2, e1[0], -(s0 * e1[2]));
    float q2 = fmaf(s0, e1[1], -(s1 * e1[0]));
    float vv = fmaf(d[0], q0, fmaf(d[1], q1, d[2] * q2)) * inv;
    if (!(vv >= 0.0f && uu + vv <= 1.0f)) return 0;
    *t = fmaf(e2[0], q0, fmaf(e2[1], q1, e2[2] * q2)) * inv;
    *u = uu;
    *v = vv;
    return 1;
}

/* Same test in FP64 on the FP32 inputs (error reference). */
int or_tri_f64(const double o[3], const double d[3], const double v0[3], const double e1[3],
               const double e2[3], double* t, double* u, double* v) {
    double p0 = d[1] * e2[2] - d[2] * e2[1];
    double p1 = d[2] * e2[0] - d[0] * e2[2];
    double p2 = d[0] * e2[1] - d[1] * e2[0];
    double det = e1[0] * p0 + e1[1] * p1 + e1[2] * p2;
    if (det == 0.0) return 0;
    double s0 = o[0] - v0[0], s1 = o[1] - v0[1], s2 = o[2] - v0[2];
    double uu = (s0 * p0 + s1 * p1 + s2 * p2) / det;
    if (!(uu >= 0.0 && uu <= 1.0)) return 0;
    double q0 = s1 * e1[2] - s2 * e1[1];
    double q1 = s2 * e1[0] - s0 * e1[2];
    double q2 = s0 * e1[1] - s1 * e1[0];
    double vv = (d[0] * q0 + d[1] * q1 + d[2] * q2) / det;
    if (!(vv >= 0.0 && uu + vv <= 1.0)) return 0;
    *t = (e2[0] * q0 + e2[1] * q1 + e2[2] * q2) / det;
    *u = uu;
    *v = vv;
    return 1;
}

/* Nearest root t > tmin of |o + t d - c|^2 = r^2 (|d| = 1), FP32: the
 * discriminant from the perpendicular offset l = s - (s.d) d, roots q and
 * c'/q with q = -b - sign(b) sqrt(disc). */
int or_sphere_f32(const float o[3], const float d[3], const float c[3], float r2, float tmin,
                  float* t) {
    float s0 = o[0] - c[0], s1 = o[1] - c[1], s2 = o[2] - c[2];
    float b = fmaf(s0, d[0], fmaf(s1, d[1], s2 * d[2]));
    float l0 = fmaf(-b, d[0], s0), l1 = fmaf(-b, d[1], s1), l2 = fmaf(-b, d[2], s2);
    float disc = r2 - fmaf(l0, l0, fmaf(l1, l1, l2 * l2));
    if (!(disc >= 0.0f)) return 0;
    float cc = fmaf(s0, s0, fmaf(s1, s1, s2 * s2)) - r2;
    float sq = sqrtf(disc);
    float q = -b - copysignf(sq, b);
    float t0, t1;
    if (q == 0.0f) {
        t0 = 0.0f;
        t1 = 0.0f;
    } else {
        float ta = cc / q, tb = q;
        t0 = fminf(ta, tb);
        t1 = fmaxf(ta, tb);
    }
    if (t0 > tmin) { *t = t0; return 1; }
    if (t1 > tmin) { *t = t1; return 1; }
    return 0;
}

/* FP64 textbook quadratic (a = |d|^2 kept general). */
int or_sphere_f64(const double o[3], const double d[3], const double c[3], double r, double tmin,
                  double* t) {
    double s0 = o[0] - c[0], s1 = o[1] - c[1], s2 = o[2] - c[2];
    double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double b = s0 * d[0] + s1 * d[1] + s2 * d[2];
    double cc = s0 * s0 + s1 * s1 + s2 * s2 - r * r;
    double disc = b * b - a * cc;
    if (disc < 0.0) return 0;
    double sq = sqrt(disc);
    double t0 = (-b - sq) / a, t1 = (-b + sq) / a;
    if (t0 > tmin) { *t = t0; return 1; }
    if (t1 > tmin) { *t = t1; return 1; }
    return 0;
}

static int in_range(float t, float tnear, float tfar) {
    return t > tnear * (1.0f + OR_EPS) && t < tfar * (1.0f - OR_EPS);
}

/* ------------------------------------------------------------------------ */
/* scene / ray views                                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    const float* verts;      /* 3*nv */
    const uint32_t* tris;    /* 3*nt */
    uint32_t nt;
    const float* spheres;    /* 4*ns (cx, cy, cz, r) */
    uint32_t ns;
} or_scene;

typedef struct {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tnear, *tfar;
} or_rays;

static void tri_edges(const or_scene* s, uint32_t k, float v0[3], float e1[3], float e2[3]) {
    const float* a = s->verts + 3 * (size_t)s->tris[3 * (size_t)k + 0];
    const float* b = s->verts + 3 * (size_t)s->tris[3 * (size_t)k + 1];
    const float* c = s->verts + 3 * (size_t)s->tris[3 * (size_t)k + 2];
    for (int i = 0; i < 3; ++i) {
        v0[i] = a[i];
        e1[i] = b[i] - a[i];
        e2[i] = c[i] - a[i];
    }
}

/* LiAR List{TriangleMesh, Sphere_0 .. Sphere_ns-1}::intersect for one ray. */
static void brute_closest_one(const or_scene* s, const float o[3], const float d[3], float tnear,
                              float tfar, uint32_t* prim, float* th, float* uh, float* vh) {
    /* TriangleMesh: nearest t > near over all triangles, then inRange. */
    float bt = INFINITY, bu = 0.0f, bv = 0.0f;
    uint32_t bid = OR_MISS;
    for (uint32_t k = 0; k < s->nt; ++k) {
        float v0[3], e1[3], e2[3], t, u, v;
        tri_edges(s, k, v0, e1, e2);
        if (or_tri_f32(o, d, v0, e1, e2, &t, &u, &v) && t > tnear) {
            if (t < bt || (t == bt && k < bid)) {
                bt = t; bu = u; bv = v; bid = k;
            }
        }
    }
    uint32_t rid = OR_MISS;
    float rt = INFINITY, ru = 0.0f, rv = 0.0f;
    if (bid != OR_MISS && in_range(bt, tnear, tfar)) {
        rid = bid; rt = bt; ru = bu; rv = bv;
    }
    /* Spheres: nearest root > near, then inRange; List keeps the minimum. */
    for (uint32_t j = 0; j < s->ns; ++j) {
        const float* sp = s->spheres + 4 * (size_t)j;
        float t;
        if (or_sphere_f32(o, d, sp, sp[3] * sp[3], tnear, &t) && in_range(t, tnear, tfar)) {
            uint32_t id = s->nt + j;
            if (t < rt || (t == rt && id < rid)) {
                rid = id; rt = t; ru = 0.0f; rv = 0.0f;
            }
        }
    }
    *prim = rid; *th = rt; *uh = ru; *vh = rv;
}

static int brute_any_one(const or_scene* s, const float o[3], const float d[3], float tnear,
                         float tfar) {
    for (uint32_t k = 0; k < s->nt; ++k) {
        float v0[3], e1[3], e2[3], t, u, v;
        tri_edges(s, k, v0, e1, e2);
        if (or_tri_f32(o, d, v0, e1, e2, &t, &u, &v) && t > tnear && t < tfar) return 1;
    }
    for (uint32_t j = 0; j < s->ns; ++j) {
        const float* sp = s->spheres + 4 * (size_t)j;
        float t;
        if (or_sphere_f32(o, d, sp, sp[3] * sp[3], tnear, &t) && in_range(t, tnear, tfar)) return 1;
    }
    return 0;
}

/* FP64 closest over the same candidates (all in FP64), for error reports. */
static void brute_closest_f64_one(const or_scene* s, const float o[3], const float d[3],
                                  float tnear, float tfar, uint32_t* prim, double* th,
                                  double* uh, double* vh) {
    double od[3] = {o[0], o[1], o[2]}, dd[3] = {d[0], d[1], d[2]};
    double bt = INFINITY, bu = 0, bv = 0;
    uint32_t bid = OR_MISS;
    for (uint32_t k = 0; k < s->nt; ++k) {
        float v0f[3], e1f[3], e2f[3];
        tri_edges(s, k, v0f, e1f, e2f);
        double v0[3] = {v0f[0], v0f[1], v0f[2]}, e1[3] = {e1f[0], e1f[1], e1f[2]},
               e2[3] = {e2f[0], e2f[1], e2f[2]}, t, u, v;
        if (or_tri_f64(od, dd, v0, e1, e2, &t, &u, &v) && t > tnear) {
            if (t < bt || (t == bt && k < bid)) {
                bt = t; bu = u; bv = v; bid = k;
            }
        }
    }
    uint32_t rid = OR_MISS;
    double rt = INFINITY, ru = 0, rv = 0;
    double nl = (double)tnear * (1.0 + 1e-6), fl = (double)tfar * (1.0 - 1e-6);
    if (bid != OR_MISS && bt > nl && bt < fl) {
        rid = bid; rt = bt; ru = bu; rv = bv;
    }
    for (uint32_t j = 0; j < s->ns; ++j) {
        const float* sp = s->spheres + 4 * (size_t)j;
        double c[3] = {sp[0], sp[1], sp[2]}, t;
        if (or_sphere_f64(od, dd, c, sp[3], tnear, &t) && t > nl && t < fl) {
            uint32_t id = s->nt + j;
            if (t < rt || (t == rt && id < rid)) {
                rid = id; rt = t; ru = 0; rv = 0;
            }
        }
    }
    *prim = rid; *th = rt; *uh = ru; *vh = rv;
}

/* ------------------------------------------------------------------------ */
/* QBVH traversal over the flattened layout (128 B nodes, 48 B prims)       */
/* ------------------------------------------------------------------------ */

typedef struct {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t reserved[4];
} or_node;

typedef struct {
    float a[4], b[4], c[4];
} or_prim;

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float safe_inv(float d) {
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}

#define OR_IZE 1.0000003576278687f
#define OR_SLACK (1.0f + 1.0f / 65536.0f)
#define OR_STACK 256

typedef struct {
    uint64_t nodes, prims;
} or_counts;

/* Child entry distances of one node (INFINITY = culled). */
static void node_children(const or_node* nd, const float o[3], const float inv[3], float tnear,
                          float cull, float t[4]) {
    for (int i = 0; i < 4; ++i) {
        if (nd->child[i] == OR_EMPTY) { t[i] = INFINITY; continue; }
        float ax = (nd->lo_x[i] - o[0]) * inv[0], bx = (nd->hi_x[i] - o[0]) * inv[0];
        float ay = (nd->lo_y[i] - o[1]) * inv[1], by = (nd->hi_y[i] - o[1]) * inv[1];
        float az = (nd->lo_z[i] - o[2]) * inv[2], bz = (nd->hi_z[i] - o[2]) * inv[2];
        float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tnear));
        float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz)) * OR_IZE;
        t[i] = (tn <= fminf(tf, cull)) ? tn : INFINITY;
    }
}

/* tied (optional): 1 when the reported hit distance is reached by two or
 * more primitives (LiAR keeps the first found of them, object_trees.h:243-244,
 * this library the smallest id).  Culling keeps every box whose entry is <=
 * the best distance (times the slack), so every primitive at that distance
 * is tested and the count is complete. */
static int qbvh_closest_one(const or_node* nodes, const or_prim* prims, uint32_t n_nodes,
                            const float o[3], const float d[3], float tnear, float tfar,
                            uint32_t* prim, float* th, float* uh, float* vh, or_counts* cnt, uint8_t* tied) {
    float inv[3] = {safe_inv(d[0]), safe_inv(d[1]), safe_inv(d[2])};
    float far_lim = tfar * (1.0f - OR_EPS), near_lim = tnear * (1.0f + OR_EPS);
    float tri_t = far_lim, tri_u = 0.0f, tri_v = 0.0f, sph_t = far_lim;
    uint32_t tri_id = OR_MISS, sph_id = OR_MISS;
    int doomed = 0, tri_tie = 0, sph_tie = 0;
    uint32_t st_e[OR_STACK];
    float st_t[OR_STACK];
    int sp = 0;
    if (n_nodes) { st_e[sp] = 0; st_t[sp] = tnear; ++sp; }
    while (sp > 0) {
        --sp;
        uint32_t e = st_e[sp];
        float te = st_t[sp];
        float cull = fminf(doomed ? INFINITY : tri_t, sph_t) * OR_SLACK;
        if (te > cull) continue;
        if (!(e & 0x80000000u)) {
            const or_node* nd = nodes + e;
            float t[4];
            int order[4] = {0, 1, 2, 3};
            cnt->nodes++;
            node_children(nd, o, inv, tnear, cull, t);
            /* push far-to-near so the nearest child is visited next */
            for (int i = 1; i < 4; ++i)
                for (int j = i; j > 0 && t[order[j]] < t[order[j - 1]]; --j) {
                    int x = order[j]; order[j] = order[j - 1]; order[j - 1] = x;
                }
            for (int i = 3; i >= 0; --i) {
                int c = order[i];
                if (t[c] == INFINITY) continue;
                if (sp >= OR_STACK) return -1;
                st_e[sp] = nd->child[c];
                st_t[sp] = t[c];
                ++sp;
            }
        } else {
            uint32_t first = (e >> 2) & 0x1FFFFFFFu, n = (e & 3u) + 1u;
            for (uint32_t k = 0; k < n; ++k) {
                const or_prim* q = prims + first + k;
                uint32_t id = fbits(q->a[3]);
                cnt->prims++;
                if (fbits(q->b[3]) == 0u) {
                    float t, u, v;
                    if (or_tri_f32(o, d, q->a, q->b, q->c, &t, &u, &v) && t > tnear && t < far_lim) {
                        if (t == tri_t && tri_id != OR_MISS && id != tri_id) tri_tie = 1;
                        if (t < tri_t || (t == tri_t && id < tri_id)) {
                            if (t < tri_t) tri_tie = 0;
                            tri_t = t; tri_u = u; tri_v = v; tri_id = id;
                            doomed = !(t > near_lim);
                        }
                    }
                } else {
                    float t;
                    if (or_sphere_f32(o, d, q->a, q->b[1], tnear, &t) && in_range(t, tnear, tfar)) {
                        if (t == sph_t && sph_id != OR_MISS && id != sph_id) sph_tie = 1;
                        if (t < sph_t || (t == sph_t && id < sph_id)) {
                            if (t < sph_t) sph_tie = 0;
                            sph_t = t; sph_id = id;
                        }
                    }
                }
            }
        }
    }
    int tri_ok = tri_id != OR_MISS && !doomed;
    int tie = 0;
    if (tri_ok && (sph_id == OR_MISS || tri_t < sph_t || (tri_t == sph_t && tri_id < sph_id))) {
        *prim = tri_id; *th = tri_t; *uh = tri_u; *vh = tri_v;
        tie = tri_tie || (sph_id != OR_MISS && sph_t == tri_t);
    } else if (sph_id != OR_MISS) {
        *prim = sph_id; *th = sph_t; *uh = 0.0f; *vh = 0.0f;
        tie = sph_tie || (tri_ok && tri_t == sph_t);
    } else {
        *prim = OR_MISS; *th = INFINITY; *uh = 0.0f; *vh = 0.0f;
    }
    /* bit 0: a tie; bit 1: LiAR's interval quirk decided the mesh's answer
     * (its nearest triangle beyond near lay within near (1 + eps), so the
     * mesh reports no hit, triangle_mesh.cpp:301-302) */
    if (tied) *tied = (uint8_t)(tie | (tri_id != OR_MISS && doomed ? 2 : 0));
    return 0;
}

static int qbvh_any_one(const or_node* nodes, const or_prim* prims, uint32_t n_nodes,
                        const float o[3], const float d[3], float tnear, float tfar, or_counts* cnt) {
    float inv[3] = {safe_inv(d[0]), safe_inv(d[1]), safe_inv(d[2])};
    float cull = tfar * OR_SLACK;
    uint32_t st_e[OR_STACK];
    int sp = 0;
    if (n_nodes) st_e[sp++] = 0;
    while (sp > 0) {
        uint32_t e = st_e[--sp];
        if (!(e & 0x80000000u)) {
            const or_node* nd = nodes + e;
            float t[4];
            cnt->nodes++;
            node_children(nd, o, inv, tnear, cull, t);
            for (int i = 3; i >= 0; --i) {
                if (t[i] == INFINITY) continue;
                if (sp >= OR_STACK) return -1;
                st_e[sp++] = nd->child[i];
            }
        } else {
            uint32_t first = (e >> 2) & 0x1FFFFFFFu, n = (e & 3u) + 1u;
            for (uint32_t k = 0; k < n; ++k) {
                const or_prim* q = prims + first + k;
                cnt->prims++;
                if (fbits(q->b[3]) == 0u) {
                    float t, u, v;
                    if (or_tri_f32(o, d, q->a, q->b, q->c, &t, &u, &v) && t > tnear && t < tfar) return 1;
                } else {
                    float t;
                    if (or_sphere_f32(o, d, q->a, q->b[1], tnear, &t) && in_range(t, tnear, tfar)) return 1;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* threaded drivers (exported)                                              */
/* ------------------------------------------------------------------------ */

enum { JOB_BRUTE_CLOSEST, JOB_BRUTE_ANY, JOB_BRUTE_F64, JOB_QBVH_CLOSEST, JOB_QBVH_ANY, JOB_INST_CLOSEST,
       JOB_INST_ANY };

/* ---- instances (src/liar/scenery/transformation.cpp:117-131) -------------- */

/* World -> local of a row-major 3x4 local-to-world affine, in FP64 and then
 * rounded to FP32 (the library does the same); 0 when singular. */
static int affine_inverse_f32(const float* m, float* w) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    const double A = e * i - f * h, B = f * g - d * i, Cc = d * h - e * g;
    const double det = a * A + b * B + c * Cc;
    if (!(fabs(det) > 0.0) || !isfinite(det)) return 0;
    const double inv[9] = {A / det, (c * h - b * i) / det, (b * f - c * e) / det,
                           B / det, (a * i - c * g) / det, (c * d - a * f) / det,
                           Cc / det, (b * g - a * h) / det, (a * e - b * d) / det};
    const double t[3] = {m[3], m[7], m[11]};
    for (int r = 0; r < 3; ++r) {
        w[4 * r + 0] = (float)inv[3 * r + 0];
        w[4 * r + 1] = (float)inv[3 * r + 1];
        w[4 * r + 2] = (float)inv[3 * r + 2];
        w[4 * r + 3] = (float)(-(inv[3 * r] * t[0] + inv[3 * r + 1] * t[1] + inv[3 * r + 2] * t[2]));
    }
    return 1;
}

/* The local ray in FP32: origin ((W00 ox + W01 oy) + W02 oz) + W03 per row,
 * direction W d / s with s = sqrt((x x + y y) + z z). */
static float local_ray(const float* w, const float o[3], const float d[3], float lo[3], float ld[3]) {
    for (int r = 0; r < 3; ++r) lo[r] = ((w[4 * r] * o[0] + w[4 * r + 1] * o[1]) + w[4 * r + 2] * o[2]) + w[4 * r + 3];
    float x = (w[0] * d[0] + w[1] * d[1]) + w[2] * d[2];
    float y = (w[4] * d[0] + w[5] * d[1]) + w[6] * d[2];
    float z = (w[8] * d[0] + w[9] * d[1]) + w[10] * d[2];
    float s = sqrtf((x * x + y * y) + z * z);
    ld[0] = x / s;
    ld[1] = y / s;
    ld[2] = z / s;
    return s;
}

typedef struct {
    int kind;
    const or_scene* scene;
    const or_node* nodes;
    const or_prim* prims;
    uint32_t n_nodes;
    or_rays rays;
    uint32_t begin, end;
    uint32_t* prim;
    float *t, *u, *v;
    double *t64, *u64, *v64;
    uint8_t* occ;
    uint8_t* tied;
    or_counts counts;
    int status;
    /* instances */
    const or_scene* geoms;
    const uint32_t* inst_geom;
    const float* inst_w;  /* n_inst x 12, world -> local */
    uint32_t n_inst;
    uint32_t* inst_out;
} or_job;

static int g_inst_plant = 0;
/* Test hook: 1 forgets the division of the local distance by the scale. */
void or_set_inst_mistake(int k) { g_inst_plant = k; }

/* List over the instances: each intersected in its local space with the
 * single-level rules, world t = t_local / s, nearest t then smaller
 * instance index then smaller primitive id. */
static void inst_closest_one(const or_job* j, const float o[3], const float d[3], float tn, float tf, uint32_t* prim,
                             float* th, float* uh, float* vh, uint32_t* inst) {
    float bt = INFINITY, bu = 0.0f, bv = 0.0f;
    uint32_t bp = OR_MISS, bi = OR_MISS;
    for (uint32_t k = 0; k < j->n_inst; ++k) {
        const or_scene* g = &j->geoms[j->inst_geom[k]];
        if (g->nt + g->ns == 0) continue;
        float lo[3], ld[3];
        const float s = local_ray(j->inst_w + 12 * (size_t)k, o, d, lo, ld);
        uint32_t p;
        float t, u, v;
        brute_closest_one(g, lo, ld, tn * s, tf * s, &p, &t, &u, &v);
        if (p == OR_MISS) continue;
        const float tw = g_inst_plant == 1 ? t : t / s;
        if (tw < bt || (tw == bt && (k < bi || (k == bi && p < bp)))) {
            bt = tw; bu = u; bv = v; bp = p; bi = k;
        }
    }
    *prim = bp; *th = bp == OR_MISS ? INFINITY : bt; *uh = bu; *vh = bv; *inst = bi;
}

static int inst_any_one(const or_job* j, const float o[3], const float d[3], float tn, float tf) {
    for (uint32_t k = 0; k < j->n_inst; ++k) {
        const or_scene* g = &j->geoms[j->inst_geom[k]];
        if (g->nt + g->ns == 0) continue;
        float lo[3], ld[3];
        const float s = local_ray(j->inst_w + 12 * (size_t)k, o, d, lo, ld);
        if (brute_any_one(g, lo, ld, tn * s, tf * s)) return 1;
    }
    return 0;
}

static void* run_job(void* arg) {
    or_job* j = (or_job*)arg;
    for (uint32_t i = j->begin; i < j->end; ++i) {
        float o[3] = {j->rays.ox[i], j->rays.oy[i], j->rays.oz[i]};
        float d[3] = {j->rays.dx[i], j->rays.dy[i], j->rays.dz[i]};
        float tn = j->rays.tnear[i], tf = j->rays.tfar[i];
        switch (j->kind) {
        case JOB_BRUTE_CLOSEST:
            brute_closest_one(j->scene, o, d, tn, tf, j->prim + i, j->t + i, j->u + i, j->v + i);
            break;
        case JOB_BRUTE_ANY:
            j->occ[i] = (uint8_t)brute_any_one(j->scene, o, d, tn, tf);
            break;
        case JOB_BRUTE_F64:
            brute_closest_f64_one(j->scene, o, d, tn, tf, j->prim + i, j->t64 + i, j->u64 + i,
                                  j->v64 + i);
            break;
        case JOB_QBVH_CLOSEST:
            if (qbvh_closest_one(j->nodes, j->prims, j->n_nodes, o, d, tn, tf, j->prim + i, j->t + i,
                                 j->u + i, j->v + i, &j->counts, j->tied ? j->tied + i : NULL))
                j->status = -1;
            break;
        case JOB_INST_CLOSEST:
            inst_closest_one(j, o, d, tn, tf, j->prim + i, j->t + i, j->u + i, j->v + i, j->inst_out + i);
            break;
        case JOB_INST_ANY:
            j->occ[i] = (uint8_t)inst_any_one(j, o, d, tn, tf);
            break;
        case JOB_QBVH_ANY: {
            int r = qbvh_any_one(j->nodes, j->prims, j->n_nodes, o, d, tn, tf, &j->counts);
            if (r < 0) j->status = -1;
            j->occ[i] = (uint8_t)(r > 0);
            break;
        }
        }
    }
    return NULL;
}

static int run_threads(or_job proto, uint32_t n, int nthreads, uint64_t* nodes_visited,
                       uint64_t* prims_tested) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > n) nthreads = n ? (int)n : 1;
    or_job jobs[256];
    pthread_t th[256];
    uint32_t chunk = (n + (uint32_t)nthreads - 1) / (uint32_t)nthreads;
    for (int k = 0; k < nthreads; ++k) {
        jobs[k] = proto;
        jobs[k].begin = (uint32_t)k * chunk < n ? (uint32_t)k * chunk : n;
        jobs[k].end = jobs[k].begin + chunk < n ? jobs[k].begin + chunk : n;
        jobs[k].counts.nodes = jobs[k].counts.prims = 0;
        jobs[k].status = 0;
        if (k > 0) pthread_create(&th[k], NULL, run_job, &jobs[k]);
    }
    run_job(&jobs[0]);
    int status = jobs[0].status;
    uint64_t nv = jobs[0].counts.nodes, np = jobs[0].counts.prims;
    for (int k = 1; k < nthreads; ++k) {
        pthread_join(th[k], NULL);
        if (jobs[k].status) status = jobs[k].status;
        nv += jobs[k].counts.nodes;
        np += jobs[k].counts.prims;
    }
    if (nodes_visited) *nodes_visited = nv;
    if (prims_tested) *prims_tested = np;
    return status;
}

static or_rays mk_rays(const float* const r[8]) {
    or_rays x = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]};
    return x;
}

/* rays[8] = {ox, oy, oz, dx, dy, dz, tnear, tfar} (host arrays of n floats) */
int or_brute_closest(const float* verts, const uint32_t* tris, uint32_t nt, const float* spheres,
                     uint32_t ns, const float* const rays[8], uint32_t n, uint32_t* prim, float* t,
                     float* u, float* v, int nthreads) {
    or_scene s = {verts, tris, nt, spheres, ns};
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = JOB_BRUTE_CLOSEST;
    j.scene = &s;
    j.rays = mk_rays(rays);
    j.prim = prim; j.t = t; j.u = u; j.v = v;
    return run_threads(j, n, nthreads, NULL, NULL);
}

/* Instanced brute force.  geoms: n_geom (verts, tris, nt, spheres, ns)
 * descriptions (layout of or_scene); inst_geom / local_to_world: per
 * instance.  any != 0: occluded only.  Returns -2 for a singular transform. */
int or_inst_brute(const or_scene* geoms, uint32_t n_geom, const uint32_t* inst_geom, const float* local_to_world,
                  uint32_t n_inst, const float* const rays[8], uint32_t n, int any, uint32_t* prim, float* t,
                  float* u, float* v, uint32_t* inst, uint8_t* occ, int nthreads) {
    float* w = (float*)malloc(sizeof(float) * 12 * (size_t)(n_inst ? n_inst : 1));
    for (uint32_t k = 0; k < n_inst; ++k) {
        if (inst_geom[k] >= n_geom || !affine_inverse_f32(local_to_world + 12 * (size_t)k, w + 12 * (size_t)k)) {
            free(w);
            return -2;
        }
    }
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = any ? JOB_INST_ANY : JOB_INST_CLOSEST;
    j.geoms = geoms;
    j.inst_geom = inst_geom;
    j.inst_w = w;
    j.n_inst = n_inst;
    j.rays = mk_rays(rays);
    j.prim = prim; j.t = t; j.u = u; j.v = v; j.inst_out = inst; j.occ = occ;
    int rc = run_threads(j, n, nthreads, NULL, NULL);
    free(w);
    return rc;
}

int or_brute_any(const float* verts, const uint32_t* tris, uint32_t nt, const float* spheres,
                 uint32_t ns, const float* const rays[8], uint32_t n, uint8_t* occ, int nthreads) {
    or_scene s = {verts, tris, nt, spheres, ns};
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = JOB_BRUTE_ANY;
    j.scene = &s;
    j.rays = mk_rays(rays);
    j.occ = occ;
    return run_threads(j, n, nthreads, NULL, NULL);
}

int or_brute_closest_f64(const float* verts, const uint32_t* tris, uint32_t nt,
                         const float* spheres, uint32_t ns, const float* const rays[8], uint32_t n,
                         uint32_t* prim, double* t, double* u, double* v, int nthreads) {
    or_scene s = {verts, tris, nt, spheres, ns};
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = JOB_BRUTE_F64;
    j.scene = &s;
    j.rays = mk_rays(rays);
    j.prim = prim; j.t64 = t; j.u64 = u; j.v64 = v;
    return run_threads(j, n, nthreads, NULL, NULL);
}

/* QBVH closest/any over the flattened arrays; also returns the total number
 * of inner nodes visited and primitives tested (for bytes-per-ray). */
int or_qbvh_closest(const void* nodes, uint32_t n_nodes, const void* prims,
                    const float* const rays[8], uint32_t n, uint32_t* prim, float* t, float* u,
                    float* v, int nthreads, uint64_t* nodes_visited, uint64_t* prims_tested) {
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = JOB_QBVH_CLOSEST;
    j.nodes = (const or_node*)nodes;
    j.prims = (const or_prim*)prims;
    j.n_nodes = n_nodes;
    j.rays = mk_rays(rays);
    j.prim = prim; j.t = t; j.u = u; j.v = v;
    return run_threads(j, n, nthreads, nodes_visited, prims_tested);
}

/* or_qbvh_closest plus tied[i]: bit 0 where ray i's hit distance is shared
 * by two or more primitives, bit 1 where the interval quirk rejected the
 * mesh's nearest triangle. */
int or_qbvh_closest_ties(const void* nodes, uint32_t n_nodes, const void* prims,
                         const float* const rays[8], uint32_t n, uint32_t* prim, float* t, float* u,
                         float* v, uint8_t* tied, int nthreads, uint64_t* nodes_visited, uint64_t* prims_tested) {
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = JOB_QBVH_CLOSEST;
    j.nodes = (const or_node*)nodes;
    j.prims = (const or_prim*)prims;
    j.n_nodes = n_nodes;
    j.rays = mk_rays(rays);
    j.prim = prim; j.t = t; j.u = u; j.v = v;
    j.tied = tied;
    return run_threads(j, n, nthreads, nodes_visited, prims_tested);
}

int or_qbvh_any(const void* nodes, uint32_t n_nodes, const void* prims, const float* const rays[8],
                uint32_t n, uint8_t* occ, int nthreads, uint64_t* nodes_visited,
                uint64_t* prims_tested) {
    or_job j;
    memset(&j, 0, sizeof(j));
    j.kind = JOB_QBVH_ANY;
    j.nodes = (const or_node*)nodes;
    j.prims = (const or_prim*)prims;
    j.n_nodes = n_nodes;
    j.rays = mk_rays(rays);
    j.occ = occ;
    return run_threads(j, n, nthreads, nodes_visited, prims_tested);
}

/* ------------------------------------------------------------------------ */
/* QBVH invariants                                                          */
/* ------------------------------------------------------------------------ */

static char g_msg[512];
const char* or_last_message(void) { return g_msg; }

static int box_has(const or_node* nd, int i, const float lo[3], const float hi[3]) {
    return nd->lo_x[i] <= lo[0] && nd->lo_y[i] <= lo[1] && nd->lo_z[i] <= lo[2] &&
           nd->hi_x[i] >= hi[0] && nd->hi_y[i] >= hi[1] && nd->hi_z[i] >= hi[2];
}

/* Checks: every node reachable exactly once from the root, every prim in
 * exactly one leaf, leaf size 1..4, each child box contains its subtree
 * (child node boxes / prim AABBs), prim records equal the input geometry.
 * Returns 0 when valid; else -1 and or_last_message(). */
int or_qbvh_check(const void* nodes_v, uint32_t n_nodes, const void* prims_v, uint32_t n_prims,
                  const float* verts, const uint32_t* tris, uint32_t nt, const float* spheres,
                  uint32_t ns, uint32_t* max_depth_out) {
    const or_node* nodes = (const or_node*)nodes_v;
    const or_prim* prims = (const or_prim*)prims_v;
    g_msg[0] = 0;
    if (n_prims != nt + ns) {
        snprintf(g_msg, sizeof g_msg, "n_prims %u != nt + ns %u", n_prims, nt + ns);
        return -1;
    }
    if (n_prims == 0) { *max_depth_out = 0; return n_nodes == 0 ? 0 : -1; }
    uint8_t* seen_node = (uint8_t*)calloc(n_nodes, 1);
    uint8_t* seen_rec = (uint8_t*)calloc(n_prims, 1);
    uint8_t* seen_id = (uint8_t*)calloc(n_prims, 1);
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * 2 * ((size_t)n_nodes + 1));
    int rc = 0;
    uint32_t maxd = 0;
    int sp = 0;
    stack[sp++] = 0; stack[sp++] = 0;
    seen_node[0] = 1;
    while (sp > 0 && rc == 0) {
        uint32_t depth = stack[--sp];
        uint32_t ni = stack[--sp];
        if (depth > maxd) maxd = depth;
        const or_node* nd = nodes + ni;
        for (int i = 0; i < 4 && rc == 0; ++i) {
            uint32_t c = nd->child[i];
            if (c == OR_EMPTY) continue;
            if (!(c & 0x80000000u)) {
                if (c >= n_nodes || seen_node[c]) {
                    snprintf(g_msg, sizeof g_msg, "node %u child %d -> bad/duplicate node %u", ni, i, c);
                    rc = -1; break;
                }
                seen_node[c] = 1;
                const or_node* ch = nodes + c;
                for (int k = 0; k < 4; ++k) {
                    if (ch->child[k] == OR_EMPTY) continue;
                    float lo[3] = {ch->lo_x[k], ch->lo_y[k], ch->lo_z[k]};
                    float hi[3] = {ch->hi_x[k], ch->hi_y[k], ch->hi_z[k]};
                    if (!box_has(nd, i, lo, hi)) {
                        snprintf(g_msg, sizeof g_msg, "node %u child %d box misses grandchild %d", ni, i, k);
                        rc = -1; break;
                    }
                }
                stack[sp++] = c; stack[sp++] = depth + 1;
            } else {
                uint32_t first = (c >> 2) & 0x1FFFFFFFu, n = (c & 3u) + 1u;
                if (first + n > n_prims) {
                    snprintf(g_msg, sizeof g_msg, "leaf range %u+%u out of %u", first, n, n_prims);
                    rc = -1; break;
                }
                for (uint32_t k = 0; k < n && rc == 0; ++k) {
                    const or_prim* q = prims + first + k;
                    uint32_t id = fbits(q->a[3]);
                    if (seen_rec[first + k] || id >= n_prims || seen_id[id]) {
                        snprintf(g_msg, sizeof g_msg, "prim record %u (id %u) duplicated/invalid", first + k, id);
                        rc = -1; break;
                    }
                    seen_rec[first + k] = 1;
                    seen_id[id] = 1;
                    float lo[3], hi[3];
                    if (id < nt) {
                        float v0[3], e1[3], e2[3];
                        or_scene s = {verts, tris, nt, spheres, ns};
                        tri_edges(&s, id, v0, e1, e2);
                        for (int a = 0; a < 3; ++a) {
                            if (q->a[a] != v0[a] || q->b[a] != e1[a] || q->c[a] != e2[a] ||
                                fbits(q->b[3]) != 0u) {
                                snprintf(g_msg, sizeof g_msg, "triangle record %u differs from input", id);
                                rc = -1;
                            }
                            const float* p0 = verts + 3 * (size_t)tris[3 * (size_t)id + 0];
                            const float* p1 = verts + 3 * (size_t)tris[3 * (size_t)id + 1];
                            const float* p2 = verts + 3 * (size_t)tris[3 * (size_t)id + 2];
                            lo[a] = fminf(p0[a], fminf(p1[a], p2[a]));
                            hi[a] = fmaxf(p0[a], fmaxf(p1[a], p2[a]));
                        }
                    } else {
                        const float* sp4 = spheres + 4 * (size_t)(id - nt);
                        if (q->a[0] != sp4[0] || q->a[1] != sp4[1] || q->a[2] != sp4[2] ||
                            q->b[0] != sp4[3] || fbits(q->b[3]) != 1u) {
                            snprintf(g_msg, sizeof g_msg, "sphere record %u differs from input", id);
                            rc = -1;
                        }
                        for (int a = 0; a < 3; ++a) {
                            lo[a] = sp4[a] - sp4[3];
                            hi[a] = sp4[a] + sp4[3];
                        }
                    }
                    if (rc == 0 && !box_has(nd, i, lo, hi)) {
                        snprintf(g_msg, sizeof g_msg, "node %u child %d box misses prim %u", ni, i, id);
                        rc = -1;
                    }
                }
            }
        }
    }
    if (rc == 0) {
        for (uint32_t i = 0; i < n_nodes; ++i)
            if (!seen_node[i]) { snprintf(g_msg, sizeof g_msg, "node %u unreachable", i); rc = -1; break; }
        for (uint32_t i = 0; i < n_prims && rc == 0; ++i)
            if (!seen_id[i]) { snprintf(g_msg, sizeof g_msg, "prim %u missing", i); rc = -1; }
    }
    free(seen_node); free(seen_rec); free(seen_id); free(stack);
    *max_depth_out = maxd;
    return rc;
}
