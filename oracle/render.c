/*
 * render.c -- CPU oracle of the spectral path-tracing integrand.  TEST
 * INFRASTRUCTURE ONLY (see oracle.c).
 *
 * A scalar, recursive restatement in FP64 of LiAR's AdjointPhotonTracer
 * (src/liar/tracers/adjoint_photon_tracer.cpp:84-292) over the flattened
 * scene description of include/liar_amd.h, with:
 *   - Observer wavelength sampling and the Single-mode XYZ estimator
 *     (src/liar/kernel/observer.cpp:57-129,283-317, spectral.cpp:212-221),
 *   - Lambert / Mirror / Dielectric / Conductor / CookTorrance(GGX, Beckmann)
 *     BSDFs (src/liar/shaders/, as cited per function),
 *   - LightArea (parallelogram), LightSky (constant) and a distant sun disc,
 *     light choice proportional to power (src/liar/kernel/light_context.cpp:272-331),
 *   - the Raster box filter, accumulated in FP64 (src/liar/output/raster.cpp:547-606).
 * Intersections are brute force over all primitives in FP64 with LiAR's
 * per-object interval rules.
 *
 * DirectLighting (src/liar/tracers/direct_lighting.cpp:119-377 with
 * RayTracer::estimateLightContribution, src/liar/kernel/ray_tracer.cpp:188-268)
 * is restated as a recursion over the ray tree with brute-force FP64 shadow
 * rays (or_render_direct).
 *
 * The random numbers are the same Philox4x32-10 streams, keyed the same way,
 * as the GPU renderer; sampling warps are the published maps named in
 * liar_amd/csrc/lr_render.h (Lass's own maps are parity unpinned).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RMISS 0xFFFFFFFFu
#define RNONE 0xFFFFFFFFu
static const double PI = 3.14159265358979323846;
static const double TOL = 1e-6;

/* Test hook, a planted mistake that a pinning test must detect; never set
 * outside tests.  1: drop both one-sample MIS weights (furnace tests);
 * 2: drop the DirectLighting light-estimate weights; 3: CookTorrance pdf
 * without the half-vector Jacobian 1 / (4 wo.h) (chi-square test);
 * 5: microfacet D without its 1/pi (normalisation test); 6: Halton BSDF
 * samples with one prime for both coordinates (convergence test). */
static int g_plant = 0;
void or_set_planted_mistake(int k) { g_plant = k; }

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---------------------------- */
static void philox(uint32_t k0, uint32_t k1, const uint32_t ctr[4], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
void or_philox(uint32_t k0, uint32_t k1, const uint32_t* ctr, uint32_t* out) { philox(k0, k1, ctr, out); }
static double u01(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

/* ---- scene description (layouts of include/liar_amd.h) ---------------- */
typedef struct { uint32_t kind, count; const float *w, *v; float value, lo, hi, temperature, scale; } or_spectrum;
typedef struct { uint32_t type, refl, trans, eta, kappa, emission, light; float roughness, roughness_v, fuzz; } or_material;
typedef struct {
    uint32_t type, radiance, double_sided;
    float origin[3], dx[3], dy[3], direction[3], angle;
    uint32_t samples; /* DirectLighting emission samples (0: 9 parallelogram or mesh, 1 otherwise) */
    uint32_t first_triangle, n_triangles; /* mesh light: its triangles' primitive ids */
} or_light;
typedef struct { float position[3], direction[3], sky[3], fov, aspect, near_limit; } or_camera;
typedef struct {
    uint32_t width, height, spp, wavelengths, max_depth, rr_depth;
    float rr_survival, lambda_min, lambda_max;
    uint64_t seed;
    uint32_t chunk_paths;
    float max_sample_luminance;
    uint32_t sampler, sampler_generations; /* 1: Halton (lr_render_params) */
    uint64_t sample_offset;
} or_params;

/* ---- Halton sampler (src/liar/samplers/halton.cpp:180-230, halton.h:41-66) -- */
uint32_t or_halton_tables(uint16_t* perm, uint32_t* off);  /* halton_tables.cpp */
static uint16_t* g_hperm;
static uint32_t g_hoff[285];
static pthread_once_t g_honce = PTHREAD_ONCE_INIT;
static void halton_init(void) {
    uint32_t n = or_halton_tables(NULL, g_hoff);
    g_hperm = (uint16_t*)malloc(sizeof(uint16_t) * n);
    or_halton_tables(g_hperm, g_hoff);
}
/* ScrambledRadicalInverse: digit d of i (least significant first) adds
 * perm[d] b^-(k+1), in FP32 in that order (perm NULL: plain). */
static float radical_inverse(uint64_t i, uint32_t b, const uint16_t* perm) {
    const float inv = 1.0f / (float)b;
    float w = inv, r = 0.0f;
    while (i) {
        uint32_t d = (uint32_t)(i % b);
        i /= b;
        r = r + (float)(perm ? perm[d] : d) * w;
        w = w * inv;
    }
    return r;
}
static float halton_sub(uint64_t j, uint32_t prime_index) {
    return radical_inverse(j, g_hoff[prime_index + 1] - g_hoff[prime_index], g_hperm + g_hoff[prime_index]);
}
/* probe: sequence index j, plain in base `base` (prime_index < 0) or on the
 * scrambled sub-sequence of prime index prime_index */
float or_halton_sample(uint64_t j, uint32_t base, int prime_index) {
    pthread_once(&g_honce, halton_init);
    if (prime_index < 0) return radical_inverse(j, base, NULL);
    return prime_index < 284 ? halton_sub(j, (uint32_t)prime_index) : -1.0f;
}

typedef struct { double x, y, z; } V;
static V v(double x, double y, double z) { V r = {x, y, z}; return r; }
static V add(V a, V b) { return v(a.x + b.x, a.y + b.y, a.z + b.z); }
static V sub(V a, V b) { return v(a.x - b.x, a.y - b.y, a.z - b.z); }
static V mul(double s, V a) { return v(s * a.x, s * a.y, s * a.z); }
static double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V cross(V a, V b) { return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static V nrm(V a) { return mul(1.0 / sqrt(dot(a, a)), a); }
static V f3(const float* p) { return v(p[0], p[1], p[2]); }

typedef struct {
    const float* verts; const uint32_t* tris; uint32_t nt;
    const float* tri_normals;  /* 9 per triangle (corner normals) or NULL; all-zero = none */
    const float* spheres; uint32_t ns;
    const uint32_t* prim_mat;
    const or_material* mats;
    const or_spectrum* spectra; uint32_t n_spectra;
    const or_light* lights; uint32_t n_lights;
    /* derived */
    double* light_sel; double* light_cdf; double* light_area; V* light_n; V* light_du; V* light_dv;
    double* light_cos; double* light_invsolid;
    uint32_t* light_ns; /* DirectLighting samples per light */
    double** light_tcdf; /* mesh lights: cumulative area fractions of their triangles */
    double* spec_reflscale;
    /* observer restricted to the domain */
    int on; double *ow, *ocdf, *oxyz;
    double R;
} Scene;

/* ---- spectra --------------------------------------------------------------- */
static double planck(double w_m, double t) {
    const double h = 6.626070040e-34, c = 299792458.0, kb = 1.3806488e-23;
    return 2 * h * c * c / (pow(w_m, 5) * expm1(h * c / (kb * w_m * t)));
}
static double spec_eval(const Scene* S, uint32_t i, double w, int reflectant) {
    if (i == RNONE) return 0.0;
    const or_spectrum* s = &S->spectra[i];
    double val = 0.0;
    switch (s->kind) {
    case 0: return s->value;
    case 1: {
        uint32_t k = 0;
        while (k < s->count && s->w[k] <= w) ++k;  /* upper_bound */
        if (k == 0 || k == s->count) return 0.0;
        double t = (w - s->w[k - 1]) / ((double)s->w[k] - s->w[k - 1]);
        val = s->v[k - 1] + t * ((double)s->v[k] - s->v[k - 1]);
        break;
    }
    case 2: {
        if (!(w >= s->lo && w < s->hi)) return 0.0;
        uint32_t k = (uint32_t)((w - s->lo) / ((double)s->hi - s->lo) * s->count);
        if (k >= s->count) k = s->count - 1;
        val = s->v[k];
        break;
    }
    case 3: {
        double peak = 2.897771955e-3 / s->temperature;
        return s->scale * planck(w * 1e-9, s->temperature) / planck(peak, s->temperature);
    }
    }
    return reflectant ? val * S->spec_reflscale[i] : val;
}

/* Observer restricted to [lo, hi]: CDF from the trapezoid of x+y+z, XYZ
 * normalised so that a unit spectrum has Y = 1 (observer.cpp:57-129). */
static int build_observer(Scene* S, const float* w, const float* xyz, int n, double lo, double hi) {
    double a = lo > w[0] ? lo : w[0], b = hi < w[n - 1] ? hi : w[n - 1];
    if (!(b > a)) return -1;
    int m = 0;
    S->ow = (double*)malloc(sizeof(double) * (n + 2));
    S->ocdf = (double*)malloc(sizeof(double) * (n + 2));
    S->oxyz = (double*)malloc(sizeof(double) * 3 * (n + 2));
    S->ow[m++] = a;
    for (int k = 0; k < n; ++k)
        if (w[k] > a && w[k] < b) S->ow[m++] = w[k];
    S->ow[m++] = b;
    for (int j = 0; j < m; ++j) {
        double wl = S->ow[j];
        int k = 0;
        while (k < n && w[k] <= wl) ++k;
        for (int c = 0; c < 3; ++c) {
            double val;
            if (k == 0) val = 0;
            else if (k == n) val = (wl == w[n - 1]) ? xyz[3 * (n - 1) + c] : 0;
            else {
                double t = (wl - w[k - 1]) / ((double)w[k] - w[k - 1]);
                val = xyz[3 * (k - 1) + c] + t * ((double)xyz[3 * k + c] - xyz[3 * (k - 1) + c]);
            }
            S->oxyz[3 * j + c] = val;
        }
    }
    S->ocdf[0] = 0;
    for (int j = 0; j + 1 < m; ++j) {
        double s0 = S->oxyz[3 * j] + S->oxyz[3 * j + 1] + S->oxyz[3 * j + 2];
        double s1 = S->oxyz[3 * j + 3] + S->oxyz[3 * j + 4] + S->oxyz[3 * j + 5];
        S->ocdf[j + 1] = S->ocdf[j] + (s0 + s1) * (S->ow[j + 1] - S->ow[j]);
    }
    double Y = 0;
    for (int j = 0; j < m; ++j) {
        double wp = S->ow[j > 0 ? j - 1 : 0], wn = S->ow[j + 1 < m ? j + 1 : m - 1];
        Y += S->oxyz[3 * j + 1] * (wn - wp) / 2;
    }
    for (int j = 0; j < m; ++j) {
        S->ocdf[j] /= S->ocdf[m - 1];
        for (int c = 0; c < 3; ++c) S->oxyz[3 * j + c] /= Y;
    }
    S->ocdf[m - 1] = 1.0;
    S->on = m;
    return 0;
}

static double obs_sample(const Scene* S, double u, double* pdf) {
    int k = 0;
    while (k < S->on && S->ocdf[k] <= u) ++k;
    if (k == 0 || k >= S->on) { *pdf = 0; return S->ow[0]; }
    double c1 = S->ocdf[k - 1], c2 = S->ocdf[k], dc = c2 - c1;
    double x = (u - c1) / dc;
    *pdf = dc / (S->ow[k] - S->ow[k - 1]);
    return S->ow[k - 1] + x * (S->ow[k] - S->ow[k - 1]);
}

static void obs_xyz(const Scene* S, double w, double out[3]) {
    int k = 0;
    while (k < S->on && S->ow[k] <= w) ++k;
    if (k == 0 || k == S->on) {
        int last = (k == S->on && w == S->ow[S->on - 1]);
        for (int c = 0; c < 3; ++c) out[c] = last ? S->oxyz[3 * (S->on - 1) + c] : 0.0;
        return;
    }
    double t = (w - S->ow[k - 1]) / (S->ow[k] - S->ow[k - 1]);
    for (int c = 0; c < 3; ++c)
        out[c] = S->oxyz[3 * (k - 1) + c] + t * (S->oxyz[3 * k + c] - S->oxyz[3 * (k - 1) + c]);
}

/* ---- intersection (FP64, LiAR per-object rules) ---------------------------- */
static int tri_hit(const Scene* S, uint32_t k, V o, V d, double* t) {
    const float* a = S->verts + 3 * (size_t)S->tris[3 * k];
    const float* b = S->verts + 3 * (size_t)S->tris[3 * k + 1];
    const float* c = S->verts + 3 * (size_t)S->tris[3 * k + 2];
    V v0 = f3(a), e1 = sub(f3(b), v0), e2 = sub(f3(c), v0);
    V p = cross(d, e2);
    double det = dot(e1, p);
    if (det == 0) return 0;
    V s = sub(o, v0);
    double u = dot(s, p) / det;
    if (!(u >= 0 && u <= 1)) return 0;
    V q = cross(s, e1);
    double vv = dot(d, q) / det;
    if (!(vv >= 0 && u + vv <= 1)) return 0;
    *t = dot(e2, q) / det;
    return 1;
}
static int sph_hit(const Scene* S, uint32_t j, V o, V d, double tmin, double* t) {
    const float* s = S->spheres + 4 * (size_t)j;
    V oc = sub(o, f3(s));
    double b = dot(oc, d), c = dot(oc, oc) - (double)s[3] * s[3];
    double disc = b * b - c;
    if (disc < 0) return 0;
    double sq = sqrt(disc), t0 = -b - sq, t1 = -b + sq;
    if (t0 > tmin) { *t = t0; return 1; }
    if (t1 > tmin) { *t = t1; return 1; }
    return 0;
}
static int in_range(double t, double tn, double tf) { return t > tn * (1 + TOL) && t < tf * (1 - TOL); }

/* List{TriangleMesh, Spheres}: returns prim id or RMISS. */
static uint32_t intersect(const Scene* S, V o, V d, double tn, double* th) {
    double bt = INFINITY;
    uint32_t bid = RMISS;
    for (uint32_t k = 0; k < S->nt; ++k) {
        double t;
        if (tri_hit(S, k, o, d, &t) && t > tn && (t < bt || (t == bt && k < bid))) { bt = t; bid = k; }
    }
    uint32_t rid = RMISS;
    double rt = INFINITY;
    if (bid != RMISS && in_range(bt, tn, INFINITY)) { rid = bid; rt = bt; }
    for (uint32_t j = 0; j < S->ns; ++j) {
        double t;
        if (sph_hit(S, j, o, d, tn, &t) && in_range(t, tn, INFINITY)) {
            uint32_t id = S->nt + j;
            if (t < rt || (t == rt && id < rid)) { rid = id; rt = t; }
        }
    }
    *th = rt;
    return rid;
}

/* Barycentric (u, v) of a triangle hit (the Moller-Trumbore values). */
static void tri_uv(const Scene* S, uint32_t k, V o, V d, double* bu, double* bv) {
    const float* a = S->verts + 3 * (size_t)S->tris[3 * k];
    const float* b = S->verts + 3 * (size_t)S->tris[3 * k + 1];
    const float* c = S->verts + 3 * (size_t)S->tris[3 * k + 2];
    V v0 = f3(a), e1 = sub(f3(b), v0), e2 = sub(f3(c), v0);
    V p = cross(d, e2);
    double det = dot(e1, p);
    V s = sub(o, v0), q = cross(s, e1);
    *bu = det != 0 ? dot(s, p) / det : 0;
    *bv = det != 0 ? dot(d, q) / det : 0;
}
static void onb(V n, V* t, V* b);

/* Surface frame of a hit (IntersectionContext, src/liar/kernel/
 * intersection_context.cpp:41-72,197-223,278-291):
 *   geometric normal ng; shading normal ns = the corner normals interpolated
 *   at (u, v) when the mesh has them (triangle_mesh.cpp:334-358), else ng;
 *   dP/du, dP/dv: triangle edges v1 - v0, v2 - v0 (barycentric parameters,
 *   the mesh having no uvs), sphere (phi / 2pi, theta / pi) derivatives
 *   (sphere.cpp:130-171); both normals flipped towards -d (flipTo); the BSDF
 *   frame u = ns.reject(dP/du), v = ns.reject(dP/dv), or any orthonormal pair
 *   when both are shorter than sqrt(1e-3), then completed by a cross product
 *   with ns from the longer one (bsdfToLocal). */
typedef struct { V x, ns, ng, tt, bb; int leaving, has_ns; } Frame;
static void hit_frame(const Scene* S, uint32_t prim, V o, V d, double tn, double t, Frame* F) {
    V x = add(o, mul(t, d)), ng, ns, dpdu, dpdv;
    if (prim < S->nt) {
        const float* a = S->verts + 3 * (size_t)S->tris[3 * prim];
        const float* b = S->verts + 3 * (size_t)S->tris[3 * prim + 1];
        const float* cc = S->verts + 3 * (size_t)S->tris[3 * prim + 2];
        /* the edges as the library's primitive records hold them (FP32) */
        dpdu = v((float)(b[0] - a[0]), (float)(b[1] - a[1]), (float)(b[2] - a[2]));
        dpdv = v((float)(cc[0] - a[0]), (float)(cc[1] - a[1]), (float)(cc[2] - a[2]));
        ng = nrm(cross(sub(f3(b), f3(a)), sub(f3(cc), f3(a))));
        F->leaving = dot(d, ng) > 0;  /* triangle_mesh.cpp:304-309 */
        ns = ng;
        F->has_ns = 0;
        if (S->tri_normals) {
            const float* n = S->tri_normals + 9 * (size_t)prim;
            V n0 = f3(n), n1 = f3(n + 3), n2 = f3(n + 6);
            if (dot(n0, n0) + dot(n1, n1) + dot(n2, n2) > 0) {
                double bu, bv;
                tri_uv(S, prim, o, d, &bu, &bv);
                ns = nrm(add(add(mul(1 - bu - bv, n0), mul(bu, n1)), mul(bv, n2)));
                F->has_ns = 1;
            }
        }
    } else {
        const float* sp = S->spheres + 4 * (size_t)(prim - S->nt);
        double r = sp[3];
        ng = nrm(sub(x, f3(sp)));
        V s0 = sub(add(o, mul(tn, d)), f3(sp));
        F->leaving = dot(s0, s0) < r * r;  /* sphere.cpp:110-111 */
        ns = ng;
        F->has_ns = 0;
        V N = ng;
        double st = sqrt(fmax(0, 1 - N.z * N.z));
        dpdu = mul(2 * PI * r, v(-N.y, N.x, 0));
        if (st != 0) {
            double cs = N.z / st;
            dpdv = mul(PI * r, v(cs * N.x, cs * N.y, -st));
        } else {
            double phi = atan2(N.y, N.x);
            dpdv = mul(PI * r, v(cos(phi), sin(phi), 0));
        }
    }
    if (dot(ns, d) > 0) ns = mul(-1, ns);
    if (dot(ng, d) > 0) ng = mul(-1, ng);
    V u = sub(dpdu, mul(dot(dpdu, ns), ns)), w = sub(dpdv, mul(dot(dpdv, ns), ns));
    double su = dot(u, u), sv = dot(w, w);
    if (fmax(su, sv) < 1e-3) onb(ns, &u, &w);
    if (su >= sv) w = cross(ns, u);
    else u = cross(w, ns);
    F->x = x;
    F->ns = ns;
    F->ng = ng;
    F->tt = nrm(u);
    F->bb = nrm(w);
}

/* ---- lights ------------------------------------------------------------------ */
static double para_hit(const Scene* S, uint32_t k, V p, V d, double tn) {
    V n = S->light_n[k];
    double dn = dot(d, n);
    if (dn == 0) return -1;
    double t = dot(sub(f3(S->lights[k].origin), p), n) / dn;
    if (!(t > tn)) return -1;
    V q = sub(add(p, mul(t, d)), f3(S->lights[k].origin));
    double u = dot(q, S->light_du[k]), vv = dot(q, S->light_dv[k]);
    if (!(u >= 0 && u <= 1 && vv >= 0 && vv <= 1)) return -1;
    return t;
}
/* The mesh light's own surface along (p, d): TriangleMesh::doIntersect with
 * near = tolerance (triangle_mesh.cpp:291-317), the nearest triangle with
 * t > near kept only if inRange; returns t (or -1) and the unit normal. */
static double mesh_light_hit(const Scene* S, uint32_t k, V p, V d, V* n) {
    const or_light* L = &S->lights[k];
    double best = INFINITY;
    uint32_t bk = RMISS;
    for (uint32_t j = 0; j < L->n_triangles; ++j) {
        double t;
        if (tri_hit(S, L->first_triangle + j, p, d, &t) && t > TOL && t < best) { best = t; bk = j; }
    }
    if (bk == RMISS || !in_range(best, TOL, INFINITY)) return -1;
    uint32_t tri = L->first_triangle + bk;
    const float* a = S->verts + 3 * (size_t)S->tris[3 * tri];
    const float* b = S->verts + 3 * (size_t)S->tris[3 * tri + 1];
    const float* c = S->verts + 3 * (size_t)S->tris[3 * tri + 2];
    *n = nrm(cross(sub(f3(b), f3(a)), sub(f3(c), f3(a))));
    return best;
}
/* LightContext::emission pdf: LightArea::doEmission through
 * SceneObject::doAngularPdf (light_area.cpp:189-203, scene_object.cpp:252-266)
 * re-intersects the surface, whose normal the IntersectionContext has
 * flipped towards p, so the single-sided test never rejects. */
static double light_pdf(const Scene* S, uint32_t k, V p, V d) {
    const or_light* L = &S->lights[k];
    if (L->type == 0) {
        double t = para_hit(S, k, p, d, TOL);
        if (t < 0) return 0;
        return t * t / (S->light_area[k] * fabs(dot(d, S->light_n[k])));
    }
    if (L->type == 3) {
        V n;
        double t = mesh_light_hit(S, k, p, d, &n);
        if (t < 0) return 0;
        return t * t / (S->light_area[k] * fabs(dot(d, n)));
    }
    if (L->type == 1) return 1.0 / (4 * PI);
    return dot(d, S->light_n[k]) >= S->light_cos[k] ? S->light_invsolid[k] : 0.0;
}
/* emission along d: area lights (already hit) emit from both faces (see
 * light_pdf), the sky always, the sun within its disc */
static int light_sees(const Scene* S, uint32_t k, V d) {
    const or_light* L = &S->lights[k];
    if (L->type == 0 || L->type == 1 || L->type == 3) return 1;
    return dot(d, S->light_n[k]) >= S->light_cos[k];
}
static void onb(V n, V* t, V* b) {  /* Duff et al., JCGT 2017 */
    double sign = copysign(1.0, n.z), a = -1.0 / (sign + n.z), bb = n.x * n.y * a;
    *t = v(1 + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    *b = v(bb, sign + n.y * n.y * a, -n.y);
}
static V uniform_sphere(double u1, double u2) {
    double z = 1 - 2 * u1, r = sqrt(fmax(0, 1 - z * z)), phi = 2 * PI * u2;
    return v(r * cos(phi), r * sin(phi), z);
}
/* sampleEmission towards point p (light_area.cpp:207-252); *dist: distance
 * to the sampled point of an area light */
static V light_sample(const Scene* S, uint32_t k, V p, double u1, double u2, double* pdf, double* dist) {
    const or_light* L = &S->lights[k];
    *dist = INFINITY;
    if (L->type == 0 || L->type == 3) {
        V q, n;
        double pdf_area;
        if (L->type == 0) {
            q = add(f3(L->origin), add(mul(u1, f3(L->dx)), mul(u2, f3(L->dy))));
            n = S->light_n[k];
            pdf_area = 1 / S->light_area[k];
        } else {  /* TriangleMesh::doSampleSurface (triangle_mesh.cpp:414-444) */
            const double* cdf = S->light_tcdf[k];
            uint32_t ii = 0;
            while (ii < L->n_triangles && cdf[ii] < u1) ++ii;  /* lower_bound */
            if (ii >= L->n_triangles) ii = L->n_triangles - 1;
            double c0 = ii > 0 ? cdf[ii - 1] : 0, tp = cdf[ii] - c0;
            double di = (u1 - c0) / tp, sq = sqrt(di), a = 1 - sq, b = sq * u2;
            uint32_t tri = L->first_triangle + ii;
            const float* va = S->verts + 3 * (size_t)S->tris[3 * tri];
            const float* vb = S->verts + 3 * (size_t)S->tris[3 * tri + 1];
            const float* vc = S->verts + 3 * (size_t)S->tris[3 * tri + 2];
            V v0 = f3(va), e1 = sub(f3(vb), v0), e2 = sub(f3(vc), v0), nv = cross(e1, e2);
            double nn = sqrt(dot(nv, nv));
            q = add(v0, add(mul(a, e1), mul(b, e2)));
            n = mul(1 / nn, nv);
            pdf_area = tp / (0.5 * nn);
        }
        V to = sub(q, p);
        double d2 = dot(to, to);
        *dist = sqrt(d2);
        to = mul(1 / *dist, to);
        double c = dot(n, to);
        *pdf = pdf_area * d2 / fabs(c);
        if (!L->double_sided && c > 0) *pdf = 0;
        if (!(*pdf < INFINITY)) *pdf = 0;
        return to;
    }
    if (L->type == 1) { *pdf = 1 / (4 * PI); return uniform_sphere(u1, u2); }
    double ct = 1 - u1 * (1 - S->light_cos[k]), st = sqrt(fmax(0, 1 - ct * ct)), phi = 2 * PI * u2;
    V t, b, n = S->light_n[k];
    onb(n, &t, &b);
    *pdf = S->light_invsolid[k];
    return add(add(mul(st * cos(phi), t), mul(st * sin(phi), b)), mul(ct, n));
}
static uint32_t light_pick(const Scene* S, double u) {
    uint32_t k = 0;
    while (k < S->n_lights && S->light_cdf[k] < u) ++k;  /* lower_bound */
    return k < S->n_lights ? k : S->n_lights - 1;
}

/* ---- BSDFs (local frame, z = shading normal towards the viewer) ------------ */
/* alpha_u, alpha_v: CookTorrance max(roughness, 1e-3)^2 along the frame's
 * tangents (cook_torrance.cpp:189-200); fuzz: Mirror (mirror.cpp:142-161). */
typedef struct { uint32_t type; double refl, trans, eta, kappa, ior, alpha_u, alpha_v, fuzz; } Bsdf;

static double fresnel_cond(double eta, double kappa, double ci) {  /* conductor.cpp:184-194 */
    double c2 = ci * ci, a = eta * eta + kappa * kappa, b = 2 * ci * eta;
    double ao = a + c2, ap = a * c2 + 1;
    return ((ao - b) / (ao + b) + (ap - b) / (ap + b)) / 2;
}
/* MicrofacetDistribution::samplePhi (src/liar/kernel/microfacet.cpp:102-123) */
static double sample_phi(double u, double au, double av) {
    if (au == av) return 2 * PI * u;
    double phi = atan(av / au * tan(2 * PI * u));
    if (u <= 0.25) return phi;
    if (u < 0.75) return phi + PI;
    return phi + 2 * PI;
}
/* D(h) and its pdf D cos(theta_h): Trowbridge-Reitz
 * (microfacet_trowbridge_reitz.cpp:46-58), Beckmann (microfacet_beckmann.cpp:50-72). */
static double mdf_d(uint32_t type, double au, double av, V h, double* pdf_h) {
    double d;
    if (type == 5) {
        double q = (h.x / au) * (h.x / au) + (h.y / av) * (h.y / av) + h.z * h.z;
        d = 1 / (PI * au * av * q * q);
    } else {
        double c2 = h.z * h.z;
        if (c2 == 0) { *pdf_h = 0; return 0; }
        d = exp(-((h.x / au) * (h.x / au) + (h.y / av) * (h.y / av)) / c2) / (PI * au * av * c2 * c2);
    }
    if (g_plant == 5) d *= PI;
    *pdf_h = d * h.z;
    return d;
}
/* sampleH: Trowbridge-Reitz (microfacet_trowbridge_reitz.cpp:62-72), Beckmann
 * (microfacet_beckmann.cpp:76-106; its anisotropic branch takes phi from
 * sample.x, the same number as the radial part, as the reference does). */
static V mdf_sample_h(uint32_t type, double au, double av, double u1, double u2) {
    if (type == 5) {
        double phi = sample_phi(u1, au, av), s = sqrt(u2 / (1 - u2));
        return nrm(v(s * au * cos(phi), s * av * sin(phi), 1));
    }
    double s = log(1 - u1), cp, sp, t2;
    if (au == av) {
        double phi = 2 * PI * u2;
        cp = cos(phi); sp = sin(phi);
        t2 = -s * au * au;
    } else {
        double phi = sample_phi(u1, au, av);
        cp = cos(phi); sp = sin(phi);
        t2 = -s / ((cp / au) * (cp / au) + (sp / av) * (sp / av));
    }
    double c2 = 1 / (1 + t2), ct = sqrt(c2), st = sqrt(fmax(1 - c2, 0));
    return v(cp * st, sp * st, ct);
}
/* BSDF value and pdf of wo, caps already filtered by the caller's geometric
 * hemisphere test (Bsdf::evaluate, bsdf.cpp:52-76): Lambert (lambert.cpp:105-114)
 * and CookTorrance (cook_torrance.cpp:262-280); delta BSDFs evaluate to 0. */
static int bsdf_eval(const Bsdf* B, V wi, V wo, double* val, double* pdf) {
    if (wo.z == 0) return 0;
    if (B->type == 1) { *val = B->refl / PI; *pdf = fabs(wo.z) / PI; return *pdf > 0; }
    if (B->type == 5 || B->type == 6) {
        if (!(wo.z > 0)) return 0;
        V h = nrm(add(wi, wo));
        double pdfh, d = mdf_d(B->type, B->alpha_u, B->alpha_v, h, &pdfh);
        double doh = fabs(dot(wi, h));
        double g = fmin(1, fmin(2 * h.z * wi.z / doh, 2 * h.z * wo.z / doh));  /* cook_torrance.cpp:234-238 */
        double woh = dot(wi, h);
        *pdf = g_plant == 3 ? pdfh : pdfh / (4 * woh);
        *val = B->refl * d * g / (4 * wi.z * wo.z) * fresnel_cond(B->eta, B->kappa, fabs(woh));
        return *pdf > 0;
    }
    return 0;
}
/* Bsdf::evaluate (bsdf.cpp:52-76): wo on the other side of the geometric
 * normal (ngl, in the local frame, towards the viewer) than wi masks the
 * reflection caps; every BSDF here evaluates only reflection (the delta
 * ones evaluate to 0 anyway), so such a wo has no value. */
static int bsdf_evaluate(const Bsdf* B, V wi, V wo, V ngl, double* val, double* pdf) {
    if (wo.z == 0) return 0;
    if ((dot(wi, ngl) > 0) != (dot(wo, ngl) > 0)) return 0;
    return bsdf_eval(B, wi, wo, val, pdf);
}
/* allow: bit 0 reflection, bit 1 transmission (the caps the caller permits;
 * only the dielectric has both and chooses among the permitted ones). */
static int bsdf_sample(const Bsdf* B, V wi, double u1, double u2, double uc, V* wo, double* val, double* pdf,
                       unsigned allow) {
    switch (B->type) {
    case 1: {  /* concentric disk + Malley */
        double a = 2 * u1 - 1, b = 2 * u2 - 1, r, phi;
        if (a == 0 && b == 0) { r = 0; phi = 0; }
        else if (a * a > b * b) { r = a; phi = (PI / 4) * (b / a); }
        else { r = b; phi = PI / 2 - (PI / 4) * (a / b); }
        double x = r * cos(phi), y = r * sin(phi), z = sqrt(fmax(0, 1 - x * x - y * y));
        *wo = v(x, y, z);
        *pdf = z / PI;
        *val = B->refl / PI;
        return *pdf > 0;
    }
    case 2:  /* mirror.cpp:142-161: value = reflectance, pdf = cos; fuzz adds
                fuzz x a uniform-sphere point to the reflection, unnormalised */
        *wo = v(-wi.x, -wi.y, wi.z);
        if (B->fuzz > 0) {
            double z = 1 - 2 * u1, r = sqrt(fmax(0, 1 - z * z)), phi = 2 * PI * u2;
            *wo = add(*wo, mul(B->fuzz, v(r * cos(phi), r * sin(phi), z)));
            *val = wo->z > 0 ? B->refl : 0;
        } else {
            *val = B->refl;
        }
        *pdf = wo->z;
        return *pdf > 0;
    case 4:
        *wo = v(-wi.x, -wi.y, wi.z);
        *val = B->refl * fresnel_cond(B->eta, B->kappa, wi.z) / wi.z;
        *pdf = 1;
        return 1;
    case 3: {  /* dielectric.cpp:209-247 */
        double ci = wi.z, ior = B->ior;
        if (ci <= 0) return 0;
        double st2 = ior * ior * (1 - ci * ci), ct = sqrt(fmax(1 - st2, 0));
        double ro = (ior * ci - ct) / (ior * ci + ct), rp = (ci - ior * ct) / (ci + ior * ct);
        double fr = (ro * ro + rp * rp) / 2;
        double pr = (allow & 1u) ? fabs(B->refl) * fr : 0, pt = (allow & 2u) ? fabs(B->trans) * (1 - fr) : 0;
        if (pr + pt == 0) return 0;
        double prob = pr / (pr + pt);
        if (uc < prob) { *wo = v(-wi.x, -wi.y, wi.z); *val = B->refl * fr / ci; *pdf = prob; }
        else { *wo = v(-ior * wi.x, -ior * wi.y, -ct); *val = B->trans * (1 - fr) / ct; *pdf = 1 - prob; }
        return *pdf > 0;
    }
    case 5:
    case 6: {  /* cook_torrance.cpp:284-294: sample h, reflect, evaluate */
        V h = mdf_sample_h(B->type, B->alpha_u, B->alpha_v, u1, u2);
        *wo = sub(mul(2 * dot(wi, h), h), wi);
        if (!(wo->z > 0)) return 0;
        return bsdf_eval(B, wi, *wo, val, pdf);
    }
    }
    return 0;
}

/* ---- probes (tests only): one BSDF / distribution / spectrum evaluation ---- */
/* prm: type, refl, trans, eta, kappa, ior, alpha_u, alpha_v, fuzz */
static Bsdf probe_bsdf(const double* prm) {
    Bsdf B = {(uint32_t)prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], prm[7], prm[8]};
    return B;
}
/* out: value, pdf; returns 1 when the BSDF has a value for (wi, wo) */
int or_bsdf_eval(const double* prm, const double* wi, const double* wo, double* out) {
    Bsdf B = probe_bsdf(prm);
    out[0] = out[1] = 0;
    return bsdf_eval(&B, v(wi[0], wi[1], wi[2]), v(wo[0], wo[1], wo[2]), &out[0], &out[1]);
}
/* u: u1, u2, component sample; out: wo (3), value, pdf */
int or_bsdf_sample(const double* prm, const double* wi, const double* u, uint32_t allow, double* out) {
    Bsdf B = probe_bsdf(prm);
    V wo = v(0, 0, 0);
    double val = 0, pdf = 0;
    int ok = bsdf_sample(&B, v(wi[0], wi[1], wi[2]), u[0], u[1], u[2], &wo, &val, &pdf, allow);
    out[0] = wo.x; out[1] = wo.y; out[2] = wo.z; out[3] = val; out[4] = pdf;
    return ok;
}
double or_fresnel_conductor(double eta, double kappa, double cos_i) { return fresnel_cond(eta, kappa, cos_i); }
/* Batched probes: n inputs of 3 doubles (wi, wo or u); out 2 (value, pdf)
 * or 5 (wo, value, pdf) doubles per input; ok[i] = the call's result. */
void or_bsdf_eval_n(const double* prm, const double* wi, const double* wo, uint32_t n, double* out, uint8_t* ok) {
    Bsdf B = probe_bsdf(prm);
    for (uint32_t i = 0; i < n; ++i) {
        out[2 * i] = out[2 * i + 1] = 0;
        ok[i] = (uint8_t)bsdf_eval(&B, v(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]),
                                   v(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), &out[2 * i], &out[2 * i + 1]);
    }
}
void or_bsdf_sample_n(const double* prm, const double* wi, const double* u, uint32_t n, uint32_t allow, double* out,
                      uint8_t* ok) {
    Bsdf B = probe_bsdf(prm);
    for (uint32_t i = 0; i < n; ++i) {
        V wo = v(0, 0, 0);
        double val = 0, pdf = 0;
        ok[i] = (uint8_t)bsdf_sample(&B, v(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), u[3 * i], u[3 * i + 1],
                                     u[3 * i + 2], &wo, &val, &pdf, allow);
        double* o = out + 5 * (size_t)i;
        o[0] = wo.x; o[1] = wo.y; o[2] = wo.z; o[3] = val; o[4] = pdf;
    }
}
/* One spectrum at n wavelengths [nm], as the renderer evaluates it. */
void or_spectrum_eval(const or_spectrum* sp, const double* w, uint32_t n, int reflectant, double* out) {
    Scene S;
    memset(&S, 0, sizeof S);
    S.spectra = sp;
    S.n_spectra = 1;
    double mx = -INFINITY, scale = 1;
    if (sp->kind == 1 || sp->kind == 2)
        for (uint32_t k = 0; k < sp->count; ++k) mx = fmax(mx, sp->v[k]);
    if (sp->kind == 1 && sp->value > 0) mx = sp->value;
    if (mx > 1) scale = 1 / mx;
    S.spec_reflscale = &scale;
    for (uint32_t i = 0; i < n; ++i) out[i] = spec_eval(&S, 0, w[i], reflectant);
}
double or_mdf_d(uint32_t type, double au, double av, const double* h, double* pdf_h) {
    return mdf_d(type, au, av, v(h[0], h[1], h[2]), pdf_h);
}
void or_mdf_sample_h(uint32_t type, double au, double av, double u1, double u2, double* h) {
    V r = mdf_sample_h(type, au, av, u1, u2);
    h[0] = r.x; h[1] = r.y; h[2] = r.z;
}

/* ---- one path -------------------------------------------------------------- */
typedef struct { uint32_t max_gen, secondary, glossy; } DLCfg;
typedef struct { const Scene* S; const or_params* P; uint32_t k0, k1; const DLCfg* dl; } Ctx;

static void rng(const Ctx* c, uint64_t path, uint32_t gen, uint32_t purpose, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)path, (uint32_t)(path >> 32), gen, purpose};
    philox(c->k0, c->k1, ctr, out);
}

/* Radiance L(lambda) of path `p` (tracePhoton, iteratively: the estimator
 * is a product of per-vertex weights plus per-vertex emission). */
static double trace_path(const Ctx* c, uint64_t p, V o, V d, double tn, double lam, uint32_t max_gen,
                         double thr) {
    const Scene* S = c->S;
    const or_params* P = c->P;
    double L = 0;
    for (uint32_t gen = 0;; ++gen) {
        double t;
        uint32_t prim = intersect(S, o, d, tn, &t);
        if (prim == RMISS) {
            if (gen > 0)
                for (uint32_t k = 0; k < S->n_lights; ++k)
                    if ((S->lights[k].type == 1 || S->lights[k].type == 2) && light_sees(S, k, d))
                        L += thr * spec_eval(S, S->lights[k].radiance, lam, 0);
            return L;
        }
        Frame F;
        hit_frame(S, prim, o, d, tn, t, &F);
        const V x = F.x, nf = F.ns, tt = F.tt, bb = F.bb;
        const int leaving = F.leaving;
        const V ngl = F.has_ns ? v(dot(F.ng, tt), dot(F.ng, bb), dot(F.ng, nf)) : v(0, 0, 1);
        const or_material* m = &S->mats[S->prim_mat[prim]];
        if (gen > 0 && m->light != RNONE && light_sees(S, m->light, d))
            L += thr * spec_eval(S, S->lights[m->light].radiance, lam, 0);
        if ((gen == 0 || m->light == RNONE) && m->emission != RNONE) L += thr * spec_eval(S, m->emission, lam, 0);
        if (m->type == 0) return L;
        Bsdf B;
        B.type = m->type;
        B.refl = spec_eval(S, m->refl, lam, 1);
        B.trans = spec_eval(S, m->trans, lam, 1);
        B.eta = fmax(spec_eval(S, m->eta, lam, 0), 1e-9);
        B.kappa = fmax(spec_eval(S, m->kappa, lam, 0), 1e-9);
        B.ior = leaving ? B.eta / B.kappa : B.kappa / B.eta;
        double ru = fmax(m->roughness, 1e-3), rv = fmax(m->roughness_v, 1e-3);
        B.alpha_u = ru * ru;
        B.alpha_v = rv * rv;
        B.fuzz = m->fuzz;
        V wiw = mul(-1, d);
        V wi = v(dot(wiw, tt), dot(wiw, bb), fmax(dot(wiw, nf), 1.1920929e-7));
        uint32_t r0[4], r1[4];
        rng(c, p, gen, 1u, r0);
        rng(c, p, gen, 2u, r1);
        double us = u01(r0[0]), u1 = u01(r0[1]), u2 = u01(r0[2]), uc = u01(r0[3]);
        double ul = u01(r1[0]), ul1 = u01(r1[1]), ul2 = u01(r1[2]), urr = u01(r1[3]);
        if (P->sampler == 1 && gen < P->sampler_generations) {
            /* the AdjointPhotonTracer's sub-sequences of this generation
             * (adjoint_photon_tracer.cpp:47-67): 1D ids 4g .. 4g+3 on the
             * first primes, 2D ids 3g .. 3g+2 after all 1D ones */
            uint64_t j = p + 1;  /* p: the path's position in the sequence */
            uint32_t b2 = 4 * P->sampler_generations + 6 * gen;
            us = halton_sub(j, 4 * gen);
            uc = halton_sub(j, 4 * gen + 1);
            ul = halton_sub(j, 4 * gen + 2);
            u1 = halton_sub(j, b2);
            u2 = halton_sub(j, g_plant == 6 ? b2 : b2 + 1);
            ul1 = halton_sub(j, b2 + 2);
            ul2 = halton_sub(j, b2 + 3);
        }
        double strategy = (m->type == 1 || m->type == 5 || m->type == 6) ? 0.5 : 1.0;
        V wo = v(0, 0, 1);
        double val = 0, pdf = 0;
        int ok = 0;
        if (us < strategy) {
            ok = bsdf_sample(&B, wi, u1, u2, uc, &wo, &val, &pdf, 3u);
            if (ok) {
                pdf *= strategy;
                double pl = 0;
                if (strategy < 1) {
                    V dw = add(add(mul(wo.x, tt), mul(wo.y, bb)), mul(wo.z, nf));
                    for (uint32_t k = 0; k < S->n_lights; ++k) pl += S->light_sel[k] * light_pdf(S, k, x, dw);
                    pl *= 1 - strategy;
                }
                if (g_plant != 1) val *= pdf * pdf / (pdf * pdf + pl * pl);
            }
        } else if (S->n_lights > 0) {
            uint32_t k = light_pick(S, ul);
            double pdfl;
            double distl;
            V dw = light_sample(S, k, x, ul1, ul2, &pdfl, &distl);
            if (S->light_sel[k] > 0 && pdfl > 0) {
                double pl = S->light_sel[k] * pdfl;
                for (uint32_t j = 0; j < S->n_lights; ++j)
                    if (j != k) pl += S->light_sel[j] * light_pdf(S, j, x, dw);
                pl *= 1 - strategy;
                wo = v(dot(dw, tt), dot(dw, bb), dot(dw, nf));
                double vb, pb;
                if (bsdf_evaluate(&B, wi, wo, ngl, &vb, &pb) && vb != 0) {
                    double pbs = pb * strategy;
                    val = g_plant == 1 ? vb : vb * pl * pl / (pl * pl + pbs * pbs);
                    pdf = pl;
                    ok = 1;
                }
            }
        }
        if (!(ok && pdf > 0)) return L;
        double ct = wo.z;
        thr *= val * fabs(ct) / pdf;
        V start = add(x, mul((ct > 0 ? 10.0 : -10.0) * TOL, F.ng));  /* adjoint_photon_tracer.cpp:177 */
        d = nrm(add(add(mul(wo.x, tt), mul(wo.y, bb)), mul(wo.z, nf)));
        o = start;
        tn = TOL;
        uint32_t next = gen + 1;
        if (next >= max_gen) return L;
        if (next >= P->rr_depth) {
            if (urr < P->rr_survival) thr /= P->rr_survival;
            else return L;
        }
        if (!(thr > 0) || !isfinite(thr)) return L;
    }
}


/* ---- DirectLighting (src/liar/tracers/direct_lighting.cpp:119-377) ---------- */

/* Scene::isIntersecting over the List{TriangleMesh, Spheres}: any triangle
 * with tn < t < tf (triangle_mesh.cpp:321-330), any sphere whose nearest
 * root beyond tn is in range (sphere.cpp:104-127). */
static int occluded(const Scene* S, V o, V d, double tn, double tf) {
    for (uint32_t k = 0; k < S->nt; ++k) {
        double t;
        if (tri_hit(S, k, o, d, &t) && t > tn && t < tf) return 1;
    }
    for (uint32_t j = 0; j < S->ns; ++j) {
        double t;
        if (sph_hit(S, j, o, d, tn, &t) && in_range(t, tn, tf)) return 1;
    }
    return 0;
}

static double sq_heuristic(double a, double b) { return a * a / (a * a + b * b); }  /* ray_tracer.cpp:180-183 */

static int is_specular(uint32_t type) { return type == 2 || type == 3 || type == 4; }

typedef struct {
    Bsdf B;
    V wi, tt, bb, nf, ngl, start;
    double lam;
    uint64_t p;
    uint32_t key;  /* generation | branch << 8: the node's random-number key */
    int hq;
} DLHit;

/* Candidate j of a hit's light estimate (RayTracer::estimateLightContribution,
 * src/liar/kernel/ray_tracer.cpp:188-268, split into one light or BSDF
 * sample per candidate): high quality takes light k's 2 n_k candidates after
 * those of lights 0..k-1, even = light sample, odd = BSDF sample; otherwise
 * candidates 2s, 2s+1 belong to secondary sample s whose light is picked by
 * power and whose estimate is divided by (samples x pick pdf)
 * (direct_lighting.cpp:259-310).  Shadow rays to a parallelogram end 1e-4
 * (relative) short of its plane. */
static uint32_t dl_pair_fifth(const uint32_t* r) {
    return ((r[0] & 0xffu) | ((r[1] & 0xffu) << 8) | ((r[2] & 0xffu) << 16)) << 8;
}

/* The two candidates of a pair (2s, 2s+1) share the light pick (purpose
 * 10 + s) and one 128-bit block (purpose 100 + s): the light sample takes
 * words 0, 1, the BSDF sample words 2, 3 and the low bytes of words 0-2 as a
 * fifth 24-bit uniform.  The even candidate draws into rs / rp, the odd one
 * reads them. */
static double dl_candidate(const Ctx* c, const DLHit* H, uint32_t j, uint32_t* rs, uint32_t* rp) {
    const Scene* S = c->S;
    uint32_t k = 0, local = 0;
    double nl = 1, scale = 1;
    if (H->hq) {
        uint32_t base = 0;
        int found = 0;
        for (uint32_t m = 0; m < S->n_lights && !found; ++m) {
            uint32_t n2 = 2 * S->light_ns[m];
            if (j < base + n2) { k = m; local = j - base; found = 1; }
            base += n2;
        }
        if (!found) return 0;
        nl = S->light_ns[k];
    } else {
        if (j >= 2 * c->dl->secondary || S->n_lights == 0) return 0;
        if ((j & 1u) == 0) rng(c, H->p, H->key, 10u + (j >> 1), rs);
        k = light_pick(S, u01(rs[0]));
        double psel = S->light_sel[k];
        if (!(psel > 0)) return 0;
        local = j & 1u;
        scale = c->dl->secondary * psel;
    }
    const or_light* L = &S->lights[k];
    double Le = spec_eval(S, L->radiance, H->lam, 0);
    if (Le == 0) return 0;
    if ((j & 1u) == 0) rng(c, H->p, H->key, 100u + (j >> 1), rp);
    double f = 0, pb = 0, pl = 0, contrib, dist = INFINITY;
    V dir;
    if ((local & 1u) == 0) {
        dir = light_sample(S, k, H->start, u01(rp[0]), u01(rp[1]), &pl, &dist);
        if (!(pl > 0)) return 0;
        /* sampleEmission with the target normal (light_area.cpp:242-252) */
        if (!(dot(dir, H->nf) > 0)) return 0;
        V wo = v(dot(dir, H->tt), dot(dir, H->bb), dot(dir, H->nf));
        if (!bsdf_evaluate(&H->B, H->wi, wo, H->ngl, &f, &pb) || f == 0) return 0;
        double wt = g_plant == 2 ? 1 : sq_heuristic(nl * pl, nl * pb);
        contrib = f * Le * wt * fabs(wo.z) / (nl * pl * scale);
    } else {
        V wo;
        if (!bsdf_sample(&H->B, H->wi, u01(rp[2]), u01(rp[3]), u01(dl_pair_fifth(rp)), &wo, &f, &pb, 3u) ||
            !(pb > 0) || f == 0)
            return 0;
        dir = nrm(add(add(mul(wo.x, H->tt), mul(wo.y, H->bb)), mul(wo.z, H->nf)));
        pl = light_pdf(S, k, H->start, dir);
        if (!(pl > 0)) return 0;
        if (L->type == 3) {
            V nn;
            dist = mesh_light_hit(S, k, H->start, dir, &nn);
        }
        double wt = (is_specular(H->B.type) || g_plant == 2) ? 1 : sq_heuristic(nl * pb, nl * pl);
        contrib = f * Le * wt * fabs(wo.z) / (nl * pb * scale);
    }
    double far = INFINITY;
    if (L->type == 0) {
        V n = S->light_n[k];
        double t = dot(sub(f3(L->origin), H->start), n) / dot(dir, n);
        if (!(t > 0)) return 0;
        far = t * (1 - 1e-4);
    } else if (L->type == 3) {
        if (!(dist > 0) || !isfinite(dist)) return 0;
        far = dist * (1 - 1e-4);
    }
    if (!(contrib != 0) || !isfinite(contrib)) return 0;
    return occluded(S, H->start, dir, TOL, far) ? 0 : contrib;
}

/* DirectLighting::castRay / doShadeSurface, recursively: radiance along
 * (o, d) for the tree node (gen, br) of path p. */
static double dl_cast(const Ctx* c, uint64_t p, V o, V d, double tn, double lam, uint32_t gen, uint32_t br,
                      int hq) {
    const Scene* S = c->S;
    const DLCfg* A = c->dl;
    double t;
    uint32_t prim = intersect(S, o, d, tn, &t);
    if (prim == RMISS) return 0;  /* alpha = 0: escaped rays add nothing */
    Frame F;
    hit_frame(S, prim, o, d, tn, t, &F);
    const V x = F.x, nf = F.ns;
    const int leaving = F.leaving;
    const or_material* m = &S->mats[S->prim_mat[prim]];
    double L = 0;
    if ((gen == 0 || m->light == RNONE) && m->emission != RNONE) L += spec_eval(S, m->emission, lam, 0);
    if (m->type == 0) return L;
    DLHit H;
    H.B.type = m->type;
    H.B.refl = spec_eval(S, m->refl, lam, 1);
    H.B.trans = spec_eval(S, m->trans, lam, 1);
    H.B.eta = fmax(spec_eval(S, m->eta, lam, 0), 1e-9);
    H.B.kappa = fmax(spec_eval(S, m->kappa, lam, 0), 1e-9);
    H.B.ior = leaving ? H.B.eta / H.B.kappa : H.B.kappa / H.B.eta;
    double rgu = fmax(m->roughness, 1e-3), rgv = fmax(m->roughness_v, 1e-3);
    H.B.alpha_u = rgu * rgu;
    H.B.alpha_v = rgv * rgv;
    H.B.fuzz = m->fuzz;
    H.nf = nf;
    H.tt = F.tt;
    H.bb = F.bb;
    H.ngl = F.has_ns ? v(dot(F.ng, F.tt), dot(F.ng, F.bb), dot(F.ng, nf)) : v(0, 0, 1);
    V wiw = mul(-1, d);
    H.wi = v(dot(wiw, H.tt), dot(wiw, H.bb), fmax(dot(wiw, nf), 1.1920929e-7));
    H.start = add(x, mul(10 * TOL, nf));
    H.lam = lam;
    H.p = p;
    H.key = gen | (br << 8);
    H.hq = hq;
    uint32_t slots = 0;
    if (S->n_lights) {
        for (uint32_t k = 0; k < S->n_lights; ++k) slots += 2 * S->light_ns[k];
        if (2 * A->secondary > slots) slots = 2 * A->secondary;
    }
    uint32_t rs[4] = {0, 0, 0, 0}, rp[4] = {0, 0, 0, 0};
    for (uint32_t j = 0; j < slots; ++j) L += dl_candidate(c, &H, j, rs, rp);
    /* specular / glossy continuations (direct_lighting.cpp:314-377) */
    int glossy = m->type == 5 || m->type == 6;
    if (gen < A->max_gen && (is_specular(m->type) || (glossy && A->glossy))) {
        for (uint32_t side = 0; side < 2; ++side) {
            if (side == 1 && m->type != 3) break;
            uint32_t r[4];
            rng(c, p, H.key, 5u + side, r);
            V wo;
            double f = 0, pdf = 0;
            if (!bsdf_sample(&H.B, H.wi, u01(r[0]), u01(r[1]), u01(r[2]), &wo, &f, &pdf, side ? 2u : 1u)) continue;
            if (!(pdf > 0) || f == 0 || (side ? !(wo.z < 0) : !(wo.z > 0))) continue;
            double w = f * fabs(wo.z) / pdf;
            if (!(w > 0) || !isfinite(w)) continue;
            V dir = nrm(add(add(mul(wo.x, H.tt), mul(wo.y, H.bb)), mul(wo.z, nf)));
            V org = add(x, mul((side ? -10.0 : 10.0) * TOL, nf));
            L += w * dl_cast(c, p, org, dir, TOL, lam, gen + 1, (br << 1) | side, hq && is_specular(m->type));
        }
    }
    return L;
}

typedef struct {
    const Ctx* c;
    V pos, dir, right, down, dirn;
    double near_limit;
    uint32_t strata;
    double* xyz;     /* 3 planes, W*H each, sums */
    double* weight;
    uint32_t row0, row1;
    uint64_t s0, s1; /* Halton: this thread's samples (into its own xyz / weight) */
} Job;

/* Halton samples [s0, s1) of the call: screen point in bases 2 and 3 of the
 * whole film, wavelength in base 13 (halton.cpp:191-212), accumulated into
 * the pixel the point falls in (raster.cpp:547-606). */
static void* run_halton(void* arg) {
    Job* J = (Job*)arg;
    const Ctx* c = J->c;
    const or_params* P = c->P;
    uint32_t max_gen = P->max_depth ? P->max_depth : 64u;
    uint64_t np = (uint64_t)P->width * P->height;
    for (uint64_t p = P->sample_offset + J->s0; p < P->sample_offset + J->s1; ++p) {
        uint64_t j = p + 1;  /* p keys the path's random streams too */
        float Xs = radical_inverse(j, 2, NULL), Ys = radical_inverse(j, 3, NULL);
        float uw = radical_inverse(j, 13, NULL);
        uint32_t px = (uint32_t)(Xs * (float)P->width), py = (uint32_t)(Ys * (float)P->height);
        if (px >= P->width) px = P->width - 1;
        if (py >= P->height) py = P->height - 1;
        uint64_t pixel = (uint64_t)py * P->width + px;
        J->weight[pixel] += 1;
        V d = nrm(add(J->dir, add(mul((double)Xs - 0.5, J->right), mul((double)Ys - 0.5, J->down))));
        double rf = 1.0 / dot(d, J->dirn);
        double tn = fmax(TOL, rf * J->near_limit);
        uint32_t rp[4];
        rng(c, p, 0xFFFEu, 0u, rp);
        double lpdf;
        double lam = obs_sample(c->S, uw, &lpdf);
        if (!(lpdf > 0) || max_gen == 0) continue;
        double thr = 1;
        if (P->rr_depth == 0) {
            if (u01(rp[1]) < P->rr_survival) thr /= P->rr_survival;
            else continue;
        }
        double L = trace_path(c, p, J->pos, d, tn, lam, max_gen, thr);
        if (L == 0) continue;
        double cmf[3];
        obs_xyz(c->S, lam, cmf);
        double sc = L / lpdf;
        if (P->max_sample_luminance > 0 && sc * cmf[1] > P->max_sample_luminance)
            sc *= P->max_sample_luminance / (sc * cmf[1]);
        J->xyz[pixel] += sc * cmf[0];
        J->xyz[np + pixel] += sc * cmf[1];
        J->xyz[2 * np + pixel] += sc * cmf[2];
    }
    return NULL;
}

static void* run_rows(void* arg) {
    Job* J = (Job*)arg;
    const Ctx* c = J->c;
    const or_params* P = c->P;
    uint32_t max_gen = P->max_depth ? P->max_depth : 64u;
    uint64_t np = (uint64_t)P->width * P->height;
    for (uint32_t py = J->row0; py < J->row1; ++py)
        for (uint32_t px = 0; px < P->width; ++px) {
            uint64_t pixel = (uint64_t)py * P->width + px;
            double X = 0, Y = 0, Z = 0;
            for (uint32_t s = 0; s < P->spp; ++s) {
                uint64_t sidx = pixel * P->spp + s;
                uint32_t rs[4];
                rng(c, sidx, 0xFFFFu, 0u, rs);
                double sx = u01(rs[0]), sy = u01(rs[1]);
                if (J->strata) {
                    sx = ((double)(s % J->strata) + sx) / J->strata;
                    sy = ((double)(s / J->strata) + sy) / J->strata;
                }
                double Xs = (px + sx) / P->width, Ys = (py + sy) / P->height;
                V d = nrm(add(J->dir, add(mul(Xs - 0.5, J->right), mul(Ys - 0.5, J->down))));
                double rf = 1.0 / dot(d, J->dirn);
                double tn = fmax(TOL, rf * J->near_limit);
                for (uint32_t l = 0; l < P->wavelengths; ++l) {
                    uint64_t p = sidx * P->wavelengths + l;
                    uint32_t rp[4];
                    rng(c, p, 0xFFFEu, 0u, rp);
                    double uw = (l + u01(rp[0])) / P->wavelengths, lpdf;
                    double lam = obs_sample(c->S, uw, &lpdf);
                    if (!(lpdf > 0) || max_gen == 0) continue;
                    double thr = 1;
                    if (!c->dl && P->rr_depth == 0) {
                        if (u01(rp[1]) < P->rr_survival) thr /= P->rr_survival;
                        else continue;
                    }
                    double L = c->dl ? dl_cast(c, p, J->pos, d, tn, lam, 0, 0, 1)
                                     : trace_path(c, p, J->pos, d, tn, lam, max_gen, thr);
                    if (L == 0) continue;
                    double cmf[3];
                    obs_xyz(c->S, lam, cmf);
                    double sc = L / lpdf;
                    if (P->max_sample_luminance > 0 && sc * cmf[1] > P->max_sample_luminance)
                        sc *= P->max_sample_luminance / (sc * cmf[1]);  /* raster.cpp:567-570 */
                    X += sc * cmf[0];
                    Y += sc * cmf[1];
                    Z += sc * cmf[2];
                }
            }
            J->xyz[pixel] += X;
            J->xyz[np + pixel] += Y;
            J->xyz[2 * np + pixel] += Z;
            J->weight[pixel] += (double)P->spp * P->wavelengths;
        }
    return NULL;
}

static double lum(const Scene* S, uint32_t i, const float* w, const float* xyz, int n) {
    double num = 0, den = 0;
    for (int k = 0; k < n; ++k) {
        double wp = w[k > 0 ? k - 1 : 0], wn = w[k + 1 < n ? k + 1 : n - 1];
        num += spec_eval(S, i, w[k], 0) * xyz[3 * k + 1] * (wn - wp) / 2;
        den += xyz[3 * k + 1] * (wn - wp) / 2;
    }
    return den > 0 ? num / den : 0;
}

static int render_impl(const float* verts, const uint32_t* tris, uint32_t nt, const float* spheres, uint32_t ns,
                       const float* tri_normals, const uint32_t* prim_mat, const or_material* mats, const or_spectrum* spectra,
                       uint32_t n_spectra, const or_light* lights, uint32_t n_lights, const float* obs_w,
                       const float* obs_xyz, uint32_t obs_n, const or_camera* cam, const or_params* P, double* xyz,
                       double* weight, int nthreads, const DLCfg* dl) {
    Scene S;
    memset(&S, 0, sizeof S);
    S.verts = verts; S.tris = tris; S.nt = nt; S.tri_normals = tri_normals; S.spheres = spheres; S.ns = ns; S.prim_mat = prim_mat;
    S.mats = mats; S.spectra = spectra; S.n_spectra = n_spectra; S.lights = lights; S.n_lights = n_lights;
    if (build_observer(&S, obs_w, obs_xyz, (int)obs_n, P->lambda_min, P->lambda_max)) return -1;
    S.spec_reflscale = (double*)malloc(sizeof(double) * (n_spectra + 1));
    for (uint32_t i = 0; i < n_spectra; ++i) {
        double mx = -INFINITY;
        if (spectra[i].kind == 1 || spectra[i].kind == 2)
            for (uint32_t k = 0; k < spectra[i].count; ++k) mx = fmax(mx, spectra[i].v[k]);
        if (spectra[i].kind == 1 && spectra[i].value > 0) mx = spectra[i].value;  /* caller's divisor */
        S.spec_reflscale[i] = mx > 1 ? 1 / mx : 1;
    }
    /* scene bounding sphere radius (half the AABB diagonal) */
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = 0; k < 3 * nt; ++k) {
        const float* p = verts + 3 * (size_t)tris[k];
        for (int a = 0; a < 3; ++a) { lo[a] = fmin(lo[a], p[a]); hi[a] = fmax(hi[a], p[a]); }
    }
    for (uint32_t j = 0; j < ns; ++j)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], spheres[4 * j + a] - spheres[4 * j + 3]);
            hi[a] = fmax(hi[a], spheres[4 * j + a] + spheres[4 * j + 3]);
        }
    double R = 0;
    if (nt + ns)
        for (int a = 0; a < 3; ++a) R += 0.25 * (hi[a] - lo[a]) * (hi[a] - lo[a]);
    R = sqrt(R);
    double far_away = R > 0 ? 1e3 * R : 1e3;
    uint32_t nl = n_lights ? n_lights : 1;
    S.light_sel = calloc(nl, sizeof(double)); S.light_cdf = calloc(nl, sizeof(double));
    S.light_area = calloc(nl, sizeof(double)); S.light_n = calloc(nl, sizeof(V));
    S.light_du = calloc(nl, sizeof(V)); S.light_dv = calloc(nl, sizeof(V));
    S.light_cos = calloc(nl, sizeof(double)); S.light_invsolid = calloc(nl, sizeof(double));
    S.light_ns = calloc(nl, sizeof(uint32_t));
    S.light_tcdf = calloc(nl, sizeof(double*));
    for (uint32_t k = 0; k < n_lights; ++k)
        S.light_ns[k] = lights[k].samples ? lights[k].samples
                                          : ((lights[k].type == 0 || lights[k].type == 3) ? 9u : 1u);
    double total = 0;
    double* power = calloc(nl, sizeof(double));
    for (uint32_t k = 0; k < n_lights; ++k) {
        const or_light* L = &lights[k];
        double Y = lum(&S, L->radiance, obs_w, obs_xyz, (int)obs_n);
        if (L->type == 0) {
            V a = f3(L->dx), b = f3(L->dy), cr = cross(a, b);
            double area = sqrt(dot(cr, cr));
            V n = mul(1 / area, cr);
            S.light_area[k] = area;
            S.light_n[k] = n;
            V u = cross(b, n), w2 = cross(n, a);
            S.light_du[k] = mul(1 / dot(a, u), u);
            S.light_dv[k] = mul(1 / dot(b, w2), w2);
            power[k] = PI * area * Y * (L->double_sided ? 2 : 1);
        } else if (L->type == 3) {
            /* TriangleMesh ctor (triangle_mesh.cpp:69-85): cumulative areas,
             * normalised, the last exactly 1 */
            double* cdf = malloc(sizeof(double) * (L->n_triangles ? L->n_triangles : 1));
            double acc = 0;
            for (uint32_t j = 0; j < L->n_triangles; ++j) {
                uint32_t tri = L->first_triangle + j;
                const float* a = verts + 3 * (size_t)tris[3 * tri];
                const float* b = verts + 3 * (size_t)tris[3 * tri + 1];
                const float* c = verts + 3 * (size_t)tris[3 * tri + 2];
                V cr = cross(sub(f3(b), f3(a)), sub(f3(c), f3(a)));
                acc += 0.5 * sqrt(dot(cr, cr));
                cdf[j] = acc;
            }
            for (uint32_t j = 0; j < L->n_triangles; ++j) cdf[j] /= acc;
            if (L->n_triangles) cdf[L->n_triangles - 1] = 1;
            S.light_tcdf[k] = cdf;
            S.light_area[k] = acc;
            power[k] = PI * acc * Y * (L->double_sided ? 2 : 1);
        } else if (L->type == 1) {
            power[k] = Y * PI * far_away * far_away * 4 * PI;
        } else {
            S.light_n[k] = nrm(f3(L->direction));
            S.light_cos[k] = cos(0.5 * L->angle);
            double solid = 2 * PI * (1 - cos(0.5 * L->angle));
            S.light_invsolid[k] = 1 / solid;
            power[k] = Y * solid * PI * far_away * far_away;
        }
        total += power[k];
    }
    double acc = 0;
    for (uint32_t k = 0; k < n_lights; ++k) {
        acc += power[k];
        S.light_cdf[k] = total > 0 ? acc / total : (double)(k + 1) / n_lights;
        S.light_sel[k] = total > 0 ? power[k] / total : 1.0 / n_lights;
    }
    if (n_lights) S.light_cdf[n_lights - 1] = 1;
    free(power);

    Ctx c = {&S, P, (uint32_t)P->seed, (uint32_t)(P->seed >> 32), dl};
    Job proto;
    memset(&proto, 0, sizeof proto);
    proto.c = &c;
    {
        V d = nrm(f3(cam->direction)), sk = f3(cam->sky);
        V r = nrm(cross(d, sk)), dn = nrm(cross(d, r));
        double width = 0.036, height = width / cam->aspect, focal = (width / 2) / tan(cam->fov / 2);
        proto.pos = f3(cam->position);
        proto.dir = mul(focal, d);
        proto.right = mul(width, r);
        proto.down = mul(height, dn);
        proto.dirn = d;
        proto.near_limit = cam->near_limit;
    }
    {
        uint32_t q = (uint32_t)lround(sqrt((double)P->spp));
        proto.strata = q * q == P->spp ? q : 0;
    }
    proto.xyz = xyz;
    proto.weight = weight;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    Job jobs[256];
    if (P->sampler == 1) {
        if (dl) return -1;  /* the Halton sampler serves the path tracer */
        pthread_once(&g_honce, halton_init);
        if (P->sampler_generations * 10 > 284) return -1;
        uint64_t np = (uint64_t)P->width * P->height, total = np * P->spp * P->wavelengths;
        uint64_t per = (total + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
        int used = 0;
        for (int k = 0; k < nthreads; ++k) {
            jobs[k] = proto;
            jobs[k].s0 = (uint64_t)k * per < total ? (uint64_t)k * per : total;
            jobs[k].s1 = jobs[k].s0 + per < total ? jobs[k].s0 + per : total;
            if (jobs[k].s0 >= jobs[k].s1) break;
            jobs[k].xyz = calloc(3 * np, sizeof(double));
            jobs[k].weight = calloc(np, sizeof(double));
            pthread_create(&th[k], NULL, run_halton, &jobs[k]);
            ++used;
        }
        for (int k = 0; k < used; ++k) {
            pthread_join(th[k], NULL);
            for (uint64_t i = 0; i < 3 * np; ++i) xyz[i] += jobs[k].xyz[i];
            for (uint64_t i = 0; i < np; ++i) weight[i] += jobs[k].weight[i];
            free(jobs[k].xyz);
            free(jobs[k].weight);
        }
        nthreads = 0;  /* done */
    }
    /* interleave rows across threads in small blocks for balance */
    uint32_t H = nthreads ? P->height : 0;
    uint32_t per = (H + (uint32_t)nthreads - 1) / (uint32_t)nthreads;
    int used = 0;
    for (int k = 0; k < nthreads; ++k) {
        jobs[k] = proto;
        jobs[k].row0 = (uint32_t)k * per < H ? (uint32_t)k * per : H;
        jobs[k].row1 = jobs[k].row0 + per < H ? jobs[k].row0 + per : H;
        if (jobs[k].row0 >= jobs[k].row1) break;
        pthread_create(&th[k], NULL, run_rows, &jobs[k]);
        ++used;
    }
    for (int k = 0; k < used; ++k) pthread_join(th[k], NULL);
    free(S.ow); free(S.ocdf); free(S.oxyz); free(S.spec_reflscale);
    free(S.light_sel); free(S.light_cdf); free(S.light_area); free(S.light_n); free(S.light_du);
    free(S.light_dv); free(S.light_cos); free(S.light_invsolid); free(S.light_ns);
    for (uint32_t k = 0; k < n_lights; ++k) free(S.light_tcdf[k]);
    free(S.light_tcdf);
    return 0;
}

/* AdjointPhotonTracer render into xyz (3*W*H doubles, += sums) and weight
 * (W*H, += counts). */
int or_render(const float* verts, const uint32_t* tris, uint32_t nt, const float* spheres, uint32_t ns,
              const float* tri_normals, const uint32_t* prim_mat, const or_material* mats, const or_spectrum* spectra,
              uint32_t n_spectra, const or_light* lights, uint32_t n_lights, const float* obs_w,
              const float* obs_xyz, uint32_t obs_n, const or_camera* cam, const or_params* P, double* xyz,
              double* weight, int nthreads) {
    return render_impl(verts, tris, nt, spheres, ns, tri_normals, prim_mat, mats, spectra, n_spectra, lights, n_lights, obs_w,
                       obs_xyz, obs_n, cam, P, xyz, weight, nthreads, NULL);
}

/* DirectLighting render (cfg: max generation, secondary light samples,
 * trace glossy); P's depth and roulette fields are ignored. */
int or_render_direct(const float* verts, const uint32_t* tris, uint32_t nt, const float* spheres, uint32_t ns,
                     const float* tri_normals, const uint32_t* prim_mat, const or_material* mats, const or_spectrum* spectra,
                     uint32_t n_spectra, const or_light* lights, uint32_t n_lights, const float* obs_w,
                     const float* obs_xyz, uint32_t obs_n, const or_camera* cam, const or_params* P,
                     const uint32_t* cfg, double* xyz, double* weight, int nthreads) {
    DLCfg dl = {cfg[0], cfg[1] ? cfg[1] : 1u, cfg[2]};
    return render_impl(verts, tris, nt, spheres, ns, tri_normals, prim_mat, mats, spectra, n_spectra, lights, n_lights, obs_w,
                       obs_xyz, obs_n, cam, P, xyz, weight, nthreads, &dl);
}
