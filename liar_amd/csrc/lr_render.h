// lr_render.h -- device-side data layout and inline functions of the spectral
// wavefront path tracer: spectra at one sampled wavelength, the standard
// observer, sampling warps, the five BSDFs and the lights.
//
// Semantics follow LiAR (paths relative to the reference repository root);
// every function cites the file it restates.  Sampling warps whose exact
// form lives in the unvendored Lass library (cosine hemisphere, uniform
// sphere) use the published maps named at each function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/liar_amd.h"
#include "lr_geom.h"

namespace lr {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTolerance = 1e-6f;  // src/liar/kernel/kernel_common.cpp:29-33

struct DevSpectrum {
    uint32_t kind, count, offset, pad;
    float value, lo, hi, refl_scale;  // refl_scale: Reflectant normalisation
    float bb_t, bb_k, pad1, pad2;     // BlackBody temperature and 1/planck(peak)*scale
};

struct DevLight {
    uint32_t type, radiance, double_sided, samples;  // samples: DirectLighting emission samples
    float o[3], area;
    float dx[3], cos_max;   // cos_max: sun cone
    float dy[3], inv_solid;  // inv_solid: 1 / sun cone solid angle
    float n[3], sel_pdf;     // n: unit front normal (parallelogram) / sun direction
    float du[3], cdf;        // du, dv: dual basis, (p - o).du = u
    float dv[3], pad2;
    uint32_t tri0, ntri;     // mesh: records tri0 .. tri0 + ntri - 1 of DevScene::light_tris
    uint32_t pad3, pad4;
};

struct DevMaterial {
    uint32_t type, refl, trans, eta, kappa, emission, light;
    float roughness, roughness_v, fuzz;  // CookTorrance roughness along u / v; Mirror fuzz
};

// Observer tables restricted to the sampling domain (n entries).
struct DevObserver {
    const float* w;    // wavelengths [nm]
    const float* cdf;  // cumulative, cdf[0] = 0, cdf[n-1] = 1
    const float* xyz;  // 3*n: x, y, z sensitivities (Y of a unit spectrum = 1)
    const float* dxyz; // 3*n: slopes d(xyz)/dw on [w[k], w[k+1])
    uint32_t n;
};

struct DevScene {
    const lr_node_q8* nodes;
    const lr_prim* prims;
    const uint32_t* prim_mat;
    const DevMaterial* mats;
    const DevSpectrum* spectra;
    const float* table;
    const DevLight* lights;
    const float4* light_tris;  // mesh lights: 3 per triangle, (v0, area cdf) (e1, 0) (e2, 0)
    uint32_t n_lights;
    uint32_t n_triangles;
    DevObserver obs;
};

// ---------------------------------------------------------------------------
// spectra
// ---------------------------------------------------------------------------

// Planck's law with the constants of src/liar/spectra/black_body.cpp:47-59.
__device__ __host__ inline float planck(float w_m, float temperature) {
    const float h = 6.626070040e-34f, c = 299792458.0f, kb = 1.3806488e-23f;
    const float c1 = 2.0f * h * c * c;
    const float c2 = h * c / kb;
    const float w5 = (w_m * w_m) * (w_m * w_m) * w_m;
    return c1 / (w5 * expm1f(c2 / (w_m * temperature)));
}

// Spectrum value at wavelength w [nm].  `reflectant` applies LiAR's
// Reflectant normalisation (src/liar/kernel/spectral.cpp:200-207).
__device__ inline float spectrum_eval(const DevScene& S, uint32_t idx, float w, bool reflectant) {
    if (idx == LR_NONE) return 0.0f;
    const DevSpectrum sp = S.spectra[idx];
    float v = 0.0f;
    switch (sp.kind) {
    case LR_SPEC_CONSTANT:
        return sp.value;
    case LR_SPEC_SAMPLED: {
        const float* ws = S.table + sp.offset;
        const float* vs = ws + sp.count;
        // upper_bound
        uint32_t lo = 0, hi = sp.count;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ws[mid] <= w) lo = mid + 1;
            else hi = mid;
        }
        if (lo == 0 || lo == sp.count) return 0.0f;
        const float t = (w - ws[lo - 1]) / (ws[lo] - ws[lo - 1]);
        v = vs[lo - 1] + t * (vs[lo] - vs[lo - 1]);
        break;
    }
    case LR_SPEC_BINNED: {
        if (!(w >= sp.lo && w < sp.hi)) return 0.0f;
        uint32_t k = (uint32_t)((w - sp.lo) / (sp.hi - sp.lo) * (float)sp.count);
        k = k < sp.count ? k : sp.count - 1;
        v = S.table[sp.offset + k];
        break;
    }
    case LR_SPEC_BLACKBODY:
        return sp.bb_k * planck(w * 1e-9f, sp.bb_t);
    default:
        return 0.0f;
    }
    return reflectant ? v * sp.refl_scale : v;
}

// Observer::sample (src/liar/kernel/observer.cpp:283-317): piecewise-constant
// pdf per table interval, proportional to the trapezoid of x+y+z.
__device__ inline float observer_sample(const DevObserver& O, float u, float& pdf) {
    uint32_t lo = 0, hi = O.n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (O.cdf[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    if (lo == 0 || lo >= O.n) {
        pdf = 0.0f;
        return O.w[0];
    }
    const float c1 = O.cdf[lo - 1], c2 = O.cdf[lo];
    const float dc = c2 - c1;
    const float x = (u - c1) / dc;
    const float w1 = O.w[lo - 1], w2 = O.w[lo];
    pdf = dc / (w2 - w1);
    return w1 + x * (w2 - w1);
}

// Observer::sensitivity (src/liar/kernel/observer.cpp:156-171).
__device__ inline void observer_xyz(const DevObserver& O, float w, float xyz[3]) {
    uint32_t lo = 0, hi = O.n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (O.w[mid] <= w) lo = mid + 1;
        else hi = mid;
    }
    if (lo == 0 || lo == O.n) {
        // w == last wavelength is the only in-domain value that lands here
        const bool last = lo == O.n && w == O.w[O.n - 1];
        for (int a = 0; a < 3; ++a) xyz[a] = last ? O.xyz[3 * (O.n - 1) + a] : 0.0f;
        return;
    }
    const uint32_t k = lo - 1;
    const float dw = w - O.w[k];
    for (int a = 0; a < 3; ++a) xyz[a] = O.xyz[3 * k + a] + O.dxyz[3 * k + a] * dw;
}

// ---------------------------------------------------------------------------
// Halton sampler (src/liar/samplers/halton.cpp:180-230, halton.h:41-66)
// ---------------------------------------------------------------------------

// ScrambledRadicalInverse of index i in base b: digit d of i (least
// significant first) contributes perm[d] b^-(k+1), summed in FP32 in that
// order; perm == nullptr is the plain radical inverse.
__device__ __host__ inline float radical_inverse(uint64_t i, uint32_t b, const uint16_t* perm) {
    const float inv = 1.0f / (float)b;
    float w = inv, r = 0.0f;
    if ((i >> 32) == 0) {
        uint32_t j = (uint32_t)i;
        while (j) {
            const uint32_t d = j % b;
            j /= b;
            r = r + (float)(perm ? perm[d] : d) * w;
            w = w * inv;
        }
        return r;
    }
    while (i) {
        const uint32_t d = (uint32_t)(i % b);
        i /= b;
        r = r + (float)(perm ? perm[d] : d) * w;
        w = w * inv;
    }
    return r;
}

// ---------------------------------------------------------------------------
// vectors and warps
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline V3 normalize(V3 a) { return (1.0f / sqrtf(dot(a, a))) * a; }
__device__ inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// Orthonormal basis around unit n: Duff et al., "Building an Orthonormal
// Basis, Revisited", JCGT 2017.
__device__ inline void onb(V3 n, V3& t, V3& b) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float bb = n.x * n.y * a;
    t = v3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    b = v3(bb, sign + n.y * n.y * a, -n.y);
}

// Cosine-weighted hemisphere via Shirley-Chiu concentric disk + Malley's
// method; pdf = cos/pi.
__device__ inline V3 cosine_hemisphere(float u1, float u2, float& pdf) {
    const float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
    float r, phi;
    if (a == 0.0f && b == 0.0f) {
        r = 0.0f;
        phi = 0.0f;
    } else if (a * a > b * b) {
        r = a;
        phi = (kPi / 4.0f) * (b / a);
    } else {
        r = b;
        phi = (kPi / 2.0f) - (kPi / 4.0f) * (a / b);
    }
    const float x = r * cosf(phi), y = r * sinf(phi);
    const float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
    pdf = z / kPi;
    return v3(x, y, z);
}

// Uniform sphere: z = 1 - 2 u1, phi = 2 pi u2; pdf = 1 / (4 pi).
__device__ inline V3 uniform_sphere(float u1, float u2) {
    const float z = 1.0f - 2.0f * u1;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 2.0f * kPi * u2;
    return v3(r * cosf(phi), r * sinf(phi), z);
}

// ---------------------------------------------------------------------------
// BSDFs in the local frame (z = shading normal flipped towards the incoming
// ray, omega_in.z > 0).  LiAR masks the caps by the geometric hemispheres
// (src/liar/kernel/bsdf.cpp:52-76): bsdf_evaluate takes the geometric
// normal in the local frame for that test.
// ---------------------------------------------------------------------------
struct BsdfParams {
    uint32_t type;
    float refl, trans;       // spectral values at the path's wavelength
    float eta, kappa;        // conductor n, k
    float ior;               // dielectric n_incident / n_transmitted
    float alpha_u, alpha_v;  // Cook-Torrance max(roughness, 1e-3)^2 along the frame's u / v
    float fuzz;              // Mirror fuzz (0: a perfect mirror)
};

__device__ inline bool bsdf_is_diffuse_or_glossy(uint32_t type) {
    return type == LR_MAT_LAMBERT || type == LR_MAT_COOK_TORRANCE || type == LR_MAT_COOK_TORRANCE_BECKMANN;
}

// Conductor Fresnel of src/liar/shaders/conductor.cpp:184-194 and
// cook_torrance.cpp:242-256.
__device__ inline float fresnel_conductor(float eta, float kappa, float cos_i) {
    const float cos2 = cos_i * cos_i;
    const float a = eta * eta + kappa * kappa;
    const float b = 2.0f * cos_i * eta;
    const float a_orth = a + cos2;
    const float a_par = a * cos2 + 1.0f;
    const float r_orth2 = (a_orth - b) / (a_orth + b);
    const float r_par2 = (a_par - b) / (a_par + b);
    return 0.5f * (r_orth2 + r_par2);
}

// MicrofacetDistribution::samplePhi (src/liar/kernel/microfacet.cpp:102-123).
__device__ inline float sample_phi(float u, float au, float av) {
    if (au == av) return 2.0f * kPi * u;
    const float phi = atanf(av / au * tanf(2.0f * kPi * u));
    if (u <= 0.25f) return phi;
    if (u < 0.75f) return phi + kPi;
    return phi + 2.0f * kPi;
}

// Trowbridge-Reitz D and pdf of h (microfacet_trowbridge_reitz.cpp:46-58).
__device__ inline float ggx_d(V3 h, float au, float av, float& pdf_h) {
    const float hx = h.x / au, hy = h.y / av;
    const float q = hx * hx + hy * hy + h.z * h.z;
    const float d = 1.0f / (kPi * au * av * q * q);
    pdf_h = d * h.z;
    return d;
}

// Beckmann D and pdf of h (microfacet_beckmann.cpp:50-72).
__device__ inline float beckmann_d(V3 h, float au, float av, float& pdf_h) {
    const float c2 = h.z * h.z;
    if (c2 == 0.0f) {
        pdf_h = 0.0f;
        return 0.0f;
    }
    const float hx = h.x / au, hy = h.y / av;
    const float d = expf(-(hx * hx + hy * hy) / c2) / (kPi * au * av * c2 * c2);
    pdf_h = d * h.z;
    return d;
}

// sampleH of the two distributions: Trowbridge-Reitz
// (microfacet_trowbridge_reitz.cpp:62-72), Beckmann (microfacet_beckmann.cpp:
// 76-106, whose anisotropic branch draws phi from u1, the number that also
// sets the radial part, as the reference does).
__device__ inline V3 sample_h(uint32_t type, float au, float av, float u1, float u2) {
    if (type == LR_MAT_COOK_TORRANCE) {
        const float phi = sample_phi(u1, au, av);
        const float s = sqrtf(u2 / (1.0f - u2));
        return normalize(v3(s * au * cosf(phi), s * av * sinf(phi), 1.0f));
    }
    const float s = logf(1.0f - u1);
    float cp, sp, tan2;
    if (au == av) {
        const float phi = 2.0f * kPi * u2;
        cp = cosf(phi);
        sp = sinf(phi);
        tan2 = -s * au * au;
    } else {
        const float phi = sample_phi(u1, au, av);
        cp = cosf(phi);
        sp = sinf(phi);
        const float a = cp / au, b = sp / av;
        tan2 = -s / (a * a + b * b);
    }
    const float cos2 = 1.0f / (1.0f + tan2);
    const float ct = sqrtf(cos2), st = sqrtf(fmaxf(1.0f - cos2, 0.0f));
    return v3(cp * st, sp * st, ct);
}

// Cook-Torrance V-cavity G (cook_torrance.cpp:234-238).
__device__ inline float ct_g(V3 wi, V3 wo, V3 h) {
    const float doh = fabsf(dot(wi, h));
    return fminf(1.0f, fminf(2.0f * h.z * wi.z / doh, 2.0f * h.z * wo.z / doh));
}

// doEvaluate of the non-delta BSDFs; returns false for "no value".
__device__ inline bool bsdf_eval(const BsdfParams& P, V3 wi, V3 wo, float& value, float& pdf) {
    if (wo.z == 0.0f) return false;
    if (P.type == LR_MAT_LAMBERT) {  // lambert.cpp:105-114
        value = P.refl / kPi;
        pdf = fabsf(wo.z) / kPi;
        return pdf > 0.0f;
    }
    if (P.type == LR_MAT_COOK_TORRANCE || P.type == LR_MAT_COOK_TORRANCE_BECKMANN) {  // cook_torrance.cpp:262-280
        if (!(wo.z > 0.0f)) return false;
        const V3 h = normalize(wi + wo);
        float pdf_h;
        const float d = P.type == LR_MAT_COOK_TORRANCE ? ggx_d(h, P.alpha_u, P.alpha_v, pdf_h)
                                                       : beckmann_d(h, P.alpha_u, P.alpha_v, pdf_h);
        const float g = ct_g(wi, wo, h);
        const float woh = dot(wi, h);
        pdf = pdf_h / (4.0f * woh);
        value = P.refl * (d * g / (4.0f * wi.z * wo.z)) *
                fresnel_conductor(P.eta, P.kappa, fabsf(woh));
        return pdf > 0.0f;
    }
    return false;  // delta BSDFs evaluate to zero (mirror.cpp:135-138 etc.)
}

// Bsdf::evaluate (bsdf.cpp:52-76): a wo on the other side of the geometric
// normal (ngl, local frame, towards the viewer) than wi masks the reflection
// caps, and every BSDF here evaluates reflection only.
__device__ inline bool bsdf_evaluate(const BsdfParams& P, V3 wi, V3 wo, V3 ngl, float& value, float& pdf) {
    if (wo.z == 0.0f) return false;
    if ((dot(wi, ngl) > 0.0f) != (dot(wo, ngl) > 0.0f)) return false;
    return bsdf_eval(P, wi, wo, value, pdf);
}

// Bsdf::sample.  u1, u2: 2D sample; uc: component sample.
// allow: bit 0 reflection, bit 1 transmission (the caps a caller permits;
// only the dielectric has both, and chooses among the permitted ones).
__device__ inline bool bsdf_sample(const BsdfParams& P, V3 wi, float u1, float u2, float uc, V3& wo, float& value,
                                   float& pdf, uint32_t allow = 3u) {
    switch (P.type) {
    case LR_MAT_LAMBERT: {  // lambert.cpp:116-124
        wo = cosine_hemisphere(u1, u2, pdf);
        value = P.refl / kPi;
        return pdf > 0.0f;
    }
    case LR_MAT_MIRROR: {  // mirror.cpp:142-161: value = reflectance, pdf = cos
        wo = v3(-wi.x, -wi.y, wi.z);
        if (P.fuzz > 0.0f) {
            // fuzz: the reflection plus fuzz x a uniform-sphere point, left
            // unnormalised (its z is the pdf, so value |cos| / pdf = refl)
            wo = wo + P.fuzz * uniform_sphere(u1, u2);
            value = wo.z > 0.0f ? P.refl : 0.0f;
        } else {
            value = P.refl;
        }
        pdf = wo.z;
        return pdf > 0.0f;
    }
    case LR_MAT_CONDUCTOR: {  // conductor.cpp:179-198
        const float cos_i = wi.z;
        wo = v3(-wi.x, -wi.y, wi.z);
        value = P.refl * fresnel_conductor(P.eta, P.kappa, cos_i) / cos_i;
        pdf = 1.0f;
        return true;
    }
    case LR_MAT_DIELECTRIC: {  // dielectric.cpp:209-247
        const float cos_i = wi.z;
        if (cos_i <= 0.0f) return false;
        const float ior = P.ior;
        const float sin_t2 = ior * ior * (1.0f - cos_i * cos_i);
        const float cos_t = sqrtf(fmaxf(1.0f - sin_t2, 0.0f));
        const float r_orth = (ior * cos_i - cos_t) / (ior * cos_i + cos_t);
        const float r_par = (cos_i - ior * cos_t) / (cos_i + ior * cos_t);
        const float fr = 0.5f * (r_orth * r_orth + r_par * r_par);
        const float pow_r = (allow & 1u) ? fabsf(P.refl) * fr : 0.0f;
        const float pow_t = (allow & 2u) ? fabsf(P.trans) * (1.0f - fr) : 0.0f;
        if (pow_r + pow_t == 0.0f) return false;
        const float prob_r = pow_r / (pow_r + pow_t);
        if (uc < prob_r) {
            wo = v3(-wi.x, -wi.y, wi.z);
            value = P.refl * fr / cos_i;
            pdf = prob_r;
        } else {
            wo = v3(-ior * wi.x, -ior * wi.y, -cos_t);
            value = P.trans * (1.0f - fr) / cos_t;
            pdf = 1.0f - prob_r;
        }
        return pdf > 0.0f;
    }
    case LR_MAT_COOK_TORRANCE:
    case LR_MAT_COOK_TORRANCE_BECKMANN: {  // cook_torrance.cpp:284-294
        const V3 h = sample_h(P.type, P.alpha_u, P.alpha_v, u1, u2);
        wo = (2.0f * dot(wi, h)) * h - wi;
        if (!(wo.z > 0.0f)) return false;
        return bsdf_eval(P, wi, wo, value, pdf);
    }
    default:
        return false;
    }
}

// ---------------------------------------------------------------------------
// hit -> surface frame (IntersectionContext, src/liar/kernel/
// intersection_context.cpp:41-72,197-223,278-291)
// ---------------------------------------------------------------------------

// Per-primitive shading geometry of a scene (by primitive id):
//   geom[p]: triangle (unit geometric normal, corner-normal record index or
//            LR_NONE as bits); sphere (centre, radius)
//   tan[p]:  triangle frame for the geometric normal: (unit vector, code),
//            code 0 = any orthonormal pair (both edges shorter than
//            sqrt(1e-3)), 1 = the vector is u (v = n x u), 2 = it is v
//            (u = v x n)
//   cn[4 r]: corner-normal record r: (n0, e1.x) (n1, e1.y) (n2, e1.z) (e2, 0)
struct GeomArgs {
    const float4* geom;
    const float4* tan;
    const float4* cn;
};

struct HitFrame {
    V3 x;       // hit point
    V3 ns;      // shading normal, towards the viewer
    V3 ng;      // geometric normal, towards the viewer
    V3 tt, bb;  // BSDF frame tangents (u, v)
    bool leaving;
    bool has_ns;  // ns comes from corner normals (else ns == ng)
};

// bsdfToLocal's choice: u = ns.reject(dP/du), v = ns.reject(dP/dv); both
// shorter than sqrt(1e-3): any orthonormal pair; then the shorter one
// is replaced by the cross product with ns of the longer one.
__device__ inline void frame_from_derivatives(V3 ns, V3 dpdu, V3 dpdv, V3& tt, V3& bb) {
    V3 u = dpdu - dot(dpdu, ns) * ns;
    V3 v = dpdv - dot(dpdv, ns) * ns;
    const float su = dot(u, u), sv = dot(v, v);
    if (fmaxf(su, sv) < 1e-3f) onb(ns, u, v);
    if (su >= sv) v = cross(ns, u);
    else u = cross(v, ns);
    tt = normalize(u);
    bb = normalize(v);
}

__device__ inline HitFrame hit_frame(const GeomArgs& G, uint32_t n_triangles, uint32_t prim, V3 o, V3 d, float tnear,
                                     float t, float bu, float bv) {
    HitFrame F;
    F.x = o + t * d;
    const float4 g = G.geom[prim];
    F.has_ns = false;
    if (prim < n_triangles) {
        const V3 ng = v3(g.x, g.y, g.z);
        F.leaving = dot(d, ng) > 0.0f;  // triangle_mesh.cpp:304-309
        const uint32_t rec = __float_as_uint(g.w);
        F.ng = dot(ng, d) > 0.0f ? -ng : ng;
        if (rec == LR_NONE) {
            F.ns = F.ng;
            const float4 tv = G.tan[prim];
            const uint32_t code = __float_as_uint(tv.w);
            const V3 a = v3(tv.x, tv.y, tv.z);
            if (code == 1u) {
                F.tt = a;
                F.bb = normalize(cross(F.ns, a));
            } else if (code == 2u) {
                F.bb = a;
                F.tt = normalize(cross(a, F.ns));
            } else {
                onb(F.ns, F.tt, F.bb);
            }
        } else {
            // interpolated corner normals (triangle_mesh.cpp:334-358)
            const float4* r = G.cn + 4 * (size_t)rec;
            const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            const float b0 = 1.0f - bu - bv;
            V3 ns = normalize(b0 * v3(r0.x, r0.y, r0.z) + bu * v3(r1.x, r1.y, r1.z) + bv * v3(r2.x, r2.y, r2.z));
            if (dot(ns, d) > 0.0f) ns = -ns;
            F.ns = ns;
            F.has_ns = true;
            frame_from_derivatives(ns, v3(r0.w, r1.w, r2.w), v3(r3.x, r3.y, r3.z), F.tt, F.bb);
        }
    } else {
        const V3 c = v3(g.x, g.y, g.z);
        const float r = g.w;
        const V3 n = normalize(F.x - c);
        const V3 s0 = (o + tnear * d) - c;
        F.leaving = dot(s0, s0) < r * r;  // sphere.cpp:110-111
        // (phi / 2pi, theta / pi) parameter derivatives (sphere.cpp:130-171)
        const float st = sqrtf(fmaxf(0.0f, 1.0f - n.z * n.z));
        const V3 dpdu = (2.0f * kPi * r) * v3(-n.y, n.x, 0.0f);
        V3 dpdv;
        if (st != 0.0f) {
            const float cs = n.z / st;
            dpdv = (kPi * r) * v3(cs * n.x, cs * n.y, -st);
        } else {
            const float phi = atan2f(n.y, n.x);
            dpdv = (kPi * r) * v3(cosf(phi), sinf(phi), 0.0f);
        }
        F.ng = dot(n, d) > 0.0f ? -n : n;
        F.ns = F.ng;
        frame_from_derivatives(F.ns, dpdu, dpdv, F.tt, F.bb);
    }
    return F;
}

// The geometric normal in the BSDF frame (bsdf_evaluate's hemisphere test).
__device__ inline V3 local_geometric_normal(const HitFrame& F) {
    return F.has_ns ? v3(dot(F.ng, F.tt), dot(F.ng, F.bb), dot(F.ng, F.ns)) : v3(0.0f, 0.0f, 1.0f);
}

// ---------------------------------------------------------------------------
// lights (src/liar/scenery/light_area.cpp:189-252, light_sky.cpp:239-297,
// src/liar/kernel/light_context.cpp:309-331)
// ---------------------------------------------------------------------------

// Ray (origin p, unit dir d, tnear) against a parallelogram; returns t or -1.
__device__ inline float parallelogram_hit(const DevLight& L, V3 p, V3 d, float tnear) {
    const V3 n = ld3(L.n);
    const float dn = dot(d, n);
    if (dn == 0.0f) return -1.0f;
    const float t = dot(ld3(L.o) - p, n) / dn;
    if (!(t > tnear)) return -1.0f;
    const V3 q = (p + t * d) - ld3(L.o);
    const float u = dot(q, ld3(L.du)), v = dot(q, ld3(L.dv));
    if (!(u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f)) return -1.0f;
    return t;
}

// A mesh light's own surface along (p, d): TriangleMesh::doIntersect with
// near = tolerance (triangle_mesh.cpp:291-317): the nearest triangle with
// t > near, kept only if inRange; returns t and its unit normal, or -1.
__device__ inline float mesh_light_hit(const DevLight& L, const float4* lt, V3 p, V3 d, V3& n) {
    float best = INFINITY;
    uint32_t bk = 0xFFFFFFFFu;
    for (uint32_t k = 0; k < L.ntri; ++k) {
        const float4* r = lt + 3 * (size_t)(L.tri0 + k);
        const float4 a = r[0], b = r[1], c = r[2];
        float t, u, v;
        if (tri_intersect(p.x, p.y, p.z, d.x, d.y, d.z, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v) &&
            t > kTolerance && t < best) {
            best = t;
            bk = k;
        }
    }
    if (bk == 0xFFFFFFFFu || !in_range(best, kTolerance, INFINITY)) return -1.0f;
    const float4* r = lt + 3 * (size_t)(L.tri0 + bk);
    n = normalize(cross(v3(r[1].x, r[1].y, r[1].z), v3(r[2].x, r[2].y, r[2].z)));
    return best;
}

// LightContext::emission pdf (solid angle) of direction d from point p:
// LightArea::doEmission through SceneObject::doAngularPdf
// (light_area.cpp:189-203, scene_object.cpp:252-266) re-intersects the
// surface; the normal it returns is flipped towards p by the
// IntersectionContext, so the single-sided test there never rejects.
__device__ inline float light_pdf(const DevLight& L, const float4* lt, V3 p, V3 d) {
    if (L.type == LR_LIGHT_PARALLELOGRAM) {
        const float t = parallelogram_hit(L, p, d, kTolerance);
        if (t < 0.0f) return 0.0f;
        return t * t / (L.area * fabsf(dot(d, ld3(L.n))));
    }
    if (L.type == LR_LIGHT_MESH) {
        V3 n;
        const float t = mesh_light_hit(L, lt, p, d, n);
        if (t < 0.0f) return 0.0f;
        return t * t / (L.area * fabsf(dot(d, n)));
    }
    if (L.type == LR_LIGHT_SKY) return 1.0f / (4.0f * kPi);
    // sun
    return dot(d, ld3(L.n)) >= L.cos_max ? L.inv_solid : 0.0f;
}

// Does light L emit along d?  For the sky and sun: any ray that escapes in
// their directions; for an area light: the caller already hit its surface,
// and (as light_pdf explains) LiAR's emission along a ray is not
// single-sided.
__device__ inline bool light_sees(const DevLight& L, V3 d) {
    if (L.type == LR_LIGHT_PARALLELOGRAM || L.type == LR_LIGHT_MESH || L.type == LR_LIGHT_SKY) return true;
    return dot(d, ld3(L.n)) >= L.cos_max;
}

// LightContext::sampleEmission: direction from p towards the light; pdf in
// solid angle (0 = failed).  dist: distance to the sampled point (area
// lights).  Single-sided area lights reject points seen from behind
// (light_area.cpp:207-252).
__device__ inline V3 light_sample(const DevLight& L, const float4* lt, V3 p, float u1, float u2, float& pdf,
                                  float& dist) {
    dist = INFINITY;
    if (L.type == LR_LIGHT_PARALLELOGRAM || L.type == LR_LIGHT_MESH) {
        V3 q, n;
        float pdf_area;
        if (L.type == LR_LIGHT_PARALLELOGRAM) {  // parallelogram.cpp:112-118
            q = ld3(L.o) + u1 * ld3(L.dx) + u2 * ld3(L.dy);
            n = ld3(L.n);
            pdf_area = 1.0f / L.area;
        } else {  // TriangleMesh::doSampleSurface (triangle_mesh.cpp:414-444)
            const float4* r = lt + 3 * (size_t)L.tri0;
            uint32_t lo = 0, hi = L.ntri;  // lower_bound of u1 on the area cdf
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (r[3 * mid].w < u1) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t k = lo < L.ntri ? lo : L.ntri - 1;
            const float c0 = k > 0 ? r[3 * (k - 1)].w : 0.0f;
            const float tp = r[3 * k].w - c0;
            const float di = (u1 - c0) / tp;
            const float sq = sqrtf(di);
            const float a = 1.0f - sq, b = sq * u2;
            const V3 v0 = v3(r[3 * k].x, r[3 * k].y, r[3 * k].z);
            const V3 e1 = v3(r[3 * k + 1].x, r[3 * k + 1].y, r[3 * k + 1].z);
            const V3 e2 = v3(r[3 * k + 2].x, r[3 * k + 2].y, r[3 * k + 2].z);
            q = v0 + a * e1 + b * e2;
            const V3 nv = cross(e1, e2);
            const float nn = sqrtf(dot(nv, nv));
            n = (1.0f / nn) * nv;
            pdf_area = tp / (0.5f * nn);
        }
        V3 to = q - p;
        const float d2 = dot(to, to);
        dist = sqrtf(d2);
        to = (1.0f / dist) * to;
        const float c = dot(n, to);
        pdf = pdf_area * d2 / fabsf(c);
        if (!L.double_sided && c > 0.0f) pdf = 0.0f;
        if (!(pdf < INFINITY)) pdf = 0.0f;
        return to;
    }
    if (L.type == LR_LIGHT_SKY) {
        pdf = 1.0f / (4.0f * kPi);
        return uniform_sphere(u1, u2);
    }
    // sun: uniform cone around L.n
    const float cos_t = 1.0f - u1 * (1.0f - L.cos_max);
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * kPi * u2;
    V3 t, b;
    const V3 n = ld3(L.n);
    onb(n, t, b);
    pdf = L.inv_solid;
    return (sin_t * cosf(phi)) * t + (sin_t * sinf(phi)) * b + cos_t * n;
}

// LightContexts::sample (light_context.cpp:309-319): lower_bound on the cdf.
__device__ inline uint32_t light_pick(const DevScene& S, float u) {
    uint32_t lo = 0, hi = S.n_lights;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (S.lights[mid].cdf < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < S.n_lights ? lo : S.n_lights - 1;
}

}  // namespace lr
