// lr_geom.h -- FP32 primitive tests and counter-based RNG shared by the host
// library code and the gfx950 kernels.  Every expression is written with
// explicit fmaf() in a fixed order and the library is built with
// -ffp-contract=off, so host and device evaluate bit-identical results.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define LR_HD __host__ __device__ __forceinline__
#else
#define LR_HD inline
#endif

namespace lr {

struct f3 {
    float x, y, z;
};

LR_HD uint32_t as_u32(float f) {
    union { float f; uint32_t u; } c;
    c.f = f;
    return c.u;
}
LR_HD float as_f32(uint32_t u) {
    union { float f; uint32_t u; } c;
    c.u = u;
    return c.f;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel Random Numbers: As
// Easy as 1, 2, 3", SC'11).  Multipliers and Weyl increments from the paper.
// ---------------------------------------------------------------------------
struct u4 {
    uint32_t x, y, z, w;
};

LR_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

LR_HD u4 philox4x32(uint32_t k0, uint32_t k1, u4 c) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, c.x, hi0, lo0);
        mulhilo(M1, c.z, hi1, lo1);
        u4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// uint32 -> float in [0, 1): top 24 bits, exact.
LR_HD float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------
// Ray-triangle: Moller-Trumbore, two-sided.  Returns true and (t, u, v) when
// the ray's line crosses the triangle (t unrestricted); the caller applies
// the interval rules.  e1 = v1 - v0, e2 = v2 - v0.
// ---------------------------------------------------------------------------
LR_HD bool tri_intersect(float ox, float oy, float oz, float dx, float dy, float dz,
                         float v0x, float v0y, float v0z, float e1x, float e1y, float e1z,
                         float e2x, float e2y, float e2z, float& t, float& u, float& v) {
    const float px = fmaf(dy, e2z, -(dz * e2y));
    const float py = fmaf(dz, e2x, -(dx * e2z));
    const float pz = fmaf(dx, e2y, -(dy * e2x));
    const float det = fmaf(e1x, px, fmaf(e1y, py, e1z * pz));
    if (det == 0.0f) return false;
    const float inv = 1.0f / det;
    const float sx = ox - v0x, sy = oy - v0y, sz = oz - v0z;
    const float uu = fmaf(sx, px, fmaf(sy, py, sz * pz)) * inv;
    if (!(uu >= 0.0f && uu <= 1.0f)) return false;
    const float qx = fmaf(sy, e1z, -(sz * e1y));
    const float qy = fmaf(sz, e1x, -(sx * e1z));
    const float qz = fmaf(sx, e1y, -(sy * e1x));
    const float vv = fmaf(dx, qx, fmaf(dy, qy, dz * qz)) * inv;
    if (!(vv >= 0.0f && uu + vv <= 1.0f)) return false;
    t = fmaf(e2x, qx, fmaf(e2y, qy, e2z * qz)) * inv;
    u = uu;
    v = vv;
    return true;
}

// ---------------------------------------------------------------------------
// Ray-sphere: nearest root t > tmin of |o + t d - c|^2 = r^2 with |d| = 1.
// Discriminant from the perpendicular offset l = oc - (oc.d) d (no
// cancellation for distant spheres), roots q and c'/q.
// ---------------------------------------------------------------------------
LR_HD bool sphere_intersect(float ox, float oy, float oz, float dx, float dy, float dz,
                            float cx, float cy, float cz, float r2, float tmin, float& t) {
    const float sx = ox - cx, sy = oy - cy, sz = oz - cz;
    const float b = fmaf(sx, dx, fmaf(sy, dy, sz * dz));
    const float lx = fmaf(-b, dx, sx), ly = fmaf(-b, dy, sy), lz = fmaf(-b, dz, sz);
    const float disc = r2 - fmaf(lx, lx, fmaf(ly, ly, lz * lz));
    if (!(disc >= 0.0f)) return false;
    const float cc = fmaf(sx, sx, fmaf(sy, sy, sz * sz)) - r2;
    const float sq = sqrtf(disc);
    const float q = -b - copysignf(sq, b);
    float t0, t1;
    if (q == 0.0f) {
        t0 = 0.0f;
        t1 = 0.0f;
    } else {
        const float ta = cc / q, tb = q;
        t0 = fminf(ta, tb);
        t1 = fmaxf(ta, tb);
    }
    if (t0 > tmin) {
        t = t0;
        return true;
    }
    if (t1 > tmin) {
        t = t1;
        return true;
    }
    return false;
}

// LiAR's BoundedRay::inRange (src/liar/kernel/bounded_ray.h:69-77).
LR_HD bool in_range(float t, float tnear, float tfar) {
    return t > tnear * (1.0f + 1e-6f) && t < tfar * (1.0f - 1e-6f);
}

}  // namespace lr
