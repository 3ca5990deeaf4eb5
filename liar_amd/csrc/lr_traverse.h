// lr_traverse.h -- device helpers shared by the QBVH traversal kernels
// (single-level in trace.hip, two-level in instance.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/liar_amd.h"

namespace lr {
namespace trav {

// Box tests use (plane - origin) * inv_dir; scaling the far distance by
// 1 + 2*gamma(3) makes the test conservative (Ize, "Robust BVH Ray
// Traversal", JCGT 2013).
constexpr float kIze = 1.0000003576278687f;
// Slack on the current best distance so that FP32 rounding of a primitive's
// hit distance can never cull the box that holds it.
constexpr float kCullSlack = 1.0f + 1.0f / 65536.0f;

struct RaysArg {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tnear, *tfar;
    const uint32_t* root;  // per-ray root node (several trees in one node array), or NULL: node 0
};
struct HitsArg {
    uint32_t* prim;
    float *t, *u, *v;
};

__device__ __forceinline__ float safe_inv(float d) {
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}

// Slab test of the four children of a node; returns per-child entry distance
// (+inf when culled).
__device__ __forceinline__ void node_test(const float4* __restrict__ np, float ox, float oy, float oz, float ix,
                                          float iy, float iz, float tnear, float cull, float t[4], uint32_t e[4]) {
    const float4 lox = np[0], hix = np[1], loy = np[2], hiy = np[3], loz = np[4], hiz = np[5];
    const uint4 ch = reinterpret_cast<const uint4*>(np)[6];
    const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
    const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
    const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
    const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float ax = (lx[i] - ox) * ix, bx = (hx[i] - ox) * ix;
        const float ay = (ly[i] - oy) * iy, by = (hy[i] - oy) * iy;
        const float az = (lz[i] - oz) * iz, bz = (hz[i] - oz) * iz;
        const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tnear));
        const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)) , fmaxf(az, bz)) * kIze;
        const bool hit = tn <= fminf(tf, cull) && c[i] != LR_CHILD_EMPTY;
        t[i] = hit ? tn : INFINITY;
        e[i] = c[i];
    }
}

#define LR_CSWAP(a, b)                     \
    {                                      \
        const bool s_ = t[b] < t[a];       \
        const float ta_ = t[a], tb_ = t[b]; \
        const uint32_t ea_ = e[a], eb_ = e[b]; \
        t[a] = s_ ? tb_ : ta_;             \
        t[b] = s_ ? ta_ : tb_;             \
        e[a] = s_ ? eb_ : ea_;             \
        e[b] = s_ ? ea_ : eb_;             \
    }

template <int BLOCK, int SHORT>
struct Stack {
    uint32_t* s_e;
    float* s_t;
    uint2* spill;
    uint32_t stride;  // threads in the grid
    uint32_t gtid, tid;
    int sp;
    int cap;

    __device__ __forceinline__ void push(uint32_t e, float t) {
        if (sp < SHORT) {
            s_e[sp * BLOCK + tid] = e;
            s_t[sp * BLOCK + tid] = t;
        } else if (sp < cap) {
            spill[(size_t)(sp - SHORT) * stride + gtid] = make_uint2(e, __float_as_uint(t));
        } else {
            return;  // unreachable: cap >= 3 * (max_depth + 1) + 1
        }
        ++sp;
    }
    __device__ __forceinline__ void pop(uint32_t& e, float& t) {
        --sp;
        if (sp < SHORT) {
            e = s_e[sp * BLOCK + tid];
            t = s_t[sp * BLOCK + tid];
        } else {
            const uint2 v = spill[(size_t)(sp - SHORT) * stride + gtid];
            e = v.x;
            t = __uint_as_float(v.y);
        }
    }
};


}  // namespace trav
}  // namespace lr
