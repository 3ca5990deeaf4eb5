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

// Slab test of the four children of a quantised node (lr_node_q8, 64 B =
// four dwordx4 loads); returns per-child entry distance (+inf when culled).
// Each plane decodes as fmaf(q, scale, origin), bit for bit the builder's
// value, then the FP32 slab test runs on those planes.
__device__ __forceinline__ void node_test(const float4* __restrict__ np, float ox, float oy, float oz, float ix,
                                          float iy, float iz, float tnear, float cull, float t[4], uint32_t e[4]) {
    const float4 g0 = np[0], g1 = np[1];
    const uint4 q = reinterpret_cast<const uint4*>(np)[2];
    const uint4 ch = reinterpret_cast<const uint4*>(np)[3];
    const float orx = g0.x, ory = g0.y, orz = g0.z, sx = g0.w, sy = g1.x, sz = g1.y;
    const uint32_t qlx = __float_as_uint(g1.z), qhx = __float_as_uint(g1.w);
    const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int sh = 8 * i;
        const float lx = fmaf((float)((qlx >> sh) & 255u), sx, orx), hx = fmaf((float)((qhx >> sh) & 255u), sx, orx);
        const float ly = fmaf((float)((q.x >> sh) & 255u), sy, ory), hy = fmaf((float)((q.y >> sh) & 255u), sy, ory);
        const float lz = fmaf((float)((q.z >> sh) & 255u), sz, orz), hz = fmaf((float)((q.w >> sh) & 255u), sz, orz);
        const float ax = (lx - ox) * ix, bx = (hx - ox) * ix;
        const float ay = (ly - oy) * iy, by = (hy - oy) * iy;
        const float az = (lz - oz) * iz, bz = (hz - oz) * iz;
        const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tnear));
        const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz)) * kIze;
        const bool hit = tn <= fminf(tf, cull) && c[i] != LR_CHILD_EMPTY;
        t[i] = hit ? tn : INFINITY;
        e[i] = c[i];
    }
}
// Marks a loaded value as used at this point (an empty asm statement, no
// instruction).  The loads of a leaf's records are all issued above it and
// awaited once, instead of the compiler sinking each record's load into the
// branch that reads it, which serialises the fetches.
__device__ __forceinline__ void keep(const float4& v) { asm volatile("" : : "v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w)); }

// Node stride in float4 units.
constexpr size_t kNodeF4 = sizeof(lr_node_q8) / 16;
static_assert(sizeof(lr_node_q8) == 64, "lr_node_q8 is 64 B");

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

// Largest ray count per launch: RayClaim's counter may run past n by up to
// 1024 per wave (at most 8192 waves), which must not wrap 32 bits.
constexpr uint32_t kMaxRays = 0xFF000000u;

__device__ __forceinline__ uint32_t min_u64_u32(uint64_t a, uint32_t b) { return a < b ? (uint32_t)a : b; }

// Ray assignment for the persistent traversal kernels.  About half of the
// rays are split statically, interleaved in groups of 64: wave w of W owns
// groups w, w + W, w + 2W, ... of the first S*W rays (S a multiple of 64 per
// wave), so the rays in flight at any time stay in one window of the ray
// arrays.  The rest are claimed from one global counter in chunks whose size
// shrinks with what is left (guided self-scheduling: remaining / (2 W), at
// most 1024, at least what the wave needs now); a wave hands out a claimed
// chunk over several refills.  One atomic on one address costs ~10 ns,
// serialised over the whole GPU, so claiming only what one refill needs made
// the counter the limit for short rays (a few node visits each).  All members
// are wave-uniform; claim() is called by the whole wave.
// With guided = false, the static share is one group of 64 per wave and each
// refill claims exactly what it needs: for long, uneven rays (the instanced
// traversal) that balances better and the atomics are rare anyway.
struct RayClaim {
    uint32_t n, n_waves, wave, static_total;
    bool guided;
    uint32_t lo, hi;    // the wave's reservation not yet handed out: [lo, hi)
    bool in_static;     // [lo, hi) counts the wave's static rays (else ray indices)
    uint32_t dyn_left;  // rays the counter had not handed out, as last seen
    bool dyn_done;      // the counter has nothing left

    __device__ __forceinline__ void init(uint32_t n_rays, uint32_t w, uint32_t waves, bool guided_share = true) {
        n = n_rays;
        n_waves = waves;
        wave = __builtin_amdgcn_readfirstlane(w);  // uniform: keep it (and what follows) in SGPRs
        guided = guided_share;
        const uint32_t per = guided ? 64u * max(1u, n / (128u * waves)) : 64u;
        static_total = min_u64_u32((uint64_t)per * waves, n);
        in_static = true;
        lo = 0;
        // per == 64 when 64 * waves may exceed n: wave w then owns [64 w, 64 w + 64) clipped to n
        hi = (uint64_t)per * waves <= n ? per : min_u64_u32(n > 64ull * w ? n - 64ull * w : 0ull, 64u);
        dyn_left = n - static_total;
        dyn_done = dyn_left == 0;
    }
    __device__ __forceinline__ bool exhausted() const { return dyn_done && lo >= hi; }
    __device__ __forceinline__ uint32_t static_ray(uint32_t j) const {
        return ((j >> 6) * n_waves + wave) * 64u + (j & 63u);
    }

    // Hands rays to the `idle` lanes set in `need`, in lane order; returns
    // this lane's ray index, or 0xFFFFFFFF when it gets none.
    __device__ __forceinline__ uint32_t claim(uint64_t need, uint32_t idle, uint32_t lane, bool is_idle,
                                              uint32_t* counter) {
        const uint32_t have = hi - lo;
        uint32_t nlo = 0, nhi = 0;
        if (have < idle && !dyn_done) {
            const uint32_t want = idle - have;
            const uint32_t chunk = guided ? max(want, min(1024u, (dyn_left / (2u * n_waves)) & ~63u)) : want;
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(counter, chunk);
            base = __builtin_amdgcn_readlane(base, leader);
            const uint64_t a = (uint64_t)static_total + base, b = a + chunk;
            nlo = min_u64_u32(a, n);
            nhi = min_u64_u32(b, n);
            dyn_left = n - nhi;
            dyn_done = b >= n;
        }
        uint32_t my = 0xFFFFFFFFu;
        if (is_idle) {
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            if (rank < have)
                my = in_static ? static_ray(lo + rank) : lo + rank;
            else if (rank - have < nhi - nlo)
                my = nlo + (rank - have);
        }
        if (have >= idle) {
            lo += idle;
        } else {
            in_static = false;
            lo = min(nlo + (idle - have), nhi);
            hi = nhi;
        }
        return my;
    }
};

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
