// instance.hip -- two-level (instanced) QBVH traversal for gfx950.
//
// A top-level QBVH over instance boxes in world space; each leaf names up to
// four instance slots.  An instance is a geometry tree (the nodes and
// primitive records of an lr_qbvh, all geometries concatenated into one
// array) plus the world-to-local transform.  Entering an instance moves the
// lane's ray into local space (LiAR's Transformation::doIntersect,
// src/liar/scenery/transformation.cpp:117-131: origin W o, direction
// normalise(W d), bounds scaled by s = |W d|), pushes an end-of-instance
// marker and walks the geometry tree with the single-level rules of
// trace.hip; popping the marker turns the instance's local result into a
// world distance t_local / s, keeps it if nearer (List semantics,
// src/liar/scenery/list.cpp:106-133) and restores the world ray.
//
// Stack entries: top-level inner node (bit 31 clear), top-level leaf
// (bit 31 set), "enter instance slot k" (bits 31..30 = 01), geometry nodes
// and leaves while in local space, and the end-of-instance marker.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lr_build.h"
#include "lr_geom.h"
#include "lr_internal.h"
#include "lr_traverse.h"

namespace lr {
namespace inst {

struct InstRec {
    float w[12];    // world -> local, row-major 3x4
    uint32_t root;  // geometry root node in the concatenated node array
    uint32_t id;    // instance index (caller's numbering)
    uint32_t pad[2];
};
static_assert(sizeof(InstRec) == 64, "instance record is one 64 B line");

}  // namespace inst
}  // namespace lr

struct lr_tlas {
    lr_context* ctx = nullptr;
    lr_node_q8* d_tnodes = nullptr;
    lr::inst::InstRec* d_inst = nullptr;
    lr_node_q8* d_nodes = nullptr;
    lr_prim* d_prims = nullptr;
    lr_tlas_info info{};
    uint32_t stack_cap = 0;
};

namespace {

using lr::inst::InstRec;
using namespace lr::trav;

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMarker = 0xFFFFFFFEu;  // end of the current instance
constexpr uint32_t kEnter = 0x40000000u;   // top level: enter instance slot (low 30 bits)
constexpr int kBlock = 256;
constexpr int kShort = 16;

template <int BLOCK, int SHORT, bool ANY>
__global__ __launch_bounds__(BLOCK) void k_trace_inst(const lr_node_q8* __restrict__ tnodes,
                                                      const InstRec* __restrict__ insts,
                                                      const lr_node_q8* __restrict__ nodes,
                                                      const lr_prim* __restrict__ prims, RaysArg rays, uint32_t n,
                                                      HitsArg hits, uint32_t* __restrict__ hit_inst,
                                                      uint8_t* __restrict__ occluded, uint2* __restrict__ spill,
                                                      int cap, uint32_t* __restrict__ counter, uint32_t refill_min) {
    __shared__ uint32_t s_e[SHORT * BLOCK];
    __shared__ float s_t[SHORT * BLOCK];
    Stack<BLOCK, SHORT> st;
    st.s_e = s_e;
    st.s_t = s_t;
    st.spill = spill;
    st.stride = gridDim.x * BLOCK;
    st.tid = threadIdx.x;
    st.gtid = blockIdx.x * BLOCK + threadIdx.x;
    st.cap = cap;
    st.sp = 0;
    const float4* __restrict__ tnf = reinterpret_cast<const float4*>(tnodes);
    const float4* __restrict__ nf = reinterpret_cast<const float4*>(nodes);
    const float4* __restrict__ pf = reinterpret_cast<const float4*>(prims);
    const uint32_t lane = threadIdx.x & 63u;
    RayClaim claim;
    claim.init(n, blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6), gridDim.x * (BLOCK / 64), false);

    uint32_t ray = kNone;
    bool no_more = claim.exhausted();  // wave-uniform
    // world ray
    float wox = 0, woy = 0, woz = 0, wdx = 0, wdy = 0, wdz = 0, tnear = 0, tfar = 0;
    // current ray (world, or the current instance's local space)
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0, nl = 0, fl = 0, sc = 1.0f;
    bool local = false;
    uint32_t cur = 0, cur_id = 0;
    // nearest hit over the instances finished so far (world distance)
    float best_t = INFINITY, best_u = 0, best_v = 0;
    uint32_t best_prim = LR_MISS, best_inst = LR_MISS;
    // the current instance's candidates (local distances)
    float tri_t = 0, tri_u = 0, tri_v = 0, sph_t = 0, far_lim = 0, near_lim = 0;
    uint32_t tri_id = LR_MISS, sph_id = LR_MISS;
    bool doomed = false;

    auto to_world = [&]() {
        ox = wox;
        oy = woy;
        oz = woz;
        dx = wdx;
        dy = wdy;
        dz = wdz;
        ix = safe_inv(dx);
        iy = safe_inv(dy);
        iz = safe_inv(dz);
        nl = tnear;
        fl = tfar;
        sc = 1.0f;
        local = false;
    };
    auto enter = [&](uint32_t slot) {
        const float4* rp = reinterpret_cast<const float4*>(insts + slot);
        const float4 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3];
        // local = W x: origin ((W00 ox + W01 oy) + W02 oz) + W03, direction W d / |W d|
        ox = ((r0.x * wox + r0.y * woy) + r0.z * woz) + r0.w;
        oy = ((r1.x * wox + r1.y * woy) + r1.z * woz) + r1.w;
        oz = ((r2.x * wox + r2.y * woy) + r2.z * woz) + r2.w;
        const float x = (r0.x * wdx + r0.y * wdy) + r0.z * wdz;
        const float y = (r1.x * wdx + r1.y * wdy) + r1.z * wdz;
        const float z = (r2.x * wdx + r2.y * wdy) + r2.z * wdz;
        sc = sqrtf((x * x + y * y) + z * z);
        dx = x / sc;
        dy = y / sc;
        dz = z / sc;
        ix = safe_inv(dx);
        iy = safe_inv(dy);
        iz = safe_inv(dz);
        nl = tnear * sc;
        fl = tfar * sc;
        if (!ANY) {
            far_lim = fl * (1.0f - 1e-6f);
            near_lim = nl * (1.0f + 1e-6f);
            tri_t = far_lim;
            sph_t = far_lim;
            tri_u = tri_v = 0.0f;
            tri_id = sph_id = LR_MISS;
            doomed = false;
        }
        cur_id = __float_as_uint(r3.y);
        st.push(kMarker, -INFINITY);
        cur = __float_as_uint(r3.x);
        local = true;
    };
    auto finalize = [&]() {
        uint32_t prim = LR_MISS;
        float tl = INFINITY, u = 0.0f, v = 0.0f;
        const bool tri_ok = tri_id != LR_MISS && !doomed;
        if (tri_ok && (sph_id == LR_MISS || tri_t < sph_t || (tri_t == sph_t && tri_id < sph_id))) {
            prim = tri_id;
            tl = tri_t;
            u = tri_u;
            v = tri_v;
        } else if (sph_id != LR_MISS) {
            prim = sph_id;
            tl = sph_t;
        }
        if (prim != LR_MISS) {
            const float tw = tl / sc;
            if (tw < best_t || (tw == best_t && (cur_id < best_inst || (cur_id == best_inst && prim < best_prim)))) {
                best_t = tw;
                best_prim = prim;
                best_inst = cur_id;
                best_u = u;
                best_v = v;
            }
        }
    };
    auto cull_now = [&]() -> float {
        if (ANY) return fl * kCullSlack;
        if (local) return fminf(fminf(best_t * sc, doomed ? INFINITY : tri_t), sph_t) * kCullSlack;
        return fminf(best_t, tfar) * kCullSlack;
    };

    for (;;) {
        if (!no_more) {
            const uint64_t need = __ballot(ray == kNone);
            const uint32_t idle = (uint32_t)__popcll(need);
            if (idle && (idle >= refill_min || idle == 64u)) {
                const uint32_t my = claim.claim(need, idle, lane, ray == kNone, counter);
                if (my != kNone) {
                    ray = my;
                    wox = rays.ox[my];
                    woy = rays.oy[my];
                    woz = rays.oz[my];
                    wdx = rays.dx[my];
                    wdy = rays.dy[my];
                    wdz = rays.dz[my];
                    tnear = rays.tnear[my];
                    tfar = rays.tfar[my];
                    to_world();
                    best_t = INFINITY;
                    best_u = best_v = 0.0f;
                    best_prim = best_inst = LR_MISS;
                    cur = 0u;
                    st.sp = 0;
                }
                no_more = claim.exhausted();
            }
        }
        if (__ballot(ray != kNone) == 0) break;
        if (ray == kNone) continue;

        bool finished = false;
        bool need_pop = true;
        bool any_hit = false;
        if (!(cur & 0x80000000u)) {
            float t[4];
            uint32_t e[4];
            node_test((local ? nf : tnf) + (size_t)cur * kNodeF4, ox, oy, oz, ix, iy, iz, nl, cull_now(), t, e);
            LR_CSWAP(0, 1);
            LR_CSWAP(2, 3);
            LR_CSWAP(0, 2);
            LR_CSWAP(1, 3);
            LR_CSWAP(1, 2);
            if (t[0] != INFINITY) {
                if (t[3] != INFINITY) st.push(e[3], t[3]);
                if (t[2] != INFINITY) st.push(e[2], t[2]);
                if (t[1] != INFINITY) st.push(e[1], t[1]);
                cur = e[0];
                need_pop = false;
            }
        } else if (!local) {
            // top-level leaf: its instance slots, the first entered now
            const uint32_t firsti = (cur >> 2) & 0x1FFFFFFFu;
            const uint32_t cnt = (cur & 3u) + 1u;
            for (uint32_t k = cnt - 1; k >= 1; --k) st.push(kEnter | (firsti + k), -INFINITY);
            enter(firsti);
            need_pop = false;
        } else {
            const uint32_t firstp = (cur >> 2) & 0x1FFFFFFFu;
            const uint32_t cnt = (cur & 3u) + 1u;
            const float4* q = pf + (size_t)firstp * 3;
            auto test = [&](const float4 a, const float4 b, const float4 c) {
                const uint32_t id = __float_as_uint(a.w);
                if (__float_as_uint(b.w) == LR_PRIM_TRIANGLE) {
                    float th, uh, vh;
                    if (lr::tri_intersect(ox, oy, oz, dx, dy, dz, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, th,
                                          uh, vh)) {
                        if (ANY) {
                            if (th > nl && th < fl) any_hit = true;
                        } else if (th > nl && th < far_lim && (th < tri_t || (th == tri_t && id < tri_id))) {
                            tri_t = th;
                            tri_u = uh;
                            tri_v = vh;
                            tri_id = id;
                            doomed = !(th > near_lim);
                        }
                    }
                } else {
                    float th;
                    if (lr::sphere_intersect(ox, oy, oz, dx, dy, dz, a.x, a.y, a.z, b.y, nl, th) &&
                        lr::in_range(th, nl, fl)) {
                        if (ANY) {
                            any_hit = true;
                        } else if (th < sph_t || (th == sph_t && id < sph_id)) {
                            sph_t = th;
                            sph_id = id;
                        }
                    }
                }
            };
            // a leaf's records (48 B each) are requested before any is tested:
            // the whole leaf for any-hit, pairs for closest-hit
            if constexpr (ANY) {
                float4 ra[4], rb[4], rc[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if ((uint32_t)k < cnt) {
                        ra[k] = q[3 * k];
                        rb[k] = q[3 * k + 1];
                        rc[k] = q[3 * k + 2];
                    }
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if ((uint32_t)k < cnt && !any_hit) test(ra[k], rb[k], rc[k]);
            } else {
                for (uint32_t k0 = 0; k0 < cnt; k0 += 2) {
                    const float4 ra0 = q[3 * k0], rb0 = q[3 * k0 + 1], rc0 = q[3 * k0 + 2];
                    float4 ra1 = ra0, rb1 = rb0, rc1 = rc0;
                    if (k0 + 1 < cnt) {
                        ra1 = q[3 * k0 + 3];
                        rb1 = q[3 * k0 + 4];
                        rc1 = q[3 * k0 + 5];
                    }
                    test(ra0, rb0, rc0);
                    if (k0 + 1 < cnt) test(ra1, rb1, rc1);
                }
            }
            if (ANY && any_hit) {
                finished = true;
                need_pop = false;
            }
        }
        if (need_pop) {
            finished = true;
            while (st.sp > 0) {
                uint32_t e;
                float te;
                st.pop(e, te);
                if (e == kMarker) {  // the current instance is done
                    if (!ANY) finalize();
                    to_world();
                    continue;
                }
                if (!local && (e & 0xC0000000u) == kEnter) {
                    enter(e & 0x3FFFFFFFu);
                    finished = false;
                    break;
                }
                if (te <= cull_now()) {
                    cur = e;
                    finished = false;
                    break;
                }
            }
        }
        if (finished) {
            if (ANY) {
                occluded[ray] = any_hit ? 1 : 0;
            } else {
                hits.prim[ray] = best_prim;
                hits.t[ray] = best_prim == LR_MISS ? INFINITY : best_t;
                hits.u[ray] = best_u;
                hits.v[ray] = best_v;
                hit_inst[ray] = best_inst;
            }
            ray = kNone;
        }
    }
}

__global__ void k_zero_counter(uint32_t* p) { p[threadIdx.x] = 0u; }

__global__ void k_fill_miss_inst(uint32_t n, HitsArg h, uint32_t* inst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        h.prim[i] = LR_MISS;
        h.t[i] = INFINITY;
        h.u[i] = 0.0f;
        h.v[i] = 0.0f;
        inst[i] = LR_MISS;
    }
}

__global__ void k_fill_zero_u8(uint32_t n, uint8_t* p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void free_tlas(lr_tlas* T) {
    if (!T) return;
    (void)hipSetDevice(T->ctx->device);
    void* ptrs[] = {T->d_tnodes, T->d_inst, T->d_nodes, T->d_prims};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete T;
}

// 3x4 affine inverse in double; false when singular.
bool affine_inverse(const float* m, double* w) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    const double A = e * i - f * h, B = f * g - d * i, Cc = d * h - e * g;
    const double det = a * A + b * B + c * Cc;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    const double inv[9] = {A / det, (c * h - b * i) / det, (b * f - c * e) / det,
                           B / det, (a * i - c * g) / det, (c * d - a * f) / det,
                           Cc / det, (b * g - a * h) / det, (a * e - b * d) / det};
    const double t[3] = {m[3], m[7], m[11]};
    for (int r = 0; r < 3; ++r) {
        w[4 * r + 0] = inv[3 * r + 0];
        w[4 * r + 1] = inv[3 * r + 1];
        w[4 * r + 2] = inv[3 * r + 2];
        w[4 * r + 3] = -(inv[3 * r] * t[0] + inv[3 * r + 1] * t[1] + inv[3 * r + 2] * t[2]);
    }
    return true;
}

template <bool ANY>
int launch_inst(lr_tlas* T, const lr_rays_soa* rays, uint32_t n, HitsArg ha, uint32_t* inst, uint8_t* occ,
                hipStream_t s) {
    if (n > kMaxRays) return lr::fail(LR_E_INVALID, "at most 4,278,190,080 rays per launch");
    lr_context* ctx = T->ctx;
    const auto fn = k_trace_inst<kBlock, kShort, ANY>;
    lr_trace_config cfg;
    int occ_blocks = 0;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        cfg = ctx->cfg;
        auto it = ctx->occupancy.find(reinterpret_cast<const void*>(fn));
        if (it != ctx->occupancy.end()) occ_blocks = it->second;
    }
    if (occ_blocks == 0) {
        LR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_blocks, reinterpret_cast<const void*>(fn),
                                                                  kBlock, 0));
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->occupancy[reinterpret_cast<const void*>(fn)] = occ_blocks;
    }
    occ_blocks = std::max(1, std::min(occ_blocks, (int)cfg.blocks_per_cu));
    const uint64_t want = ((uint64_t)n + kBlock - 1) / kBlock;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)ctx->n_cu * occ_blocks));
    const int cap = (int)T->stack_cap;
    size_t bytes = 256;
    if (cap > kShort) bytes += (size_t)(cap - kShort) * grid * kBlock * sizeof(uint2);
    void* p = nullptr;
    if (int rc = ctx->scratch(s, bytes, &p)) return rc;
    uint32_t* counter = static_cast<uint32_t*>(p);
    uint2* spill = cap > kShort ? reinterpret_cast<uint2*>(static_cast<char*>(p) + 256) : nullptr;
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, nullptr};
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, counter);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, T->d_tnodes, T->d_inst, T->d_nodes, T->d_prims, ra, n, ha,
                       inst, occ, spill, cap, counter, cfg.refill_min);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

int check_rays_inst(const lr_rays_soa* rays, uint32_t n) {
    if (!rays) return lr::fail(LR_E_INVALID, "rays is NULL");
    if (n && (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->tnear ||
              !rays->tfar))
        return lr::fail(LR_E_INVALID, "a ray array pointer is NULL");
    return LR_OK;
}

}  // namespace

extern "C" {

int lr_tlas_create(lr_context* ctx, lr_qbvh* const* geometry, uint32_t n_geometry, const lr_instance* instances,
                   uint32_t n_instances, lr_tlas** out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !out) return lr::fail(LR_E_INVALID, "need a context and out");
    *out = nullptr;
    if (n_geometry && !geometry) return lr::fail(LR_E_INVALID, "geometry is NULL");
    if (n_instances && !instances) return lr::fail(LR_E_INVALID, "instances is NULL");
    if (n_instances >= (1u << 29)) return lr::fail(LR_E_INVALID, "too many instances (limit 2^29 - 1)");
    uint64_t nodes = 0, prims = 0;
    uint32_t geo_stack = 1;
    for (uint32_t g = 0; g < n_geometry; ++g) {
        const lr_qbvh* q = geometry[g];
        if (!q) return lr::fail(LR_E_INVALID, "geometry " + std::to_string(g) + " is NULL");
        if (q->ctx != ctx) return lr::fail(LR_E_INVALID, "geometry " + std::to_string(g) + " belongs to another context");
        nodes += q->host.nodes.size();
        prims += q->host.prims.size();
        geo_stack = std::max(geo_stack, q->info.max_stack);
    }
    if (nodes >= (1ull << 30) || prims >= (1ull << 29))
        return lr::fail(LR_E_INVALID, "geometry too large for the node encoding (2^30 nodes, 2^29 primitives)");
    // concatenated geometry trees (inner children += node base, leaf first += prim base)
    std::vector<lr_node> hn;
    std::vector<lr_prim> hp;
    std::vector<uint32_t> roots(n_geometry, LR_CHILD_EMPTY);
    hn.reserve(nodes);
    hp.reserve(prims);
    for (uint32_t g = 0; g < n_geometry; ++g) {
        const lr::BuiltQbvh& H = geometry[g]->host;
        if (H.nodes.empty()) continue;
        const uint32_t nb = (uint32_t)hn.size(), pb = (uint32_t)hp.size();
        roots[g] = nb;
        for (const lr_node& nd : H.nodes) {
            lr_node m = nd;
            for (int c = 0; c < 4; ++c) {
                const uint32_t e = nd.child[c];
                if (e == LR_CHILD_EMPTY) continue;
                if (e & 0x80000000u) {
                    const uint32_t f = ((e >> 2) & 0x1FFFFFFFu) + pb;
                    m.child[c] = 0x80000000u | (f << 2) | (e & 3u);
                } else {
                    m.child[c] = e + nb;
                }
            }
            hn.push_back(m);
        }
        hp.insert(hp.end(), H.prims.begin(), H.prims.end());
    }
    // instances with geometry: world -> local transforms and world boxes
    std::vector<InstRec> recs;
    std::vector<float> boxes;
    for (uint32_t k = 0; k < n_instances; ++k) {
        const lr_instance& I = instances[k];
        if (I.geometry >= n_geometry)
            return lr::fail(LR_E_INVALID, "instance " + std::to_string(k) + " references geometry " +
                                              std::to_string(I.geometry) + " >= " + std::to_string(n_geometry));
        for (int c = 0; c < 12; ++c)
            if (!std::isfinite(I.local_to_world[c]))
                return lr::fail(LR_E_INVALID, "instance " + std::to_string(k) + " has a non-finite transform");
        double w[12];
        if (!affine_inverse(I.local_to_world, w))
            return lr::fail(LR_E_INVALID, "instance " + std::to_string(k) + " has a singular transform");
        if (roots[I.geometry] == LR_CHILD_EMPTY) continue;  // empty geometry: nothing to hit
        InstRec r;
        std::memset(&r, 0, sizeof(r));
        for (int c = 0; c < 12; ++c) r.w[c] = (float)w[c];
        r.root = roots[I.geometry];
        r.id = k;
        recs.push_back(r);
        const lr_qbvh_info& gi = geometry[I.geometry]->info;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int corner = 0; corner < 8; ++corner) {
            const double p[3] = {(corner & 1) ? gi.scene_hi[0] : gi.scene_lo[0],
                                 (corner & 2) ? gi.scene_hi[1] : gi.scene_lo[1],
                                 (corner & 4) ? gi.scene_hi[2] : gi.scene_lo[2]};
            for (int a = 0; a < 3; ++a) {
                const float* m = I.local_to_world + 4 * a;
                const double v = (double)m[0] * p[0] + (double)m[1] * p[1] + (double)m[2] * p[2] + (double)m[3];
                lo[a] = std::min(lo[a], v);
                hi[a] = std::max(hi[a], v);
            }
        }
        for (int a = 0; a < 3; ++a) boxes.push_back(std::nextafter((float)lo[a], -INFINITY));
        for (int a = 0; a < 3; ++a) boxes.push_back(std::nextafter((float)hi[a], INFINITY));
    }
    lr::BuiltQbvh top;
    std::vector<uint32_t> order;
    std::string err;
    if (lr::build_boxes(boxes.data(), (uint32_t)recs.size(), top, order, err))
        return lr::fail(LR_E_INVALID, "lr_tlas_create: " + err);
    std::vector<InstRec> leaf_recs(recs.size());
    for (size_t s = 0; s < order.size(); ++s) leaf_recs[s] = recs[order[s]];
    if (int rc = lr::bind(ctx)) return rc;
    lr_tlas* T = new (std::nothrow) lr_tlas();
    if (!T) return lr::fail(LR_E_NOMEM, "out of host memory");
    T->ctx = ctx;
    T->info.n_instances = (uint32_t)recs.size();
    T->info.n_nodes = (uint32_t)top.nodes.size();
    T->info.max_depth = top.max_depth;
    const uint32_t top_stack = 3 * (top.max_depth + 1) + 1;
    T->stack_cap = top_stack + 3 + 1 + geo_stack;
    T->info.max_stack = T->stack_cap;
    T->info.n_geometry_nodes = hn.size();
    T->info.n_geometry_prims = hp.size();
    for (int a = 0; a < 3; ++a) {
        T->info.scene_lo[a] = top.scene_lo[a];
        T->info.scene_hi[a] = top.scene_hi[a];
    }
    hipError_t e = hipSuccess;
    if (!top.nodes.empty()) {
        const std::vector<lr_node_q8> tq = lr::quantize_nodes(top.nodes), gq = lr::quantize_nodes(hn);
        e = hipMalloc(&T->d_tnodes, sizeof(lr_node_q8) * tq.size());
        if (e == hipSuccess) e = hipMalloc(&T->d_inst, sizeof(InstRec) * leaf_recs.size());
        if (e == hipSuccess) e = hipMalloc(&T->d_nodes, sizeof(lr_node_q8) * gq.size());
        if (e == hipSuccess) e = hipMalloc(&T->d_prims, sizeof(lr_prim) * hp.size());
        if (e == hipSuccess)
            e = hipMemcpy(T->d_tnodes, tq.data(), sizeof(lr_node_q8) * tq.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(T->d_inst, leaf_recs.data(), sizeof(InstRec) * leaf_recs.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(T->d_nodes, gq.data(), sizeof(lr_node_q8) * gq.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(T->d_prims, hp.data(), sizeof(lr_prim) * hp.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        free_tlas(T);
        return lr::fail(LR_E_DEVICE, std::string("lr_tlas_create: ") + hipGetErrorString(e));
    }
    T->info.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = T;
    return LR_OK;
}

int lr_tlas_destroy(lr_tlas* tlas) {
    free_tlas(tlas);
    return LR_OK;
}

int lr_tlas_get_info(const lr_tlas* tlas, lr_tlas_info* info) {
    if (!tlas || !info) return lr::fail(LR_E_INVALID, "NULL argument");
    *info = tlas->info;
    return LR_OK;
}

int lr_trace_closest_instanced(lr_tlas* T, const lr_rays_soa* rays, uint32_t n, lr_hits_soa* hits, uint32_t* instance,
                               void* stream) {
    if (!T) return lr::fail(LR_E_INVALID, "NULL tlas");
    if (int rc = check_rays_inst(rays, n)) return rc;
    if (!hits || !instance) return lr::fail(LR_E_INVALID, "NULL hits or instance array");
    if (n && (!hits->prim || !hits->t || !hits->u || !hits->v)) return lr::fail(LR_E_INVALID, "a hit array pointer is NULL");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(T->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const HitsArg ha{hits->prim, hits->t, hits->u, hits->v};
    if (T->info.n_instances == 0) {
        hipLaunchKernelGGL(k_fill_miss_inst, dim3((n + 255) / 256), dim3(256), 0, s, n, ha, instance);
        LR_HIP_CHECK(hipGetLastError());
        return LR_OK;
    }
    return launch_inst<false>(T, rays, n, ha, instance, nullptr, s);
}

int lr_trace_any_instanced(lr_tlas* T, const lr_rays_soa* rays, uint32_t n, uint8_t* occluded, void* stream) {
    if (!T) return lr::fail(LR_E_INVALID, "NULL tlas");
    if (int rc = check_rays_inst(rays, n)) return rc;
    if (!occluded) return lr::fail(LR_E_INVALID, "NULL occluded array");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(T->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (T->info.n_instances == 0) {
        hipLaunchKernelGGL(k_fill_zero_u8, dim3((n + 255) / 256), dim3(256), 0, s, n, occluded);
        LR_HIP_CHECK(hipGetLastError());
        return LR_OK;
    }
    const HitsArg none{nullptr, nullptr, nullptr, nullptr};
    return launch_inst<true>(T, rays, n, none, nullptr, occluded, s);
}

}  // extern "C"
