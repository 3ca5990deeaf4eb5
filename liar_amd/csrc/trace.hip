// trace.hip -- gfx950 QBVH closest-hit / any-hit traversal and the C-ABI for
// contexts, QBVH handles and ray batches.
//
// Execution model: persistent wave64 kernel, one ray per lane, grid sized to
// blocks_per_cu x #CUs.  Each lane walks the 4-wide QBVH front to back with a
// short stack of (entry, entry distance) pairs in LDS; pushes past the LDS
// depth spill to a per-thread global region, so any tree depth is handled.
//
// Closest-hit keeps two candidates, following LiAR's per-object filtering
// (src/liar/scenery/list.cpp:106-133): the nearest triangle with t > tnear
// (src/liar/scenery/triangle_mesh.cpp:291-317, whose interval check is
// applied once at the end) and the nearest in-range sphere
// (src/liar/scenery/sphere.cpp:104-118).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "lr_geom.h"
#include "lr_internal.h"
#include "lr_traverse.h"

namespace {

using namespace lr::trav;

// Lane-refill persistent traversal.  Every lane runs one traversal step per
// iteration (a node visit or a leaf's primitive tests); lanes whose rays are
// finished take the next ray indices from their wave's reservation (a static
// share, then chunks claimed from a global counter: RayClaim), so no lane
// idles while rays remain.
template <int BLOCK, int SHORT, bool ANY, bool PHASED>
__global__ __launch_bounds__(BLOCK) void k_trace(const lr_node_q8* __restrict__ nodes,
                                                 const lr_prim* __restrict__ prims, RaysArg rays, uint32_t n,
                                                 HitsArg hits, uint8_t* __restrict__ occluded,
                                                 uint2* __restrict__ spill, int cap,
                                                 uint32_t* __restrict__ counter, uint32_t refill_min,
                                                 const uint32_t* __restrict__ n_dev) {
    if (n_dev) n = min(*n_dev, n);  // queue length produced on the device (n: its capacity)
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
    const float4* __restrict__ nf = reinterpret_cast<const float4*>(nodes);
    const float4* __restrict__ pf = reinterpret_cast<const float4*>(prims);
    const uint32_t lane = threadIdx.x & 63u;

    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t ray = kNone;
    RayClaim claim;
    claim.init(n, blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6), gridDim.x * (BLOCK / 64));
    bool no_more = claim.exhausted();  // wave-uniform
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0;
    float tnear = 0, tfar = 0, far_lim = 0, near_lim = 0;
    float tri_t = 0, tri_u = 0, tri_v = 0, sph_t = 0;
    uint32_t tri_id = LR_MISS, sph_id = LR_MISS, cur = 0;
    bool doomed = false;
    // cur: the next entry to visit (inner node, leaf, or kEmpty); leaf: the
    // postponed leaf (or kEmpty); cur_t / leaf_t: their entry distances.
    constexpr uint32_t kEmpty = LR_CHILD_EMPTY;
    uint32_t leaf = kEmpty;
    float cur_t = 0.0f, leaf_t = 0.0f;

    for (;;) {
        if (!no_more) {
            const uint64_t need = __ballot(ray == kNone);
            const uint32_t idle = (uint32_t)__popcll(need);
            // Refill only when enough lanes are idle (or all are): each refill
            // costs the wave a ray-load round trip.
            if (idle && (idle >= refill_min || idle == 64u)) {
                const uint32_t my = claim.claim(need, idle, lane, ray == kNone, counter);
                if (my != kNone) {
                    ray = my;
                    ox = rays.ox[my];
                    oy = rays.oy[my];
                    oz = rays.oz[my];
                    dx = rays.dx[my];
                    dy = rays.dy[my];
                    dz = rays.dz[my];
                    tnear = rays.tnear[my];
                    tfar = rays.tfar[my];
                    ix = safe_inv(dx);
                    iy = safe_inv(dy);
                    iz = safe_inv(dz);
                    far_lim = tfar * (1.0f - 1e-6f);
                    near_lim = tnear * (1.0f + 1e-6f);
                    tri_t = far_lim;
                    sph_t = far_lim;
                    tri_u = tri_v = 0.0f;
                    tri_id = sph_id = LR_MISS;
                    doomed = false;
                    cur = rays.root ? rays.root[my] : 0u;
                    cur_t = tnear;
                    leaf = kEmpty;
                    st.sp = 0;
                }
                no_more = claim.exhausted();
            }
        }
        if (__ballot(ray != kNone) == 0) break;
        // STEP: each lane visits its node or its leaf every iteration (the two
        // code paths share the iteration).  PHASED: the phase of an iteration
        // is the same for the whole wave -- node steps while any lane is still
        // searching for a leaf (no leaf postponed and an inner node to visit),
        // else leaf steps; a lane that reaches a leaf postpones it and visits
        // nodes on (speculatively) until every lane has one, so each phase runs
        // with most lanes active.
        bool leaf_phase = false;
        if constexpr (PHASED) {
            const bool searching = ray != kNone && leaf == kEmpty && !(cur & 0x80000000u);
            leaf_phase = __ballot(searching) == 0;
        }
        if (ray == kNone) continue;

        bool finished = false;
        bool any_hit = false;
        const float cull = ANY ? tfar * kCullSlack : fminf(doomed ? INFINITY : tri_t, sph_t) * kCullSlack;
        // Visits inner node `cur`: pushes the far children, returns whether any
        // child was hit (then cur = the nearest, its entry distance in t0).
        auto visit_node = [&](float& t0) -> bool {
            float t[4];
            uint32_t e[4];
            node_test(nf + (size_t)cur * kNodeF4, ox, oy, oz, ix, iy, iz, tnear, cull, t, e);
            LR_CSWAP(0, 1);
            LR_CSWAP(2, 3);
            LR_CSWAP(0, 2);
            LR_CSWAP(1, 3);
            LR_CSWAP(1, 2);
            if (t[0] == INFINITY) return false;
            if (t[3] != INFINITY) st.push(e[3], t[3]);
            if (t[2] != INFINITY) st.push(e[2], t[2]);
            if (t[1] != INFINITY) st.push(e[1], t[1]);
            cur = e[0];
            t0 = t[0];
            return true;
        };
        // Tests the primitives of leaf entry `lf`.  The records are requested
        // before they are tested (one memory round trip per group instead of
        // two per primitive): any-hit asks for all four at once; closest-hit,
        // which keeps more state live, two at a time.
        auto visit_leaf = [&](uint32_t lf) {
            const uint32_t first = (lf >> 2) & 0x1FFFFFFFu;
            const uint32_t cnt = (lf & 3u) + 1u;
            const float4* q = pf + (size_t)first * 3;
            auto test = [&](const float4 a, const float4 b, const float4 c) {
                const uint32_t id = __float_as_uint(a.w);
                if (__float_as_uint(b.w) == LR_PRIM_TRIANGLE) {
                    float th, uh, vh;
                    if (lr::tri_intersect(ox, oy, oz, dx, dy, dz, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, th,
                                          uh, vh)) {
                        if (ANY) {
                            if (th > tnear && th < tfar) any_hit = true;
                        } else if (th > tnear && th < far_lim && (th < tri_t || (th == tri_t && id < tri_id))) {
                            tri_t = th;
                            tri_u = uh;
                            tri_v = vh;
                            tri_id = id;
                            doomed = !(th > near_lim);
                        }
                    }
                } else {
                    float th;
                    if (lr::sphere_intersect(ox, oy, oz, dx, dy, dz, a.x, a.y, a.z, b.y, tnear, th) &&
                        lr::in_range(th, tnear, tfar)) {
                        if (ANY) {
                            any_hit = true;
                        } else if (th < sph_t || (th == sph_t && id < sph_id)) {
                            sph_t = th;
                            sph_id = id;
                        }
                    }
                }
            };
            // Slots past the leaf's count re-read its last record (same
            // address, no extra memory traffic) so that every load is
            // unconditional.
            if constexpr (ANY) {
                float4 ra[4], rb[4], rc[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t r = 3u * min((uint32_t)k, cnt - 1u);
                    ra[k] = q[r];
                    rb[k] = q[r + 1];
                    rc[k] = q[r + 2];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    keep(ra[k]);
                    keep(rb[k]);
                    keep(rc[k]);
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if ((uint32_t)k < cnt && !any_hit) test(ra[k], rb[k], rc[k]);
            } else {
                for (uint32_t k0 = 0; k0 < cnt; k0 += 2) {
                    const uint32_t r1 = 3u * min(k0 + 1u, cnt - 1u);
                    const float4 ra0 = q[3 * k0], rb0 = q[3 * k0 + 1], rc0 = q[3 * k0 + 2];
                    const float4 ra1 = q[r1], rb1 = q[r1 + 1], rc1 = q[r1 + 2];
                    keep(ra0);
                    keep(rb0);
                    keep(rc0);
                    keep(ra1);
                    keep(rb1);
                    keep(rc1);
                    test(ra0, rb0, rc0);
                    if (k0 + 1 < cnt) test(ra1, rb1, rc1);
                }
            }
        };
        // Pops the next entry that can still hold a better hit (false: none).
        auto pop_next = [&](uint32_t& e_out, float& t_out) -> bool {
            const float cull2 = ANY ? cull : fminf(doomed ? INFINITY : tri_t, sph_t) * kCullSlack;
            while (st.sp > 0) {
                uint32_t e;
                float te;
                st.pop(e, te);
                if (te <= cull2) {
                    e_out = e;
                    t_out = te;
                    return true;
                }
            }
            return false;
        };

        if constexpr (!PHASED) {
            bool need_pop = true;
            if (!(cur & 0x80000000u)) {
                float t0;
                need_pop = !visit_node(t0);
            } else {
                visit_leaf(cur);
                if (ANY && any_hit) {
                    finished = true;
                    need_pop = false;
                }
            }
            if (need_pop) {
                float te;
                finished = !pop_next(cur, te);
            }
        } else {
            if (!leaf_phase) {
                if (!(cur & 0x80000000u) && !visit_node(cur_t)) cur = kEmpty;  // inner node (kEmpty has bit 31 set)
            } else if (leaf != kEmpty) {
                // a leaf postponed behind a hit found since is skipped
                if (leaf_t <= cull) visit_leaf(leaf);
                leaf = kEmpty;
                if (ANY && any_hit) finished = true;
            }
            if (!finished) {
                // when there is no entry to visit, pop one; a leaf moves to the
                // free leaf slot and the next entry is popped at once
                if (cur == kEmpty && !pop_next(cur, cur_t)) cur = kEmpty;
                if (cur != kEmpty && (cur & 0x80000000u) && leaf == kEmpty) {
                    leaf = cur;
                    leaf_t = cur_t;
                    if (!pop_next(cur, cur_t)) cur = kEmpty;
                }
                finished = cur == kEmpty && leaf == kEmpty;
            }
        }
        if (finished) {
            if (ANY) {
                occluded[ray] = any_hit ? 1 : 0;
            } else {
                uint32_t prim = LR_MISS;
                float t = INFINITY, u = 0.0f, v = 0.0f;
                const bool tri_ok = tri_id != LR_MISS && !doomed;
                if (tri_ok && (sph_id == LR_MISS || tri_t < sph_t || (tri_t == sph_t && tri_id < sph_id))) {
                    prim = tri_id;
                    t = tri_t;
                    u = tri_u;
                    v = tri_v;
                } else if (sph_id != LR_MISS) {
                    prim = sph_id;
                    t = sph_t;
                }
                hits.prim[ray] = prim;
                hits.t[ray] = t;
                hits.u[ray] = u;
                hits.v[ray] = v;
            }
            ray = kNone;
        }
    }
}
__global__ void k_fill_miss(uint32_t n, HitsArg hits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        hits.prim[i] = LR_MISS;
        hits.t[i] = INFINITY;
        hits.u[i] = 0.0f;
        hits.v[i] = 0.0f;
    }
}

__global__ void k_fill_u8(uint32_t n, uint8_t* p, uint8_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_u32(uint32_t* p) { p[threadIdx.x] = 0u; }

using TraceFn = void (*)(const lr_node_q8*, const lr_prim*, RaysArg, uint32_t, HitsArg, uint8_t*, uint2*, int,
                         uint32_t*, uint32_t, const uint32_t*);

struct KernelEntry {
    uint32_t block, lds_stack;
    TraceFn closest[2];  // [schedule - 1]: LR_TRACE_SCHEDULE_STEP, LR_TRACE_SCHEDULE_PHASED
    TraceFn any[2];
};

#define LR_K(B, S) \
    {B, S, {k_trace<B, S, false, false>, k_trace<B, S, false, true>}, {k_trace<B, S, true, false>, k_trace<B, S, true, true>}}
const KernelEntry kKernels[] = {
    LR_K(64, 8),  LR_K(64, 16),  LR_K(64, 24),  LR_K(128, 8), LR_K(128, 16),
    LR_K(128, 24), LR_K(256, 8), LR_K(256, 16), LR_K(256, 24),
};
#undef LR_K

const KernelEntry* find_kernel(uint32_t block, uint32_t lds_stack) {
    for (const auto& k : kKernels)
        if (k.block == block && k.lds_stack == lds_stack) return &k;
    return nullptr;
}

int check_rays(const lr_rays_soa* rays, uint32_t n) {
    if (!rays) return lr::fail(LR_E_INVALID, "rays is NULL");
    if (n && (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->tnear ||
              !rays->tfar))
        return lr::fail(LR_E_INVALID, "a ray array pointer is NULL");
    return LR_OK;
}

struct Launch {
    TraceFn fn;
    uint32_t block;
    uint32_t grid;
    int cap;
    uint2* spill;
    uint32_t* counter;
    uint32_t refill_min;
};

// Scratch layout per stream: [work counter, 256 B][spill stack: (cap - lds) x threads x 8 B].
constexpr size_t kCounterBytes = 256;

// Chooses the kernel and the persistent grid of a launch over n rays; the
// scratch pointers are bound by the caller (prepare, prepare_fixed).
// cache_occupancy = false: a missing occupancy entry is queried but not
// stored (the map insertion would allocate host memory).
int plan(lr_qbvh* bvh, uint32_t n, bool any, Launch& L, uint32_t& lds_stack, bool cache_occupancy = true) {
    lr_context* ctx = bvh->ctx;
    lr_trace_config cfg;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        cfg = ctx->cfg;
    }
    const KernelEntry* k = find_kernel(cfg.block, cfg.lds_stack);
    if (!k) return lr::fail(LR_E_INVALID, "no traversal kernel for this block / lds_stack");
    L.fn = any ? k->any[cfg.schedule - 1] : k->closest[cfg.schedule - 1];
    L.block = cfg.block;
    L.refill_min = cfg.refill_min;
    int occ = 0;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->occupancy.find(reinterpret_cast<const void*>(L.fn));
        if (it != ctx->occupancy.end()) occ = it->second;
    }
    if (occ == 0) {
        LR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(L.fn),
                                                                  (int)cfg.block, 0));
        if (cache_occupancy) {
            std::lock_guard<std::mutex> g(ctx->mu);
            ctx->occupancy[reinterpret_cast<const void*>(L.fn)] = occ;
        }
    }
    occ = std::max(1, std::min(occ, (int)cfg.blocks_per_cu));
    const uint64_t want = ((uint64_t)n + cfg.block - 1) / cfg.block;
    const uint64_t maxg = (uint64_t)ctx->n_cu * (uint64_t)occ;
    L.grid = (uint32_t)std::max<uint64_t>(1, std::min(want, maxg));
    L.cap = (int)bvh->info.max_stack;
    lds_stack = cfg.lds_stack;
    return LR_OK;
}

// Bytes of scratch a launch needs: the work counter, then the spilled stack
// entries below the LDS part, (cap - lds_stack) x grid x block x 8 B.
size_t scratch_bytes(const Launch& L, uint32_t lds_stack) {
    size_t bytes = kCounterBytes;
    if ((uint32_t)L.cap > lds_stack) bytes += (size_t)(L.cap - lds_stack) * L.grid * L.block * sizeof(uint2);
    return bytes;
}

void bind_scratch(Launch& L, uint32_t lds_stack, void* p) {
    L.counter = static_cast<uint32_t*>(p);
    L.spill = (uint32_t)L.cap > lds_stack ? reinterpret_cast<uint2*>(static_cast<char*>(p) + kCounterBytes) : nullptr;
}

int prepare(lr_qbvh* bvh, uint32_t n, bool any, hipStream_t stream, Launch& L) {
    uint32_t lds = 0;
    if (int rc = plan(bvh, n, any, L, lds)) return rc;
    void* p = nullptr;
    if (int rc = bvh->ctx->scratch(stream, scratch_bytes(L, lds), &p)) return rc;
    bind_scratch(L, lds, p);
    return LR_OK;
}

// The fixed scratch of lr_trace_closest_noalloc: one array per GPU, alive
// until the process ends.  Host code never uses this name as an address:
// lr_context_create looks up the device address of the current GPU's copy
// (lr_context::fixed_scratch) and launches receive that pointer.
constexpr size_t kFixedScratchBytes = 32u << 20;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

// As prepare, on the context's fixed scratch: the persistent grid shrinks
// until its spill slices fit (the kernel's spill stride is its own grid), and
// a tree too deep for even one workgroup's slice is rejected.
int prepare_fixed(lr_qbvh* bvh, uint32_t n, Launch& L) {
    lr_context* ctx = bvh->ctx;
    if (!ctx->fixed_scratch) return lr::fail(LR_E_DEVICE, "fixed traversal scratch was not set up for this context");
    uint32_t lds = 0;
    if (int rc = plan(bvh, n, false, L, lds, false)) return rc;
    if ((uint32_t)L.cap > lds) {
        const size_t per_block = (size_t)(L.cap - lds) * L.block * sizeof(uint2);
        const size_t fit = (kFixedScratchBytes - kCounterBytes) / per_block;
        if (fit == 0)
            return lr::fail(LR_E_INVALID, "lr_trace_closest_noalloc: a traversal stack of " + std::to_string(L.cap) +
                                              " entries needs more than the fixed " +
                                              std::to_string(kFixedScratchBytes >> 20) + " MiB scratch holds");
        L.grid = (uint32_t)std::min<size_t>(L.grid, fit);
    }
    if (scratch_bytes(L, lds) > kFixedScratchBytes)
        return lr::fail(LR_E_DEVICE, "lr_trace_closest_noalloc: launch does not fit the fixed scratch");
    bind_scratch(L, lds, ctx->fixed_scratch);
    return LR_OK;
}

int launch(lr_qbvh* bvh, const Launch& L, const RaysArg& ra, uint32_t n, const HitsArg& ha, uint8_t* occ,
           hipStream_t s, const uint32_t* n_dev = nullptr) {
    // the claim counter runs past n by at most 1024 per wave
    if (n > kMaxRays) return lr::fail(LR_E_INVALID, "at most 4,278,190,080 rays per launch");
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, L.counter);
    hipLaunchKernelGGL(L.fn, dim3(L.grid), dim3(L.block), 0, s, bvh->d_nodes, bvh->d_prims, ra, n, ha, occ,
                       L.spill, L.cap, L.counter, L.refill_min, n_dev);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

thread_local std::string g_last_error;

}  // namespace

namespace lr {

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int bind(const lr_context* ctx) {
    LR_HIP_CHECK(hipSetDevice(ctx->device));
    return LR_OK;
}

}  // namespace lr

namespace lr {

int trace_closest_queue(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n_max, const uint32_t* n_dev,
                        lr_hits_soa* hits, hipStream_t s, const uint32_t* root) {
    if (n_max == 0 || bvh->info.n_nodes == 0) return LR_OK;  // callers treat an empty scene separately
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, root};
    const HitsArg ha{hits->prim, hits->t, hits->u, hits->v};
    Launch L;
    if (int rc = prepare(bvh, n_max, false, s, L)) return rc;
    return launch(bvh, L, ra, n_max, ha, nullptr, s, n_dev);
}

int trace_any_queue(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n_max, const uint32_t* n_dev, uint8_t* occluded,
                    hipStream_t s) {
    if (n_max == 0 || bvh->info.n_nodes == 0) return LR_OK;
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, nullptr};
    const HitsArg none{nullptr, nullptr, nullptr, nullptr};
    Launch L;
    if (int rc = prepare(bvh, n_max, true, s, L)) return rc;
    return launch(bvh, L, ra, n_max, none, occluded, s, n_dev);
}

int trace_reserve(lr_qbvh* bvh, uint32_t n_max, hipStream_t s) {
    if (n_max == 0 || bvh->info.n_nodes == 0) return LR_OK;
    Launch L;
    return prepare(bvh, n_max, false, s, L);
}

}  // namespace lr

int lr_context::scratch(hipStream_t stream, size_t bytes, void** out) {
    std::lock_guard<std::mutex> g(mu);
    lr_scratch& s = spill[stream];
    if (s.bytes < bytes) {
        if (s.ptr) {
            LR_HIP_CHECK(hipStreamSynchronize(stream));
            LR_HIP_CHECK(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        LR_HIP_CHECK(hipMalloc(&s.ptr, bytes));
        s.bytes = bytes;
    }
    *out = s.ptr;
    return LR_OK;
}

void lr_context::release(hipStream_t stream) {
    std::lock_guard<std::mutex> g(mu);
    auto it = spill.find(stream);
    if (it == spill.end()) return;
    if (it->second.ptr) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(it->second.ptr);
    }
    spill.erase(it);
}

extern "C" {

const char* lr_last_error(void) { return g_last_error.c_str(); }

const char* lr_version(void) { return "liar_amd 0.1.0 (gfx950)"; }

int lr_context_create(int device, lr_context** out) {
    if (!out) return lr::fail(LR_E_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    LR_HIP_CHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return lr::fail(LR_E_INVALID, "device " + std::to_string(device) + " out of range (" +
                                          std::to_string(count) + " devices)");
    LR_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    LR_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
        return lr::fail(LR_E_UNSUPPORTED, std::string("device is ") + prop.gcnArchName + ", built for gfx950 only");
    // device address of this GPU's copy of the fixed scratch (the device is
    // current); lr_trace_closest_noalloc uses it and never the array's name
    void* fixed = nullptr;
    LR_HIP_CHECK(hipGetSymbolAddress(&fixed, HIP_SYMBOL(scratch)));
    if (!fixed) return lr::fail(LR_E_DEVICE, "no device address for the fixed traversal scratch");
    lr_context* ctx = new (std::nothrow) lr_context();
    if (!ctx) return lr::fail(LR_E_NOMEM, "out of host memory");
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    ctx->fixed_scratch = fixed;
    *out = ctx;
    return LR_OK;
}

int lr_context_destroy(lr_context* ctx) {
    if (!ctx) return LR_OK;
    (void)hipSetDevice(ctx->device);
    // streams in the map may be gone by now (the caller's, or a scene's):
    // wait for the whole device instead of for each of them
    (void)hipDeviceSynchronize();
    for (auto& kv : ctx->spill)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    delete ctx;
    return LR_OK;
}

int lr_set_trace_config(lr_context* ctx, const lr_trace_config* cfg) {
    if (!ctx || !cfg) return lr::fail(LR_E_INVALID, "NULL argument");
    lr_trace_config c;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        c = ctx->cfg;
    }
    if (cfg->block) c.block = cfg->block;
    if (cfg->blocks_per_cu) c.blocks_per_cu = cfg->blocks_per_cu;
    if (cfg->lds_stack) c.lds_stack = cfg->lds_stack;
    if (cfg->refill_min) c.refill_min = cfg->refill_min;
    if (cfg->schedule) c.schedule = cfg->schedule;
    if (c.refill_min > 64) return lr::fail(LR_E_INVALID, "refill_min must be in 1..64");
    if (c.schedule != LR_TRACE_SCHEDULE_STEP && c.schedule != LR_TRACE_SCHEDULE_PHASED)
        return lr::fail(LR_E_INVALID, "schedule must be LR_TRACE_SCHEDULE_STEP or LR_TRACE_SCHEDULE_PHASED");
    if (!find_kernel(c.block, c.lds_stack))
        return lr::fail(LR_E_INVALID, "unsupported (block, lds_stack) = (" + std::to_string(c.block) + ", " +
                                          std::to_string(c.lds_stack) +
                                          "); block in {64,128,256}, lds_stack in {8,16,24}");
    if (c.blocks_per_cu > 64) return lr::fail(LR_E_INVALID, "blocks_per_cu must be <= 64");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->cfg = c;
    return LR_OK;
}

int lr_get_trace_config(const lr_context* ctx, lr_trace_config* cfg) {
    if (!ctx || !cfg) return lr::fail(LR_E_INVALID, "NULL argument");
    *cfg = ctx->cfg;
    return LR_OK;
}

int lr_autotune_trace(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n, lr_hits_soa* hits, uint8_t* occluded,
                      uint32_t repeats, lr_tune_entry* entries, uint32_t max_entries, uint32_t* n_entries,
                      lr_trace_config* best, void* stream) {
    if (!bvh || !bvh->ctx) return lr::fail(LR_E_INVALID, "NULL or host-only qbvh");
    if (int rc = check_rays(rays, n)) return rc;
    const bool any = hits == nullptr;
    if (any && !occluded) return lr::fail(LR_E_INVALID, "need hits (closest) or occluded (any)");
    if (!any && (!hits->prim || !hits->t || !hits->u || !hits->v))
        return lr::fail(LR_E_INVALID, "a hit array pointer is NULL");
    if (entries && max_entries && !n_entries) return lr::fail(LR_E_INVALID, "entries given without n_entries");
    if (n == 0 || bvh->info.n_nodes == 0) return lr::fail(LR_E_INVALID, "nothing to tune: no rays or empty scene");
    if (int rc = lr::bind(bvh->ctx)) return rc;
    lr_context* ctx = bvh->ctx;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t reps = repeats ? repeats : 3;
    lr_trace_config saved;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        saved = ctx->cfg;
    }
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, nullptr};
    const HitsArg ha = any ? HitsArg{nullptr, nullptr, nullptr, nullptr} : HitsArg{hits->prim, hits->t, hits->u, hits->v};
    hipEvent_t e0, e1;
    LR_HIP_CHECK(hipEventCreate(&e0));
    LR_HIP_CHECK(hipEventCreate(&e1));
    static const uint32_t kBpc[] = {2, 4, 8, 16};
    static const uint32_t kRefill[] = {8, 16, 32, 48};
    static const uint32_t kSchedule[] = {LR_TRACE_SCHEDULE_STEP, LR_TRACE_SCHEDULE_PHASED};
    lr_trace_config win = saved;
    float win_ms = INFINITY;
    uint32_t count = 0;
    int rc = LR_OK;
    for (const auto& k : kKernels) {
        for (uint32_t bpc : kBpc) {
            for (uint32_t rfs = 0; rfs < 8 && !rc; ++rfs) {
                const uint32_t rf = kRefill[rfs & 3], sched = kSchedule[rfs >> 2];
                const lr_trace_config c{k.block, bpc, k.lds_stack, rf, sched};
                {
                    std::lock_guard<std::mutex> g(ctx->mu);
                    ctx->cfg = c;
                }
                Launch L;
                if ((rc = prepare(bvh, n, any, s, L))) break;
                if ((rc = launch(bvh, L, ra, n, ha, any ? occluded : nullptr, s))) break;  // warm-up
                hipError_t he = hipEventRecord(e0, s);
                for (uint32_t r = 0; r < reps && !rc; ++r) rc = launch(bvh, L, ra, n, ha, any ? occluded : nullptr, s);
                if (he == hipSuccess) he = hipEventRecord(e1, s);
                if (he == hipSuccess) he = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
                if (he != hipSuccess) rc = lr::fail(LR_E_DEVICE, std::string("lr_autotune_trace: ") + hipGetErrorString(he));
                if (rc) break;
                ms /= (float)reps;
                if (entries && count < max_entries) entries[count] = lr_tune_entry{c, ms};
                ++count;
                if (ms < win_ms) {
                    win_ms = ms;
                    win = c;
                }
            }
            if (rc) break;
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->cfg = rc ? saved : win;
    }
    if (rc) return rc;
    if (n_entries) *n_entries = std::min(count, max_entries);
    if (best) *best = win;
    return LR_OK;
}

int lr_qbvh_create(lr_context* ctx, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                   const float* spheres, uint32_t ns, lr_qbvh** out) {
    if (!out) return lr::fail(LR_E_INVALID, "out is NULL");
    *out = nullptr;
    if ((nt && (!tris || !verts)) || (ns && !spheres)) return lr::fail(LR_E_INVALID, "NULL geometry array");
    lr_qbvh* b = new (std::nothrow) lr_qbvh();
    if (!b) return lr::fail(LR_E_NOMEM, "out of host memory");
    b->ctx = ctx;
    std::string err;
    if (lr::build_qbvh(verts, nv, tris, nt, spheres, ns, b->host, err)) {
        delete b;
        return lr::fail(LR_E_INVALID, "lr_qbvh_create: " + err);
    }
    lr_qbvh_info& I = b->info;
    I.n_nodes = (uint32_t)b->host.nodes.size();
    I.n_prims = (uint32_t)b->host.prims.size();
    I.n_triangles = nt;
    I.n_spheres = ns;
    I.max_depth = b->host.max_depth;
    I.max_stack = 3 * (b->host.max_depth + 1) + 1;
    I.sah_cost = b->host.sah_cost;
    I.build_seconds = b->host.build_seconds;
    for (int k = 0; k < 3; ++k) {
        I.scene_lo[k] = b->host.scene_lo[k];
        I.scene_hi[k] = b->host.scene_hi[k];
    }
    b->host_q8 = lr::quantize_nodes(b->host.nodes);
    int rc = ctx ? lr::bind(ctx) : LR_OK;
    if (!rc && ctx && I.n_nodes) {
        hipError_t e = hipMalloc(&b->d_nodes, sizeof(lr_node_q8) * I.n_nodes);
        if (e == hipSuccess) e = hipMalloc(&b->d_prims, sizeof(lr_prim) * I.n_prims);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_nodes, b->host_q8.data(), sizeof(lr_node_q8) * I.n_nodes, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_prims, b->host.prims.data(), sizeof(lr_prim) * I.n_prims, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = lr::fail(LR_E_DEVICE, std::string("lr_qbvh_create upload: ") + hipGetErrorString(e));
    }
    if (rc) {
        lr_qbvh_destroy(b);
        return rc;
    }
    *out = b;
    return LR_OK;
}

int lr_qbvh_destroy(lr_qbvh* b) {
    if (!b) return LR_OK;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    if (b->d_nodes) (void)hipFree(b->d_nodes);
    if (b->d_prims) (void)hipFree(b->d_prims);
    delete b;
    return LR_OK;
}

int lr_qbvh_get_info(const lr_qbvh* b, lr_qbvh_info* info) {
    if (!b || !info) return lr::fail(LR_E_INVALID, "NULL argument");
    *info = b->info;
    return LR_OK;
}

int lr_qbvh_export(const lr_qbvh* b, lr_node* nodes, lr_prim* prims) {
    if (!b) return lr::fail(LR_E_INVALID, "NULL qbvh");
    if (nodes && !b->host.nodes.empty())
        std::memcpy(nodes, b->host.nodes.data(), sizeof(lr_node) * b->host.nodes.size());
    if (prims && !b->host.prims.empty())
        std::memcpy(prims, b->host.prims.data(), sizeof(lr_prim) * b->host.prims.size());
    return LR_OK;
}

int lr_qbvh_export_q8(const lr_qbvh* b, lr_node_q8* nodes) {
    if (!b) return lr::fail(LR_E_INVALID, "NULL qbvh");
    if (nodes && !b->host_q8.empty()) std::memcpy(nodes, b->host_q8.data(), sizeof(lr_node_q8) * b->host_q8.size());
    return LR_OK;
}

int lr_trace_closest(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n, lr_hits_soa* hits, void* stream) {
    if (!bvh) return lr::fail(LR_E_INVALID, "NULL qbvh");
    if (!bvh->ctx) return lr::fail(LR_E_INVALID, "QBVH was built host-only (NULL context); it cannot be traced");
    if (int rc = check_rays(rays, n)) return rc;
    if (!hits || (n && (!hits->prim || !hits->t || !hits->u || !hits->v)))
        return lr::fail(LR_E_INVALID, "a hit array pointer is NULL");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(bvh->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, nullptr};
    const HitsArg ha{hits->prim, hits->t, hits->u, hits->v};
    if (bvh->info.n_nodes == 0) {
        hipLaunchKernelGGL(k_fill_miss, dim3((n + 255) / 256), dim3(256), 0, s, n, ha);
        LR_HIP_CHECK(hipGetLastError());
        return LR_OK;
    }
    Launch L;
    if (int rc = prepare(bvh, n, false, s, L)) return rc;
    return launch(bvh, L, ra, n, ha, nullptr, s);
}

int lr_trace_closest_noalloc(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n, lr_hits_soa* hits, void* stream) {
    if (!bvh) return lr::fail(LR_E_INVALID, "NULL qbvh");
    if (!bvh->ctx) return lr::fail(LR_E_INVALID, "QBVH was built host-only (NULL context); it cannot be traced");
    if (int rc = check_rays(rays, n)) return rc;
    if (!hits || (n && (!hits->prim || !hits->t || !hits->u || !hits->v)))
        return lr::fail(LR_E_INVALID, "a hit array pointer is NULL");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(bvh->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, nullptr};
    const HitsArg ha{hits->prim, hits->t, hits->u, hits->v};
    if (bvh->info.n_nodes == 0) {
        hipLaunchKernelGGL(k_fill_miss, dim3((n + 255) / 256), dim3(256), 0, s, n, ha);
        LR_HIP_CHECK(hipGetLastError());
        return LR_OK;
    }
    Launch L;
    if (int rc = prepare_fixed(bvh, n, L)) return rc;
    // launch() clears the work counter with a kernel (k_zero_u32) first
    return launch(bvh, L, ra, n, ha, nullptr, s);
}

int lr_trace_any(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n, uint8_t* occluded, void* stream) {
    if (!bvh) return lr::fail(LR_E_INVALID, "NULL qbvh");
    if (!bvh->ctx) return lr::fail(LR_E_INVALID, "QBVH was built host-only (NULL context); it cannot be traced");
    if (int rc = check_rays(rays, n)) return rc;
    if (n && !occluded) return lr::fail(LR_E_INVALID, "occluded is NULL");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(bvh->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const RaysArg ra{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tnear, rays->tfar, nullptr};
    if (bvh->info.n_nodes == 0) {
        hipLaunchKernelGGL(k_fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, n, occluded, (uint8_t)0);
        LR_HIP_CHECK(hipGetLastError());
        return LR_OK;
    }
    Launch L;
    if (int rc = prepare(bvh, n, true, s, L)) return rc;
    const HitsArg none{nullptr, nullptr, nullptr, nullptr};
    return launch(bvh, L, ra, n, none, occluded, s);
}

}  // extern "C"
