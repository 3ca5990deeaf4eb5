// render.hip -- spectral wavefront path tracer for gfx950 and its C-ABI.
//
// Integrand: LiAR's AdjointPhotonTracer (src/liar/tracers/adjoint_photon_tracer.cpp:77-292)
// in SPECTRAL_MODE=Single: every path carries one sampled wavelength; at each
// surface a one-sample MIS choice between BSDF sampling and light sampling
// picks the next direction (power heuristic over both pdfs), emission is
// added when a later generation hits a light, Russian roulette applies from
// a configurable generation on.
//
// Wavefront: per chunk of paths (whole pixels), generation g runs
//   trace  : closest hit of queue[g&1] (persistent QBVH kernel, length read
//            from device memory)
//   shade  : emission + BSDF/light scatter; survivors appended to
//            queue[(g+1)&1] with one atomic per wave (ballot + mbcnt)
// and a final per-pixel reduction adds the paths' XYZ to the film in a fixed
// order, so renders are deterministic for a seed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "lr_internal.h"
#include "lr_render.h"

using namespace lr;

namespace {

std::mutex g_obs_mu;
uint64_t g_obs_version = 1;  // bumped by lr_set_standard_observer (scenes rebuild their tables)
std::vector<float> g_obs_w, g_obs_xyz;  // standard observer (nm, xyz interleaved)

struct DevCamera {
    float pos[3], dir[3], right[3], down[3], dir_n[3];
    float near_limit;
};

struct Params {
    uint32_t width, height, spp, nl, max_gen, rr_depth;
    float rr_p;
    float max_lum;  // 0: no clamp
    uint32_t seed_lo, seed_hi;
    uint32_t strata;  // sqrt(spp) if spp is a square, else 0
    // Halton sampler (LR_SAMPLER_HALTON): path p of the call is sequence
    // position hoff + p (index hoff + p + 1, and the key of its random
    // streams); generations < hgens draw from the sub-sequences whose digit
    // permutations start at perm + perm_off[prime index]
    uint32_t halton, hgens;
    uint64_t hoff;
    const uint16_t* perm;
    const uint32_t* perm_off;
};

// The Halton sub-sequence numbers of generation g (halton.cpp:180-230 with
// the AdjointPhotonTracer's requests, adjoint_photon_tracer.cpp:47-67): per
// generation the 1D strategy, component, light choice and scatter samples,
// then (after all 1D ones) the 2D BSDF, light and phase samples, each
// coordinate on the next prime.
__device__ inline float halton_sub(const Params& P, uint64_t index, uint32_t prime_index) {
    const uint32_t off = P.perm_off[prime_index];
    const uint32_t prime = P.perm_off[prime_index + 1] - off;
    return radical_inverse(index, prime, P.perm + off);
}

struct Queue {
    float* r[8];  // ox oy oz dx dy dz tnear tfar
    uint32_t* pid;
    uint32_t* root;  // batch mode: root node of the path's scene in the shared node array
};

// -- batch mode: one wavefront over many scenes ------------------------------
// A chunk is a run of segments, each a range of whole pixels of one scene;
// chunk-local path i of segment k is scene path pixel0 * ppp + i - path_begin.
struct Segment {
    uint32_t scene;
    uint32_t path_begin;  // first chunk-local path of the segment
    uint32_t pix_begin;   // first chunk-local pixel of the segment
    uint32_t pixel0;      // first pixel (scene-local linear index)
};

struct SceneFilm {
    float* xyz;
    float* w;
};

struct BatchArgs {
    const Segment* seg;  // this chunk's segments, ascending path_begin / pix_begin
    uint32_t n_seg;
    const DevScene* scenes;
    const Params* params;
    const DevCamera* cams;
    const GeomArgs* geoms;       // per scene: shading geometry
    const uint32_t* roots;
    const SceneFilm* films;
};

// Last segment whose path_begin (by_pixel: pix_begin) is <= x.
__device__ inline uint32_t find_segment(const BatchArgs& B, uint32_t x, bool by_pixel) {
    uint32_t lo = 0, hi = B.n_seg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t key = by_pixel ? B.seg[mid].pix_begin : B.seg[mid].path_begin;
        if (key <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct PathState {
    float* thr;
    float* L;
    float* lam;
    float* lpdf;
};

struct Hits {
    uint32_t* prim;
    float *t, *u, *v;
};

__device__ inline lr::u4 rng(const Params& P, uint64_t path, uint32_t gen, uint32_t purpose) {
    return lr::philox4x32(P.seed_lo, P.seed_hi, lr::u4{(uint32_t)path, (uint32_t)(path >> 32), gen, purpose});
}

// Block size of the kernels that append to a ray queue with block_append:
// one atomic per block, so 1024 threads make 4x fewer than 256.  The guided
// and batched k_shade (72 VGPRs) fit one 16-wave block per CU, 4 waves per
// SIMD, so they run 512-thread blocks instead (6 waves per SIMD).
constexpr uint32_t kAppendBlock = 1024;
constexpr uint32_t kAppendBlockWide = 512;

// Slot of this lane's entry in a queue that the whole workgroup appends to:
// wave counts by ballot, one atomic per workgroup (a single counter address
// serialises its atomics, so fewer of them matter).  Every thread of the
// block must call it; blockDim.x is a multiple of 64, at most 1024.
__device__ inline uint32_t block_append(bool alive, uint32_t* count) {
    __shared__ uint32_t s_cnt[16];
    __shared__ uint32_t s_base;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t mask = __ballot(alive);
    if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t c = s_cnt[w];
            s_cnt[w] = tot;
            tot += c;
        }
        s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    return s_base + s_cnt[wave] +
           __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Start of a generation: queue `cur` is about to be traced (its rays are
// counted), queue `nxt` starts empty.
__global__ void k_next_gen(uint32_t* counts, uint32_t cur, uint32_t nxt, uint64_t* rays_total) {
    if (threadIdx.x == 0) {
        *rays_total += counts[cur];
        counts[nxt] = 0u;
    }
}

// -- generation 0: camera rays (perspective_camera.cpp:406-434) -------------
template <bool BATCH>
__global__ void k_camera(DevScene S0, DevCamera C0, Params P0, uint64_t path0, uint32_t n, PathState ps, Queue q,
                         uint32_t* count, BatchArgs B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    V3 d = v3(0, 0, 1);
    float ray_factor = 1.0f;
    DevCamera C = C0;
    uint32_t root = 0;
    if (i < n) {
    DevScene S = S0;
    Params P = P0;
    uint64_t p = P0.hoff + path0 + i;  // Halton: the path's index in the sequence
    if (BATCH) {
        const Segment sg = B.seg[find_segment(B, i, false)];
        S = B.scenes[sg.scene];
        C = B.cams[sg.scene];
        P = B.params[sg.scene];
        p = (uint64_t)sg.pixel0 * (P.spp * P.nl) + (i - sg.path_begin);
        root = B.roots[sg.scene];
    }
    const lr::u4 rp = rng(P, p, 0xFFFEu, 0u);     // per path
    float X, Y, uw;
    if (P.halton) {
        // one global sequence over the film: screen in bases 2 and 3,
        // wavelength in base 13 (halton.cpp:191-212)
        const uint64_t j = p + 1;
        X = radical_inverse(j, 2u, nullptr);
        Y = radical_inverse(j, 3u, nullptr);
        uw = radical_inverse(j, 13u, nullptr);
    } else {
        const uint32_t l = (uint32_t)(p % P.nl);
        const uint64_t sidx = p / P.nl;
        const uint32_t s = (uint32_t)(sidx % P.spp);
        const uint64_t pixel = sidx / P.spp;
        const uint32_t px = (uint32_t)(pixel % P.width), py = (uint32_t)(pixel / P.width);
        const lr::u4 rs = rng(P, sidx, 0xFFFFu, 0u);  // per camera sample (shared by its wavelengths)
        float sx = lr::u01(rs.x), sy = lr::u01(rs.y);
        if (P.strata) {
            sx = ((float)(s % P.strata) + sx) / (float)P.strata;
            sy = ((float)(s / P.strata) + sy) / (float)P.strata;
        }
        X = ((float)px + sx) / (float)P.width;
        Y = ((float)py + sy) / (float)P.height;
        // stratified wavelength (stratifier.cpp:191-196 with nl strata per sample)
        uw = ((float)l + lr::u01(rp.x)) / (float)P.nl;
    }
    d = ld3(C.dir) + (X - 0.5f) * ld3(C.right) + (Y - 0.5f) * ld3(C.down);
    d = normalize(d);
    ray_factor = 1.0f / dot(d, ld3(C.dir_n));
    float lpdf;
    const float lam = observer_sample(S.obs, uw, lpdf);
    float thr = lpdf > 0.0f ? 1.0f : 0.0f;
    // generation guard / roulette of tracePhoton(0) (adjoint_photon_tracer.cpp:84-101)
    if (P.max_gen == 0) thr = 0.0f;
    if (thr > 0.0f && P.rr_depth == 0) {
        if (lr::u01(rp.y) < P.rr_p) thr /= P.rr_p;
        else thr = 0.0f;
    }
    ps.thr[i] = thr;
    ps.L[i] = 0.0f;
    ps.lam[i] = lam;
    ps.lpdf[i] = lpdf;
    alive = thr > 0.0f;
    }
    const uint32_t pos = block_append(alive, count);
    {
        if (alive) {
            q.r[0][pos] = C.pos[0];
            q.r[1][pos] = C.pos[1];
            q.r[2][pos] = C.pos[2];
            q.r[3][pos] = d.x;
            q.r[4][pos] = d.y;
            q.r[5][pos] = d.z;
            q.r[6][pos] = fmaxf(kTolerance, ray_factor * C.near_limit);
            q.r[7][pos] = INFINITY;
            q.pid[pos] = i;
            if (BATCH) q.root[pos] = root;
        }
    }
}

__global__ void k_zero_words(uint32_t* c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0u;
}

__global__ void k_fill_miss_queue(const uint32_t* count, Hits h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *count) {
        h.prim[i] = LR_MISS;
        h.t[i] = INFINITY;
    }
}

// -- guided mode ---------------------------------------------------------------
// The guide network proposes incident directions at diffuse / glossy
// vertices; its mixture (8 lobes x (weight, mean xyz, kappa)) is inferred per
// queue entry before shading.
struct GuideArgs {
    const float* in;   // [queue][8] network inputs (written by k_guide_inputs)
    const float* mix;  // [48][mix_ld] mixture, strategy probabilities, radiance (lr_guide_infer_columns)
    uint32_t mix_ld;
    float sel;         // guide share of the non-light strategies
    float light_share; // probability of light sampling
    uint32_t learned;  // strategy probabilities from the network (mix[40..42]) instead of sel / light_share
    uint32_t balance;  // balance heuristic instead of the power heuristic
    uint32_t image_weight;  // training weights times the throughput before the vertex
    uint32_t record_depth;  // record vertices of generations < record_depth only
    float lo[3], inv_ext[3];
    float lam_lo, inv_lam_ext;
    // training records (cap entries): inputs [8] (network inputs 0..5, then
    // the light and BSDF sampling pdfs of the taken direction), direction [3],
    // aux [4] = (f|cos| / p_mix, L accumulated at the vertex, throughput after
    // scatter before roulette, path id bits)
    uint32_t record, cap;
    uint32_t* count;
    float* r_in;
    float* r_dir;
    float* r_aux;
    // radiance-driven Russian roulette: est[path] = throughput x the network's
    // reflected-radiance estimate at the path's first diffuse / glossy vertex
    // (0 before it); boost[path] = the product of the path's 1 / q so far,
    // kept <= 1 / rr_min
    uint32_t rr, rr_start;
    float rr_scale, rr_min;
    float* est;
    float* boost;
};

// The mixture of one queue entry: value c at m[c * ld] (columns of the
// inferred mixture).
__device__ inline float vmf_mix_pdf(const float* m, uint32_t ld, V3 w) {
    float p = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const float kap = m[(5 * k + 4) * (size_t)ld];
        const float c = m[(5 * k + 1) * (size_t)ld] * w.x + m[(5 * k + 2) * (size_t)ld] * w.y +
                        m[(5 * k + 3) * (size_t)ld] * w.z;
        p += m[5 * k * (size_t)ld] * kap / (2.0f * kPi * (-expm1f(-2.0f * kap))) * expf(kap * (c - 1.0f));
    }
    return p;
}

// Lobe chosen by weight, then Jakob's stable vMF sampling ("Numerically
// stable sampling of the von Mises-Fisher distribution on S^2", 2012).
__device__ inline V3 vmf_mix_sample(const float* m, uint32_t ld, float u0, float u1, float u2) {
    int k = 0;
    float acc = m[0];
    while (k < 7 && u0 >= acc) {
        ++k;
        acc += m[5 * k * (size_t)ld];
    }
    const V3 mu = v3(m[(5 * k + 1) * (size_t)ld], m[(5 * k + 2) * (size_t)ld], m[(5 * k + 3) * (size_t)ld]);
    const float kap = m[(5 * k + 4) * (size_t)ld];
    const float w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap;
    const float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const float phi = 2.0f * kPi * u2;
    V3 t, b;
    onb(mu, t, b);
    return normalize((s * cosf(phi)) * t + (s * sinf(phi)) * b + w * mu);
}

__global__ void k_guide_inputs(DevScene S, Queue q, const uint32_t* count, Hits h, PathState ps, GuideArgs G,
                               float* in) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t prim = h.prim[i];
    if (prim != LR_MISS) {
        const V3 o = v3(q.r[0][i], q.r[1][i], q.r[2][i]);
        const V3 d = v3(q.r[3][i], q.r[4][i], q.r[5][i]);
        const V3 x = o + h.t[i] * d;
        v[0] = fminf(fmaxf((x.x - G.lo[0]) * G.inv_ext[0], 0.0f), 1.0f);
        v[1] = fminf(fmaxf((x.y - G.lo[1]) * G.inv_ext[1], 0.0f), 1.0f);
        v[2] = fminf(fmaxf((x.z - G.lo[2]) * G.inv_ext[2], 0.0f), 1.0f);
        // outgoing direction (towards the previous vertex) in spherical coordinates
        v[3] = acosf(fminf(fmaxf(-d.z, -1.0f), 1.0f)) / kPi;
        v[4] = (atan2f(-d.y, -d.x) + kPi) / (2.0f * kPi);
        v[5] = fminf(fmaxf((ps.lam[q.pid[i]] - G.lam_lo) * G.inv_lam_ext, 0.0f), 1.0f);
    }
    float4* o4 = reinterpret_cast<float4*>(in + (size_t)i * 8);
    o4[0] = make_float4(v[0], v[1], v[2], v[3]);
    o4[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// Training weights of the recorded vertices: w = f|cos|/p_mix * L_i, with
// the incident radiance estimate L_i = (L_path - L_vertex) / throughput.
// The record count is read on the device (capped at cap); fixed grid.
__global__ void k_guide_targets(uint32_t cap, const uint32_t* __restrict__ count, const float* __restrict__ aux,
                                const float* __restrict__ L, float* __restrict__ w) {
    const uint32_t n = min(*count, cap);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const float fcp = aux[4 * r], lv = aux[4 * r + 1], tn = aux[4 * r + 2];
        const uint32_t pid = __float_as_uint(aux[4 * r + 3]);
        const float rest = L[pid] - lv;
        const float v = (tn > 0.0f && rest > 0.0f) ? fcp * rest / tn : 0.0f;
        w[r] = isfinite(v) ? v : 0.0f;
    }
}

// -- one generation: emission + scatter (adjoint_photon_tracer.cpp:140-292) --
template <bool GUIDED, bool BATCH>
__global__ void k_shade(DevScene S0, Params P0, uint32_t gen, uint64_t path0, Queue q, const uint32_t* count,
                        Hits h, GeomArgs GA0, PathState ps, Queue qn, uint32_t* count_next,
                        GuideArgs G, BatchArgs B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < *count;
    bool alive = false;
    V3 start = v3(0, 0, 0), dir = v3(0, 0, 1);
    uint32_t pid = 0, root = 0;
    if (valid) {
        pid = q.pid[i];
        DevScene S = S0;
        Params P = P0;
        GeomArgs Gm = GA0;
        // the path's index in its scene (random streams; Halton: in the sequence)
        uint64_t gp = P0.hoff + path0 + pid;
        if (BATCH) {
            const Segment sg = B.seg[find_segment(B, pid, false)];
            S = B.scenes[sg.scene];
            P = B.params[sg.scene];
            Gm = B.geoms[sg.scene];
            gp = (uint64_t)sg.pixel0 * (P.spp * P.nl) + (pid - sg.path_begin);
            root = q.root[i];
        }
        const V3 o = v3(q.r[0][i], q.r[1][i], q.r[2][i]);
        const V3 d = v3(q.r[3][i], q.r[4][i], q.r[5][i]);
        const float tnear = q.r[6][i];
        float thr = ps.thr[pid];
        const float thr_in = thr;
        if (GUIDED && G.rr && gen == 0) {  // no roulette reference or weight yet
            G.est[pid] = 0.0f;
            G.boost[pid] = 1.0f;
        }
        float L = ps.L[pid];
        const float lam = ps.lam[pid];
        const uint32_t prim = h.prim[i];
        if (prim == LR_MISS) {
            // escaped: the sky / sun are lights, counted from generation 1 on
            if (gen > 0) {
                for (uint32_t k = 0; k < S.n_lights; ++k) {
                    const DevLight& Lk = S.lights[k];
                    if ((Lk.type == LR_LIGHT_SKY || Lk.type == LR_LIGHT_SUN) && light_sees(Lk, d))
                        L += thr * spectrum_eval(S, Lk.radiance, lam, false);
                }
            }
        } else {
            const HitFrame F = hit_frame(Gm, S.n_triangles, prim, o, d, tnear, h.t[i], h.u[i], h.v[i]);
            const V3 x = F.x, nf = F.ns, tt = F.tt, bb = F.bb;
            const bool leaving = F.leaving;
            const DevMaterial m = S.mats[S.prim_mat[prim]];
            if (gen > 0 && m.light != LR_NONE) {
                const DevLight& Lk = S.lights[m.light];
                if (light_sees(Lk, d)) L += thr * spectrum_eval(S, Lk.radiance, lam, false);
            }
            if ((gen == 0 || m.light == LR_NONE) && m.emission != LR_NONE)
                L += thr * spectrum_eval(S, m.emission, lam, false);

            if (m.type != LR_MAT_NONE) {
                BsdfParams B;
                B.type = m.type;
                B.refl = spectrum_eval(S, m.refl, lam, true);
                B.trans = spectrum_eval(S, m.trans, lam, true);
                B.eta = fmaxf(spectrum_eval(S, m.eta, lam, false), 1e-9f);
                B.kappa = fmaxf(spectrum_eval(S, m.kappa, lam, false), 1e-9f);
                B.ior = leaving ? B.eta / B.kappa : B.kappa / B.eta;  // dielectric.cpp:150-155
                const float ru = fmaxf(m.roughness, 1e-3f), rv = fmaxf(m.roughness_v, 1e-3f);
                B.alpha_u = ru * ru;
                B.alpha_v = rv * rv;
                B.fuzz = m.fuzz;
                const V3 ngl = local_geometric_normal(F);
                const V3 wiw = -d;
                const V3 wi = v3(dot(wiw, tt), dot(wiw, bb), fmaxf(dot(wiw, nf), 1.1920929e-7f));
                const lr::u4 r0 = rng(P, gp, gen, 1u);
                const lr::u4 r1 = rng(P, gp, gen, 2u);
                float u_strat = lr::u01(r0.x), u1 = lr::u01(r0.y), u2 = lr::u01(r0.z);
                float u_comp = lr::u01(r0.w), u_light = lr::u01(r1.x), ul1 = lr::u01(r1.y), ul2 = lr::u01(r1.z);
                const float u_rr = lr::u01(r1.w);
                if (!BATCH && !GUIDED && P.halton && gen < P.hgens) {
                    const uint64_t j = gp + 1;
                    const uint32_t n1 = 4u * P.hgens, b2 = n1 + 6u * gen;
                    u_strat = halton_sub(P, j, 4u * gen);
                    u_comp = halton_sub(P, j, 4u * gen + 1u);
                    u_light = halton_sub(P, j, 4u * gen + 2u);
                    u1 = halton_sub(P, j, b2);
                    u2 = halton_sub(P, j, b2 + 1u);
                    ul1 = halton_sub(P, j, b2 + 2u);
                    ul2 = halton_sub(P, j, b2 + 3u);
                }
                const float strategy = bsdf_is_diffuse_or_glossy(m.type) ? 0.5f : 1.0f;
                V3 wo;
                float value = 0.0f, pdf = 0.0f;
                bool ok = false;
                float rec_fcp = 0.0f, rec_pl = 0.0f, rec_pb = 0.0f;
                if (GUIDED && strategy < 1.0f) {
                    // one-sample MIS over BSDF, light and guide sampling
                    // (power or balance heuristic over c_i p_i): the strategy
                    // probabilities are the network's (learned) or light
                    // sampling takes `light_share` and the guide `sel` of the rest
                    const float* mx = G.mix + i;  // column layout: value c at mx[c * G.mix_ld]
                    float cl, cg, cb;
                    if (G.learned) {
                        cl = mx[40 * (size_t)G.mix_ld];
                        cg = mx[41 * (size_t)G.mix_ld];
                        cb = mx[42 * (size_t)G.mix_ld];
                    } else {
                        cl = G.light_share;
                        cg = (1.0f - cl) * G.sel;
                        cb = (1.0f - cl) * (1.0f - G.sel);
                    }
                    if (S.n_lights == 0) {  // no light to sample: its share goes to the others
                        const float rest = cg + cb;
                        cg /= rest;
                        cb /= rest;
                        cl = 0.0f;
                    }
                    // the strategy picks the direction; the BSDF value and the
                    // three densities of that direction are then evaluated
                    // once, in code every lane shares (the same BSDF and
                    // light-pdf evaluation whichever strategy sampled it)
                    V3 dw = v3(0, 0, 1);
                    float f = 0.0f, pb = 0.0f, pl = 0.0f, pg = 0.0f, pc = 0.0f;
                    int chosen = -1;
                    uint32_t lk = 0xFFFFFFFFu;  // the sampled light (light strategy)
                    float pdf_lk = 0.0f;        // its density of the sampled direction
                    if (u_strat < cb) {
                        float fs, ps;
                        if (bsdf_sample(B, wi, u1, u2, u_comp, wo, fs, ps)) {
                            dw = wo.x * tt + wo.y * bb + wo.z * nf;
                            chosen = 0;
                        }
                    } else if (u_strat < cb + cl || !(cg > 0.0f)) {  // (cb + cl may round below 1 when cg = 0)
                        if (S.n_lights > 0) {
                            lk = light_pick(S, u_light);
                            const DevLight& Lk = S.lights[lk];
                            float dist_l;
                            dw = light_sample(Lk, S.light_tris, x, ul1, ul2, pdf_lk, dist_l);
                            if (Lk.sel_pdf > 0.0f && pdf_lk > 0.0f) {
                                wo = v3(dot(dw, tt), dot(dw, bb), dot(dw, nf));
                                chosen = 1;
                            }
                        }
                    } else {
                        const lr::u4 r2 = rng(P, gp, gen, 3u);
                        dw = vmf_mix_sample(mx, G.mix_ld, lr::u01(r2.x), lr::u01(r2.y), lr::u01(r2.z));
                        wo = v3(dot(dw, tt), dot(dw, bb), dot(dw, nf));
                        chosen = 2;
                    }
                    ok = chosen >= 0 && bsdf_evaluate(B, wi, wo, ngl, f, pb) && f != 0.0f;
                    if (ok) {
                        // the sampled light's own density (re-evaluating it can
                        // miss a direction at the edge of a small source such as
                        // the sun's cone), the others' evaluated
                        for (uint32_t k = 0; k < S.n_lights; ++k)
                            pl += S.lights[k].sel_pdf *
                                  (k == lk ? pdf_lk : light_pdf(S.lights[k], S.light_tris, x, dw));
                        pg = cg > 0.0f ? vmf_mix_pdf(mx, G.mix_ld, dw) : 0.0f;  // no guide share: no mixture
                        pc = chosen == 0 ? cb * pb : (chosen == 1 ? cl * pl : cg * pg);
                    }
                    if (ok && pc > 0.0f) {
                        const float qb = cb * pb, ql = cl * pl, qg = cg * pg;
                        value = G.balance ? f * pc / (qb + ql + qg) : f * (pc * pc) / (qb * qb + ql * ql + qg * qg);
                        pdf = pc;
                        rec_fcp = f * fabsf(wo.z) / (qb + ql + qg);
                        rec_pl = pl;
                        rec_pb = pb;
                    } else {
                        ok = false;
                    }
                } else if (u_strat < strategy) {
                    // BSDF sampling, MIS-weighted against all lights
                    ok = bsdf_sample(B, wi, u1, u2, u_comp, wo, value, pdf);
                    if (ok) {
                        pdf *= strategy;
                        float pdf_lights = 0.0f;
                        if (strategy < 1.0f) {
                            const V3 dw = wo.x * tt + wo.y * bb + wo.z * nf;
                            for (uint32_t k = 0; k < S.n_lights; ++k)
                                pdf_lights += S.lights[k].sel_pdf * light_pdf(S.lights[k], S.light_tris, x, dw);
                            pdf_lights *= (1.0f - strategy);
                        }
                        value *= (pdf * pdf) / (pdf * pdf + pdf_lights * pdf_lights);
                    }
                } else if (S.n_lights > 0) {
                    // light sampling
                    const uint32_t k = light_pick(S, u_light);
                    const DevLight& Lk = S.lights[k];
                    float pdf_l = 0.0f;
                    float dist_l;
                    const V3 dw = light_sample(Lk, S.light_tris, x, ul1, ul2, pdf_l, dist_l);
                    if (Lk.sel_pdf > 0.0f && pdf_l > 0.0f) {
                        float pl = Lk.sel_pdf * pdf_l;
                        for (uint32_t j = 0; j < S.n_lights; ++j)
                            if (j != k) pl += S.lights[j].sel_pdf * light_pdf(S.lights[j], S.light_tris, x, dw);
                        pl *= (1.0f - strategy);
                        wo = v3(dot(dw, tt), dot(dw, bb), dot(dw, nf));
                        float vb, pb;
                        if (bsdf_evaluate(B, wi, wo, ngl, vb, pb) && vb != 0.0f) {
                            const float pbs = pb * strategy;
                            value = vb * (pl * pl) / (pl * pl + pbs * pbs);
                            pdf = pl;
                            ok = true;
                        }
                    }
                }
                if (ok && pdf > 0.0f) {
                    const float cos_t = wo.z;
                    thr *= value * fabsf(cos_t) / pdf;
                    start = x + ((cos_t > 0.0f ? 10.0f : -10.0f) * kTolerance) * F.ng;  // adjoint_photon_tracer.cpp:177
                    dir = normalize(wo.x * tt + wo.y * bb + wo.z * nf);
                    if (GUIDED && G.record && gen < G.record_depth && rec_fcp > 0.0f && isfinite(thr) && thr > 0.0f) {
                        const uint32_t slot = atomicAdd(G.count, 1u);
                        if (slot < G.cap) {
                            const float4* src = reinterpret_cast<const float4*>(G.in + (size_t)i * 8);
                            float4* dst = reinterpret_cast<float4*>(G.r_in + (size_t)slot * 8);
                            const float4 tail = src[1];
                            dst[0] = src[0];
                            dst[1] = make_float4(tail.x, tail.y, rec_pl, rec_pb);
                            G.r_dir[3 * (size_t)slot] = dir.x;
                            G.r_dir[3 * (size_t)slot + 1] = dir.y;
                            G.r_dir[3 * (size_t)slot + 2] = dir.z;
                            G.r_aux[4 * (size_t)slot] = G.image_weight ? rec_fcp * thr_in : rec_fcp;
                            G.r_aux[4 * (size_t)slot + 1] = L;
                            G.r_aux[4 * (size_t)slot + 2] = thr;
                            G.r_aux[4 * (size_t)slot + 3] = __uint_as_float(pid);
                        }
                    }
                    const uint32_t next = gen + 1;
                    if (next >= P.max_gen) thr = 0.0f;
                    else if (next >= P.rr_depth) {
                        if (u_rr < P.rr_p) thr /= P.rr_p;
                        else thr = 0.0f;
                    } else if (GUIDED && G.rr && strategy < 1.0f) {
                        // radiance-driven roulette: the expected contribution
                        // of the rest of the path, thr_in x L_out, against the
                        // one expected at its first diffuse / glossy vertex;
                        // the path's total roulette weight stays <= 1 / rr_min
                        // (an estimate that drops along a path would otherwise
                        // compound the weights of successive survivals)
                        const float lo = thr_in * G.mix[43 * (size_t)G.mix_ld + i];
                        const float ref = gen == 0 ? 0.0f : G.est[pid];
                        if (ref > 0.0f) {
                            if (gen >= G.rr_start) {
                                const float b = G.boost[pid];
                                const float q = fminf(1.0f, fmaxf(b * G.rr_min, lo / (G.rr_scale * ref)));
                                if (u_rr < q) {
                                    thr /= q;
                                    if (q < 1.0f) G.boost[pid] = b / q;
                                } else {
                                    thr = 0.0f;
                                }
                            }
                        } else {
                            G.est[pid] = lo;
                        }
                    }
                    alive = thr > 0.0f && isfinite(thr);
                }
            }
        }
        ps.L[pid] = L;
        ps.thr[pid] = thr;
    }
    const uint32_t pos = block_append(alive, count_next);
    {
        if (alive) {
            qn.r[0][pos] = start.x;
            qn.r[1][pos] = start.y;
            qn.r[2][pos] = start.z;
            qn.r[3][pos] = dir.x;
            qn.r[4][pos] = dir.y;
            qn.r[5][pos] = dir.z;
            qn.r[6][pos] = kTolerance;
            qn.r[7][pos] = INFINITY;
            qn.pid[pos] = pid;
            if (BATCH) qn.root[pos] = root;
        }
    }
}

// -- film: Raster box filter, fixed-order per-pixel sums (raster.cpp:547-606)
// Film splat in tiles: a block takes up to pb = max(1, 256 / ppp) pixels of
// one scene (ppp = spp x wavelengths paths each, stored pixel after pixel).
// Each thread turns one path's (L, lambda, pdf) into its XYZ contribution in
// LDS, then one thread per pixel adds its paths' contributions up in path
// order -- the same terms, in the same order, as a thread walking the
// pixel's paths, so films are the same bit for bit -- while the table lookups
// and loads run one path per lane instead of one pixel per lane.
__device__ void splat_tile(const DevScene& S, const Params& P, uint32_t npb, uint32_t path_begin, uint64_t px0,
                           const PathState& ps, float* film_xyz, float* film_w) {
    __shared__ float c[3][256];
    __shared__ uint8_t has[256];
    const uint32_t ppp = P.spp * P.nl;
    const uint32_t tile = ppp >= 256u ? 256u : (256u / ppp) * ppp;  // paths per pass
    const uint32_t total = npb * ppp;
    const uint32_t t = threadIdx.x;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    for (uint32_t t0 = 0; t0 < total; t0 += tile) {
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        bool h = false;
        const uint32_t k = t0 + t;
        if (t < tile && k < total) {
            const uint32_t i = path_begin + k;
            const float L = ps.L[i];
            if (L != 0.0f) {
                float xyz[3];
                observer_xyz(S.obs, ps.lam[i], xyz);
                float sc = L / ps.lpdf[i];  // Spectral::xyz, spectral.cpp:212-221
                if (P.max_lum > 0.0f && sc * xyz[1] > P.max_lum) sc *= P.max_lum / (sc * xyz[1]);  // raster.cpp:567-570
                cx = sc * xyz[0];
                cy = sc * xyz[1];
                cz = sc * xyz[2];
                h = true;
            }
        }
        c[0][t] = cx;
        c[1][t] = cy;
        c[2][t] = cz;
        has[t] = h ? 1 : 0;
        __syncthreads();
        if (t < npb) {
            const uint32_t lo = max(t * ppp, t0), hi = min((t + 1) * ppp, t0 + tile);
            for (uint32_t kk = lo; kk < hi; ++kk) {
                if (!has[kk - t0]) continue;
                X += c[0][kk - t0];
                Y += c[1][kk - t0];
                Z += c[2][kk - t0];
            }
        }
        __syncthreads();
    }
    if (t < npb) {
        const uint64_t px = px0 + t;
        const uint64_t np = (uint64_t)P.width * P.height;
        film_xyz[px] += X;
        film_xyz[np + px] += Y;
        film_xyz[2 * np + px] += Z;
        film_w[px] += (float)ppp;
    }
}

// One scene's chunk: block b takes pixels [b pb, b pb + pb) of the chunk.
__global__ __launch_bounds__(256) void k_splat_tiles(DevScene S, Params P, uint64_t pix0, uint32_t npix,
                                                     PathState ps, float* film_xyz, float* film_w) {
    const uint32_t ppp = P.spp * P.nl;
    const uint32_t pb = ppp >= 256u ? 1u : 256u / ppp;
    const uint32_t pixb = blockIdx.x * pb;
    if (pixb >= npix) return;  // the whole block
    splat_tile(S, P, min(pb, npix - pixb), pixb * ppp, pix0 + pixb, ps, film_xyz, film_w);
}

// Halton chunk: each path's contribution goes to the pixel its screen point
// falls in (Raster::doWriteRender, raster.cpp:547-606), with float atomics.
__global__ __launch_bounds__(256) void k_splat_scatter(DevScene S, Params P, uint64_t path0, uint32_t n, PathState ps,
                                                       float* film_xyz, float* film_w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = P.hoff + path0 + i + 1;
    const float X = radical_inverse(j, 2u, nullptr), Y = radical_inverse(j, 3u, nullptr);
    const uint32_t px = min((uint32_t)(X * (float)P.width), P.width - 1u);
    const uint32_t py = min((uint32_t)(Y * (float)P.height), P.height - 1u);
    const uint64_t pix = (uint64_t)py * P.width + px;
    const uint64_t np = (uint64_t)P.width * P.height;
    atomicAdd(&film_w[pix], 1.0f);
    const float L = ps.L[i];
    if (L == 0.0f) return;
    float xyz[3];
    observer_xyz(S.obs, ps.lam[i], xyz);
    float sc = L / ps.lpdf[i];  // Spectral::xyz, spectral.cpp:212-221
    if (P.max_lum > 0.0f && sc * xyz[1] > P.max_lum) sc *= P.max_lum / (sc * xyz[1]);  // raster.cpp:567-570
    atomicAdd(&film_xyz[pix], sc * xyz[0]);
    atomicAdd(&film_xyz[np + pix], sc * xyz[1]);
    atomicAdd(&film_xyz[2 * np + pix], sc * xyz[2]);
}

// A batch chunk: block b takes the pixels of tile b, all of one segment.
struct SplatTile {
    uint32_t seg;   // chunk-local segment
    uint32_t pix;   // first chunk-local pixel
    uint32_t npix;  // pixels (<= max(1, 256 / ppp))
};

__global__ __launch_bounds__(256) void k_splat_tiles_batch(BatchArgs B, const SplatTile* __restrict__ tiles,
                                                           PathState ps) {
    const SplatTile T = tiles[blockIdx.x];
    const Segment sg = B.seg[T.seg];
    const Params P = B.params[sg.scene];
    const uint32_t ppp = P.spp * P.nl;
    const uint32_t off = T.pix - sg.pix_begin;
    splat_tile(B.scenes[sg.scene], P, T.npix, sg.path_begin + off * ppp, (uint64_t)sg.pixel0 + off, ps,
               B.films[sg.scene].xyz, B.films[sg.scene].w);
}

inline uint32_t splat_blocks(uint32_t npix, uint32_t ppp) {
    const uint32_t pb = ppp >= 256u ? 1u : 256u / ppp;
    return (npix + pb - 1) / pb;
}

// -- host helpers --------------------------------------------------------------

template <class T>
int upload(std::vector<T>& v, T** d) {
    *d = nullptr;
    if (v.empty()) return LR_OK;
    LR_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(d), sizeof(T) * v.size()));
    LR_HIP_CHECK(hipMemcpy(*d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return LR_OK;
}

float host_planck(float w_m, float temp) { return lr::planck(w_m, temp); }

// Host evaluation of a user spectrum (for luminance / normalisation).
double host_spectrum(const lr_spectrum& s, double w) {
    switch (s.kind) {
    case LR_SPEC_CONSTANT:
        return s.value;
    case LR_SPEC_SAMPLED: {
        const float* ws = s.w;
        const uint32_t n = s.count;
        const float* it = std::upper_bound(ws, ws + n, (float)w);
        if (it == ws || it == ws + n) return 0.0;
        const size_t k = (size_t)(it - ws);
        const double t = (w - ws[k - 1]) / ((double)ws[k] - ws[k - 1]);
        return s.v[k - 1] + t * ((double)s.v[k] - s.v[k - 1]);
    }
    case LR_SPEC_BINNED: {
        if (!(w >= s.lo && w < s.hi)) return 0.0;
        uint32_t k = (uint32_t)((w - s.lo) / ((double)s.hi - s.lo) * s.count);
        return s.v[std::min(k, s.count - 1)];
    }
    case LR_SPEC_BLACKBODY: {
        const float peak = 2.897771955e-3f / s.temperature;
        return (double)s.scale * host_planck((float)(w * 1e-9), s.temperature) / host_planck(peak, s.temperature);
    }
    }
    return 0.0;
}

// Luminance Y of a spectrum under the standard observer (trapezoid over the
// observer table, normalised so that Y(1) = 1; observer.cpp:105-123).
double host_luminance(const lr_spectrum& s) {
    std::lock_guard<std::mutex> g(g_obs_mu);
    const size_t n = g_obs_w.size();
    double num = 0.0, den = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double wp = g_obs_w[k > 0 ? k - 1 : 0], wn = g_obs_w[k + 1 < n ? k + 1 : n - 1];
        const double dw = (wn - wp) / 2.0;
        num += host_spectrum(s, g_obs_w[k]) * g_obs_xyz[3 * k + 1] * dw;
        den += g_obs_xyz[3 * k + 1] * dw;
    }
    return den > 0.0 ? num / den : 0.0;
}

}  // namespace

struct lr_scene {
    lr_context* ctx = nullptr;
    lr_qbvh* bvh = nullptr;
    uint32_t n_triangles = 0, n_lights = 0;
    uint32_t* d_prim_mat = nullptr;
    float4* d_prim_geom = nullptr;  // GeomArgs::geom
    float4* d_prim_tan = nullptr;   // GeomArgs::tan
    float4* d_corner = nullptr;     // GeomArgs::cn (lr_scene_set_normals)
    DevMaterial* d_mats = nullptr;
    DevSpectrum* d_spectra = nullptr;
    float* d_table = nullptr;
    DevLight* d_lights = nullptr;
    float4* d_light_tris = nullptr;  // mesh lights (DevScene::light_tris)
    uint16_t* d_halton_perm = nullptr;  // Halton sampler tables (ensure_halton)
    uint32_t* d_halton_off = nullptr;
    // observer restricted to the last render's sampling domain
    float obs_lo = -1.0f, obs_hi = -1.0f;
    uint32_t obs_n = 0;
    uint64_t obs_version = 0;
    float* d_obs = nullptr;
    // wavefront workspace (grow-only)
    size_t cap = 0;
    float* d_work = nullptr;
    uint32_t* d_counts = nullptr;  // [0..1] queue lengths, [4..5] u64 ray counter, [8] records
    // guided mode: inputs [cap][8], mixture [cap][40], records
    float* d_gwork = nullptr;
    size_t gcap = 0, grec = 0;
    // graph mode: capture stream, ordering events, instantiated graph
    hipStream_t cstream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t gexec = nullptr;
    // DirectLighting: two branch queues, hits and the shadow-ray queue
    float* d_dlwork = nullptr;
    uint8_t* d_dlocc = nullptr;
    size_t dl_qcap = 0, dl_scap = 0;
    uint32_t light_samples = 0;  // sum of the lights' DirectLighting samples
    bool has_specular = false, has_glossy = false;  // materials that continue DirectLighting rays
};

namespace {

void free_scene(lr_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    if (s->cstream) {
        (void)hipStreamSynchronize(s->cstream);
        s->ctx->release(s->cstream);
    }
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->ev_in) (void)hipEventDestroy(s->ev_in);
    if (s->ev_out) (void)hipEventDestroy(s->ev_out);
    if (s->cstream) (void)hipStreamDestroy(s->cstream);
    void* ptrs[] = {s->d_prim_mat, s->d_prim_geom, s->d_prim_tan, s->d_corner, s->d_mats,   s->d_spectra,
                    s->d_table,    s->d_lights,    s->d_obs,      s->d_work,   s->d_counts, s->d_gwork,
                    s->d_dlwork,   s->d_dlocc,     s->d_light_tris, s->d_halton_perm, s->d_halton_off};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete s;
}

GeomArgs geom_args(const lr_scene* sc) { return GeomArgs{sc->d_prim_geom, sc->d_prim_tan, sc->d_corner}; }

int build_observer(lr_scene* sc, float lo, float hi) {
    {
        std::lock_guard<std::mutex> g(g_obs_mu);
        if (sc->d_obs && sc->obs_lo == lo && sc->obs_hi == hi && sc->obs_version == g_obs_version) return LR_OK;
    }
    std::vector<double> w, x, y, z;
    uint64_t version = 0;
    {
        std::lock_guard<std::mutex> g(g_obs_mu);
        version = g_obs_version;
        const size_t n = g_obs_w.size();
        if (n < 2) return lr::fail(LR_E_INVALID, "no standard observer: call lr_set_standard_observer first");
        auto sens = [&](double wl, int a) -> double {
            const float* it = std::upper_bound(g_obs_w.data(), g_obs_w.data() + n, (float)wl);
            if (it == g_obs_w.data()) return 0.0;
            if (it == g_obs_w.data() + n) return wl == g_obs_w[n - 1] ? g_obs_xyz[3 * (n - 1) + a] : 0.0;
            const size_t k = (size_t)(it - g_obs_w.data()) - 1;
            const double t = (wl - g_obs_w[k]) / ((double)g_obs_w[k + 1] - g_obs_w[k]);
            return g_obs_xyz[3 * k + a] + t * ((double)g_obs_xyz[3 * (k + 1) + a] - g_obs_xyz[3 * k + a]);
        };
        const double a = std::max<double>(lo, g_obs_w[0]), b = std::min<double>(hi, g_obs_w[n - 1]);
        if (!(b > a)) return lr::fail(LR_E_INVALID, "wavelength domain does not overlap the observer");
        w.push_back(a);
        for (size_t k = 0; k < n; ++k)
            if (g_obs_w[k] > a && g_obs_w[k] < b) w.push_back(g_obs_w[k]);
        w.push_back(b);
        for (double wl : w) {
            x.push_back(sens(wl, 0));
            y.push_back(sens(wl, 1));
            z.push_back(sens(wl, 2));
        }
    }
    const size_t n = w.size();
    std::vector<double> cdf(n, 0.0);
    for (size_t k = 0; k + 1 < n; ++k)
        cdf[k + 1] = cdf[k] + (x[k] + y[k] + z[k] + x[k + 1] + y[k + 1] + z[k + 1]) * (w[k + 1] - w[k]);
    double Y = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double wp = w[k > 0 ? k - 1 : 0], wn = w[k + 1 < n ? k + 1 : n - 1];
        Y += y[k] * (wn - wp) / 2.0;
    }
    if (!(Y > 0.0) || !(cdf[n - 1] > 0.0)) return lr::fail(LR_E_INVALID, "observer has no sensitivity in domain");
    std::vector<float> buf(8 * n);
    float* W = buf.data();
    float* CDF = W + n;
    float* XYZ = CDF + n;
    float* DXYZ = XYZ + 3 * n;
    for (size_t k = 0; k < n; ++k) {
        W[k] = (float)w[k];
        CDF[k] = (float)(cdf[k] / cdf[n - 1]);
        XYZ[3 * k + 0] = (float)(x[k] / Y);
        XYZ[3 * k + 1] = (float)(y[k] / Y);
        XYZ[3 * k + 2] = (float)(z[k] / Y);
    }
    CDF[n - 1] = 1.0f;
    for (size_t k = 0; k < n; ++k) {
        for (int a = 0; a < 3; ++a) {
            DXYZ[3 * k + a] = k + 1 < n ? (XYZ[3 * (k + 1) + a] - XYZ[3 * k + a]) / (W[k + 1] - W[k]) : 0.0f;
        }
    }
    if (sc->d_obs) LR_HIP_CHECK(hipFree(sc->d_obs));
    sc->d_obs = nullptr;
    LR_HIP_CHECK(hipMalloc(&sc->d_obs, sizeof(float) * buf.size()));
    LR_HIP_CHECK(hipMemcpy(sc->d_obs, buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
    sc->obs_lo = lo;
    sc->obs_hi = hi;
    sc->obs_n = (uint32_t)n;
    sc->obs_version = version;
    return LR_OK;
}

// Guided-mode work memory: per path inputs [8] + mixture [48]; per record
// inputs [8], direction [3], aux [4], weight [1]; per path the roulette
// reference and weight [2].
int ensure_guide_work(lr_scene* sc, size_t paths, size_t records, hipStream_t s) {
    if (sc->d_gwork && sc->gcap >= paths && sc->grec >= records) return LR_OK;
    LR_HIP_CHECK(hipStreamSynchronize(s));
    paths = std::max(paths, sc->gcap);
    records = std::max(records, sc->grec);
    // capacities drop to zero with the buffer, so a failed allocation below
    // leaves a state that the next call grows again instead of trusting
    if (sc->d_gwork) LR_HIP_CHECK(hipFree(sc->d_gwork));
    sc->d_gwork = nullptr;
    sc->gcap = sc->grec = 0;
    const size_t words = 58 * paths + 16 * records;
    LR_HIP_CHECK(hipMalloc(&sc->d_gwork, sizeof(float) * words));
    sc->gcap = paths;
    sc->grec = records;
    return LR_OK;
}

int ensure_work(lr_scene* sc, size_t paths, hipStream_t s) {
    if (sc->d_work && sc->d_counts && sc->cap >= paths) return LR_OK;
    LR_HIP_CHECK(hipStreamSynchronize(s));
    if (sc->d_work) LR_HIP_CHECK(hipFree(sc->d_work));
    sc->d_work = nullptr;
    sc->cap = 0;  // a failed allocation below leaves no stale capacity behind
    // path state 4, two queues 2 * 9, hits 4 (32-bit words per path)
    LR_HIP_CHECK(hipMalloc(&sc->d_work, sizeof(float) * 26 * paths));
    if (!sc->d_counts) LR_HIP_CHECK(hipMalloc(&sc->d_counts, 64));
    sc->cap = paths;
    return LR_OK;
}

}  // namespace

namespace lr {

// lr_spectrum descriptions -> device records and their value table (sampled
// wavelengths then values, binned values), with LiAR's Reflectant scale
// (spectral.cpp:200-207) and the BlackBody peak normalisation.
int build_spectra(const lr_spectrum* spectra, uint32_t n_spectra, std::vector<DevSpectrum>& dspec,
                  std::vector<float>& table) {
    dspec.assign(n_spectra, DevSpectrum{});
    table.clear();
    for (uint32_t i = 0; i < n_spectra; ++i) {
        const lr_spectrum& s = spectra[i];
        DevSpectrum& d = dspec[i];
        std::memset(&d, 0, sizeof(d));
        d.kind = s.kind;
        d.refl_scale = 1.0f;
        switch (s.kind) {
        case LR_SPEC_CONSTANT:
            d.value = s.value;
            break;
        case LR_SPEC_SAMPLED: {
            if (s.count < 2 || !s.w || !s.v) return lr::fail(LR_E_INVALID, "sampled spectrum needs >= 2 points");
            for (uint32_t k = 0; k + 1 < s.count; ++k)
                if (!(s.w[k + 1] > s.w[k])) return lr::fail(LR_E_INVALID, "sampled spectrum wavelengths must increase");
            d.count = s.count;
            d.offset = (uint32_t)table.size();
            table.insert(table.end(), s.w, s.w + s.count);
            table.insert(table.end(), s.v, s.v + s.count);
            // Reflectant divisor: the values' maximum, or the caller's (value > 0)
            const float mx = s.value > 0.0f ? s.value : *std::max_element(s.v, s.v + s.count);
            d.refl_scale = mx > 1.0f ? 1.0f / mx : 1.0f;
            break;
        }
        case LR_SPEC_BINNED: {
            if (s.count < 1 || !s.v || !(s.hi > s.lo)) return lr::fail(LR_E_INVALID, "binned spectrum needs bins and lo < hi");
            d.count = s.count;
            d.lo = s.lo;
            d.hi = s.hi;
            d.offset = (uint32_t)table.size();
            table.insert(table.end(), s.v, s.v + s.count);
            const float mx = *std::max_element(s.v, s.v + s.count);
            d.refl_scale = mx > 1.0f ? 1.0f / mx : 1.0f;
            break;
        }
        case LR_SPEC_BLACKBODY: {
            if (!(s.temperature > 0.0f)) return lr::fail(LR_E_INVALID, "blackbody temperature must be > 0");
            d.bb_t = s.temperature;
            d.bb_k = s.scale / host_planck(2.897771955e-3f / s.temperature, s.temperature);
            break;
        }
        default:
            return lr::fail(LR_E_INVALID, "unknown spectrum kind " + std::to_string(s.kind));
        }
    }
    return LR_OK;
}

}  // namespace lr

extern "C" {

int lr_set_standard_observer(const float* w, const float* xyz, uint32_t n) {
    if (!w || !xyz || n < 2) return lr::fail(LR_E_INVALID, "observer needs at least two wavelengths");
    for (uint32_t k = 0; k + 1 < n; ++k)
        if (!(w[k + 1] > w[k])) return lr::fail(LR_E_INVALID, "observer wavelengths must be strictly increasing");
    std::lock_guard<std::mutex> g(g_obs_mu);
    g_obs_w.assign(w, w + n);
    ++g_obs_version;
    g_obs_xyz.assign(xyz, xyz + 3 * (size_t)n);
    return LR_OK;
}

int lr_scene_create(lr_context* ctx, lr_qbvh* bvh, const uint32_t* prim_material, const lr_material* materials,
                    uint32_t n_materials, const lr_spectrum* spectra, uint32_t n_spectra, const lr_light* lights,
                    uint32_t n_lights, lr_scene** out) {
    if (!ctx || !bvh || !out) return lr::fail(LR_E_INVALID, "NULL context, geometry or out");
    *out = nullptr;
    if (bvh->ctx != ctx) return lr::fail(LR_E_INVALID, "geometry belongs to another context");
    const uint32_t n_prims = bvh->info.n_prims;
    if (n_prims && !prim_material) return lr::fail(LR_E_INVALID, "prim_material is NULL");
    if ((n_materials && !materials) || (n_spectra && !spectra) || (n_lights && !lights))
        return lr::fail(LR_E_INVALID, "NULL materials / spectra / lights array");
    {
        std::lock_guard<std::mutex> g(g_obs_mu);
        if (g_obs_w.size() < 2) return lr::fail(LR_E_INVALID, "no standard observer: call lr_set_standard_observer first");
    }
    auto spec_ok = [&](uint32_t i) { return i == LR_NONE || i < n_spectra; };
    // spectra -> device table
    std::vector<DevSpectrum> dspec;
    std::vector<float> table;
    if (int rc = lr::build_spectra(spectra, n_spectra, dspec, table)) return rc;
    std::vector<DevMaterial> dmat(n_materials);
    for (uint32_t i = 0; i < n_materials; ++i) {
        const lr_material& m = materials[i];
        if (m.type > LR_MAT_COOK_TORRANCE_BECKMANN) return lr::fail(LR_E_INVALID, "unknown material type");
        if (!spec_ok(m.reflectance) || !spec_ok(m.transmittance) || !spec_ok(m.eta) || !spec_ok(m.kappa) ||
            !spec_ok(m.emission))
            return lr::fail(LR_E_INVALID, "material " + std::to_string(i) + " references a missing spectrum");
        if (m.light != LR_NONE && m.light >= n_lights)
            return lr::fail(LR_E_INVALID, "material " + std::to_string(i) + " references a missing light");
        if (!(m.fuzz >= 0.0f) || !std::isfinite(m.fuzz))
            return lr::fail(LR_E_INVALID, "material " + std::to_string(i) + ": fuzz must be finite and >= 0");
        dmat[i] = DevMaterial{m.type,     m.reflectance, m.transmittance, m.eta,        m.kappa,
                              m.emission, m.light,       m.roughness,     m.roughness_v, m.fuzz};
    }
    std::vector<uint32_t> pm(prim_material, prim_material + n_prims);
    for (uint32_t i = 0; i < n_prims; ++i)
        if (pm[i] >= n_materials) return lr::fail(LR_E_INVALID, "prim " + std::to_string(i) + " has no material");
    // per-prim shading geometry by primitive id (GeomArgs): triangle unit
    // normal + corner-normal record (none yet) and its BSDF frame's fixed
    // tangent (intersection_context.cpp:197-223 with dP/du, dP/dv = the
    // edges); sphere (c, r)
    std::vector<float4> geom(n_prims), tan(n_prims, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    const float kNoRecord = lr::as_f32(LR_NONE);
    for (const lr_prim& q : bvh->host.prims) {
        const uint32_t id = lr::as_u32(q.a[3]);
        if (lr::as_u32(q.b[3]) == LR_PRIM_TRIANGLE) {
            const double e1[3] = {q.b[0], q.b[1], q.b[2]}, e2[3] = {q.c[0], q.c[1], q.c[2]};
            double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (len > 0) {
                for (double& c : n) c /= len;
            } else {
                n[0] = n[1] = 0.0;
                n[2] = 1.0;
            }
            geom[id] = make_float4((float)n[0], (float)n[1], (float)n[2], kNoRecord);
            double u[3], v[3];
            const double du = e1[0] * n[0] + e1[1] * n[1] + e1[2] * n[2];
            const double dv = e2[0] * n[0] + e2[1] * n[1] + e2[2] * n[2];
            for (int k = 0; k < 3; ++k) {
                u[k] = e1[k] - du * n[k];
                v[k] = e2[k] - dv * n[k];
            }
            const double su = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            const double sv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (std::max(su, sv) >= 1e-3) {
                const double* a = su >= sv ? u : v;
                const double la = std::sqrt(su >= sv ? su : sv);
                tan[id] = make_float4((float)(a[0] / la), (float)(a[1] / la), (float)(a[2] / la),
                                      lr::as_f32(su >= sv ? 1u : 2u));
            }
        } else {
            geom[id] = make_float4(q.a[0], q.a[1], q.a[2], q.b[0]);
        }
    }
    // lights: geometry, power (light_area.cpp:290-294, light_sky.cpp:93-134)
    const lr_qbvh_info& I = bvh->info;
    double R = 0.0;
    for (int k = 0; k < 3; ++k) R += 0.25 * ((double)I.scene_hi[k] - I.scene_lo[k]) * ((double)I.scene_hi[k] - I.scene_lo[k]);
    R = std::sqrt(R);
    const double far_away = R > 0 ? 1e3 * R : 1e3;
    std::vector<DevLight> dl(n_lights);
    std::vector<double> power(n_lights, 0.0);
    std::vector<float4> light_tris;   // mesh lights' triangles
    std::vector<uint32_t> prim_rec;   // primitive id -> record index (mesh lights only)
    for (uint32_t i = 0; i < n_lights; ++i) {
        const lr_light& L = lights[i];
        DevLight& d = dl[i];
        std::memset(&d, 0, sizeof(d));
        d.type = L.type;
        d.radiance = L.radiance;
        d.double_sided = L.double_sided ? 1u : 0u;
        d.samples = L.samples ? L.samples : ((L.type == LR_LIGHT_PARALLELOGRAM || L.type == LR_LIGHT_MESH) ? 9u : 1u);
        if (d.samples > 64) return lr::fail(LR_E_INVALID, "light " + std::to_string(i) + ": at most 64 samples");
        if (L.radiance >= n_spectra) return lr::fail(LR_E_INVALID, "light " + std::to_string(i) + " has no radiance");
        const double lum = host_luminance(spectra[L.radiance]);
        if (L.type == LR_LIGHT_PARALLELOGRAM) {
            const double ax = L.dx[0], ay = L.dx[1], az = L.dx[2], bx = L.dy[0], by = L.dy[1], bz = L.dy[2];
            const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            const double area = std::sqrt(cx * cx + cy * cy + cz * cz);
            if (!(area > 0)) return lr::fail(LR_E_INVALID, "degenerate parallelogram light");
            for (int k = 0; k < 3; ++k) {
                d.o[k] = L.origin[k];
                d.dx[k] = L.dx[k];
                d.dy[k] = L.dy[k];
            }
            d.n[0] = (float)(cx / area);
            d.n[1] = (float)(cy / area);
            d.n[2] = (float)(cz / area);
            d.area = (float)area;
            // dual basis: du = (dy x n) / (dx . (dy x n)), dv = (n x dx) / (dy . (n x dx))
            const double nx = cx / area, ny = cy / area, nz = cz / area;
            const double ux = by * nz - bz * ny, uy = bz * nx - bx * nz, uz = bx * ny - by * nx;
            const double ud = ax * ux + ay * uy + az * uz;
            const double vx = ny * az - nz * ay, vy = nz * ax - nx * az, vz = nx * ay - ny * ax;
            const double vd = bx * vx + by * vy + bz * vz;
            d.du[0] = (float)(ux / ud);
            d.du[1] = (float)(uy / ud);
            d.du[2] = (float)(uz / ud);
            d.dv[0] = (float)(vx / vd);
            d.dv[1] = (float)(vy / vd);
            d.dv[2] = (float)(vz / vd);
            power[i] = 3.14159265358979 * area * lum * (L.double_sided ? 2.0 : 1.0);
        } else if (L.type == LR_LIGHT_MESH) {
            // TriangleMesh surface sampling (triangle_mesh.cpp:69-85,414-444):
            // triangles picked by a cumulative area table, uniform inside
            if (L.n_triangles == 0 || L.n_triangles > 4096 || L.first_triangle > bvh->info.n_triangles ||
                L.n_triangles > bvh->info.n_triangles - L.first_triangle)
                return lr::fail(LR_E_INVALID, "light " + std::to_string(i) +
                                                  ": mesh triangles out of range (1 .. 4096 of the scene's triangles)");
            if (prim_rec.empty()) {
                prim_rec.assign(n_prims, LR_NONE);
                for (size_t r = 0; r < bvh->host.prims.size(); ++r)
                    prim_rec[lr::as_u32(bvh->host.prims[r].a[3])] = (uint32_t)r;
            }
            d.tri0 = (uint32_t)(light_tris.size() / 3);
            d.ntri = L.n_triangles;
            std::vector<double> acc(L.n_triangles);
            double total = 0.0;
            for (uint32_t k = 0; k < L.n_triangles; ++k) {
                const lr_prim& q = bvh->host.prims[prim_rec[L.first_triangle + k]];
                const double cx = (double)q.b[1] * q.c[2] - (double)q.b[2] * q.c[1];
                const double cy = (double)q.b[2] * q.c[0] - (double)q.b[0] * q.c[2];
                const double cz = (double)q.b[0] * q.c[1] - (double)q.b[1] * q.c[0];
                total += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
                acc[k] = total;
            }
            if (!(total > 0)) return lr::fail(LR_E_INVALID, "light " + std::to_string(i) + ": mesh has no area");
            for (uint32_t k = 0; k < L.n_triangles; ++k) {
                const lr_prim& q = bvh->host.prims[prim_rec[L.first_triangle + k]];
                const float cdf = k + 1 == L.n_triangles ? 1.0f : (float)(acc[k] / total);
                light_tris.push_back(make_float4(q.a[0], q.a[1], q.a[2], cdf));
                light_tris.push_back(make_float4(q.b[0], q.b[1], q.b[2], 0.0f));
                light_tris.push_back(make_float4(q.c[0], q.c[1], q.c[2], 0.0f));
            }
            d.area = (float)total;
            power[i] = 3.14159265358979 * total * lum * (L.double_sided ? 2.0 : 1.0);
        } else if (L.type == LR_LIGHT_SKY) {
            power[i] = lum * 3.14159265358979 * far_away * far_away * 4.0 * 3.14159265358979;
        } else if (L.type == LR_LIGHT_SUN) {
            double dx = L.direction[0], dy = L.direction[1], dz = L.direction[2];
            const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (!(len > 0) || !(L.angle > 0 && L.angle < 3.14159265)) return lr::fail(LR_E_INVALID, "bad sun light");
            d.n[0] = (float)(dx / len);
            d.n[1] = (float)(dy / len);
            d.n[2] = (float)(dz / len);
            d.cos_max = (float)std::cos(0.5 * L.angle);
            const double solid = 2.0 * 3.14159265358979 * (1.0 - std::cos(0.5 * L.angle));
            d.inv_solid = (float)(1.0 / solid);
            power[i] = lum * solid * 3.14159265358979 * far_away * far_away;
        } else {
            return lr::fail(LR_E_INVALID, "unknown light type");
        }
    }
    double total = 0.0;
    for (double p : power) total += p;
    double acc = 0.0;
    for (uint32_t i = 0; i < n_lights; ++i) {
        acc += power[i];
        dl[i].cdf = total > 0 ? (float)(acc / total) : (float)(i + 1) / n_lights;
        dl[i].sel_pdf = total > 0 ? (float)(power[i] / total) : 1.0f / n_lights;
    }
    if (n_lights) dl[n_lights - 1].cdf = 1.0f;

    lr_scene* sc = new (std::nothrow) lr_scene();
    if (!sc) return lr::fail(LR_E_NOMEM, "out of host memory");
    sc->ctx = ctx;
    sc->bvh = bvh;
    sc->n_triangles = bvh->info.n_triangles;
    sc->n_lights = n_lights;
    for (const DevLight& d : dl) sc->light_samples += d.samples;
    for (const DevMaterial& m : dmat) {
        if (m.type == LR_MAT_MIRROR || m.type == LR_MAT_CONDUCTOR || m.type == LR_MAT_DIELECTRIC) sc->has_specular = true;
        if (m.type == LR_MAT_COOK_TORRANCE || m.type == LR_MAT_COOK_TORRANCE_BECKMANN) sc->has_glossy = true;
    }
    int rc = lr::bind(ctx);
    if (!rc) rc = upload(pm, &sc->d_prim_mat);
    if (!rc) rc = upload(geom, &sc->d_prim_geom);
    if (!rc) rc = upload(tan, &sc->d_prim_tan);
    if (!rc) rc = upload(dmat, &sc->d_mats);
    if (!rc) rc = upload(dspec, &sc->d_spectra);
    if (!rc) rc = upload(table, &sc->d_table);
    if (!rc) rc = upload(dl, &sc->d_lights);
    if (!rc) rc = upload(light_tris, &sc->d_light_tris);
    if (rc) {
        free_scene(sc);
        return rc;
    }
    *out = sc;
    return LR_OK;
}

int lr_scene_set_normals(lr_scene* sc, const float* cn, uint32_t n_triangles) {
    if (!sc) return lr::fail(LR_E_INVALID, "NULL scene");
    if (n_triangles != sc->n_triangles)
        return lr::fail(LR_E_INVALID, "lr_scene_set_normals: " + std::to_string(n_triangles) +
                                          " triangles given, the scene has " + std::to_string(sc->n_triangles));
    if (int rc = lr::bind(sc->ctx)) return rc;
    std::vector<float4> geom(sc->bvh->info.n_prims);
    if (!geom.empty())
        LR_HIP_CHECK(hipMemcpy(geom.data(), sc->d_prim_geom, sizeof(float4) * geom.size(), hipMemcpyDeviceToHost));
    // records (n0, e1.x) (n1, e1.y) (n2, e1.z) (e2, 0) for triangles with
    // normals; edges as the primitive records hold them
    std::vector<float4> rec;
    std::vector<uint32_t> slot(n_triangles, LR_NONE);
    if (cn) {
        for (uint32_t t = 0; t < n_triangles; ++t) {
            const float* n = cn + 9 * (size_t)t;
            bool any = false;
            for (int k = 0; k < 9; ++k) {
                if (!std::isfinite(n[k]))
                    return lr::fail(LR_E_INVALID, "lr_scene_set_normals: non-finite normal of triangle " +
                                                      std::to_string(t));
                any = any || n[k] != 0.0f;
            }
            if (any) {
                slot[t] = (uint32_t)(rec.size() / 4);
                rec.resize(rec.size() + 4);
            }
        }
        for (const lr_prim& q : sc->bvh->host.prims) {
            const uint32_t id = lr::as_u32(q.a[3]);
            if (lr::as_u32(q.b[3]) != LR_PRIM_TRIANGLE || slot[id] == LR_NONE) continue;
            const float* n = cn + 9 * (size_t)id;
            float4* r = rec.data() + 4 * (size_t)slot[id];
            r[0] = make_float4(n[0], n[1], n[2], q.b[0]);
            r[1] = make_float4(n[3], n[4], n[5], q.b[1]);
            r[2] = make_float4(n[6], n[7], n[8], q.b[2]);
            r[3] = make_float4(q.c[0], q.c[1], q.c[2], 0.0f);
        }
    }
    for (uint32_t t = 0; t < n_triangles; ++t) geom[t].w = lr::as_f32(slot[t]);
    LR_HIP_CHECK(hipDeviceSynchronize());
    if (sc->d_corner) LR_HIP_CHECK(hipFree(sc->d_corner));
    sc->d_corner = nullptr;
    if (int rc = upload(rec, &sc->d_corner)) return rc;
    if (!geom.empty())
        LR_HIP_CHECK(hipMemcpy(sc->d_prim_geom, geom.data(), sizeof(float4) * geom.size(), hipMemcpyHostToDevice));
    return LR_OK;
}

int lr_scene_destroy(lr_scene* sc) {
    free_scene(sc);
    return LR_OK;
}

}  // extern "C"

// Camera frame of an lr_camera (perspective_camera.cpp:126-128, 259, 466-468).
static int make_camera(const lr_camera* cam, DevCamera& C) {
    if (!(cam->fov > 0.0f && cam->fov < 3.14159f) || !(cam->aspect > 0.0f))
        return lr::fail(LR_E_INVALID, "camera fov must be in (0, pi) and aspect > 0");
    double d[3] = {cam->direction[0], cam->direction[1], cam->direction[2]};
    double sk[3] = {cam->sky[0], cam->sky[1], cam->sky[2]};
    auto nrm = [](double* v) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (l > 0)
            for (int k = 0; k < 3; ++k) v[k] /= l;
        return l;
    };
    if (!(nrm(d) > 0)) return lr::fail(LR_E_INVALID, "camera direction is zero");
    double r[3] = {d[1] * sk[2] - d[2] * sk[1], d[2] * sk[0] - d[0] * sk[2], d[0] * sk[1] - d[1] * sk[0]};
    if (!(nrm(r) > 0)) return lr::fail(LR_E_INVALID, "camera sky is parallel to its direction");
    double dn[3] = {d[1] * r[2] - d[2] * r[1], d[2] * r[0] - d[0] * r[2], d[0] * r[1] - d[1] * r[0]};
    nrm(dn);
    const double width = 0.036, height = width / cam->aspect;
    const double focal = (width / 2) / std::tan(cam->fov / 2);
    for (int k = 0; k < 3; ++k) {
        C.pos[k] = cam->position[k];
        C.dir[k] = (float)(d[k] * focal);
        C.right[k] = (float)(r[k] * width);
        C.down[k] = (float)(dn[k] * height);
        C.dir_n[k] = (float)d[k];
    }
    C.near_limit = cam->near_limit;
    return LR_OK;
}

// Validated kernel parameters of an lr_render_params.
// The Halton sampler's sub-sequence primes (17, 19, ..., 1889: the 284
// primes after 13) and their digit permutations as LiAR's
// Halton::initialize draws them (halton.cpp:102-118): one default-seeded
// std::mt19937 shuffling 0 .. p-1 for each prime in order.
constexpr uint32_t kHaltonPrimes = 284;

struct HaltonTables {
    std::vector<uint16_t> perm;  // concatenated permutations
    std::vector<uint32_t> off;   // kHaltonPrimes + 1 offsets (off[k+1] - off[k] = prime k)
};

const HaltonTables& halton_tables() {
    static HaltonTables T;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<uint32_t> primes;
        for (uint32_t n = 17; primes.size() < kHaltonPrimes; ++n) {
            bool prime = true;
            for (uint32_t d = 2; d * d <= n; ++d)
                if (n % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) primes.push_back(n);
        }
        std::mt19937 rng;
        T.off.push_back(0);
        std::vector<size_t> v;
        for (uint32_t p : primes) {
            v.resize(p);
            for (uint32_t k = 0; k < p; ++k) v[k] = k;
            std::shuffle(v.begin(), v.end(), rng);
            for (size_t x : v) T.perm.push_back((uint16_t)x);
            T.off.push_back((uint32_t)T.perm.size());
        }
    });
    return T;
}

// Device copies of the Halton tables for the scene's device (made once).
int ensure_halton(lr_scene* sc) {
    if (sc->d_halton_perm) return LR_OK;
    const HaltonTables& T = halton_tables();
    LR_HIP_CHECK(hipMalloc(&sc->d_halton_perm, sizeof(uint16_t) * T.perm.size()));
    LR_HIP_CHECK(hipMemcpy(sc->d_halton_perm, T.perm.data(), sizeof(uint16_t) * T.perm.size(), hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMalloc(&sc->d_halton_off, sizeof(uint32_t) * T.off.size()));
    LR_HIP_CHECK(hipMemcpy(sc->d_halton_off, T.off.data(), sizeof(uint32_t) * T.off.size(), hipMemcpyHostToDevice));
    return LR_OK;
}

static int make_params(const lr_render_params& R, Params& P) {
    if (!R.width || !R.height || !R.spp || !R.wavelengths)
        return lr::fail(LR_E_INVALID, "width, height, spp and wavelengths must be > 0");
    if (!(R.rr_survival > 0.0f && R.rr_survival <= 1.0f)) return lr::fail(LR_E_INVALID, "rr_survival must be in (0, 1]");
    if (!(R.lambda_max > R.lambda_min)) return lr::fail(LR_E_INVALID, "lambda_max must exceed lambda_min");
    if (!(R.max_sample_luminance >= 0.0f)) return lr::fail(LR_E_INVALID, "max_sample_luminance must be >= 0");
    P.width = R.width;
    P.height = R.height;
    P.spp = R.spp;
    P.nl = R.wavelengths;
    P.max_gen = R.max_depth ? R.max_depth : 64u;
    P.rr_depth = R.rr_depth;
    P.rr_p = R.rr_survival;
    P.max_lum = R.max_sample_luminance;
    P.seed_lo = (uint32_t)R.seed;
    P.seed_hi = (uint32_t)(R.seed >> 32);
    const uint32_t q = (uint32_t)std::lround(std::sqrt((double)R.spp));
    P.strata = q * q == R.spp ? q : 0u;
    if (R.sampler > LR_SAMPLER_HALTON) return lr::fail(LR_E_INVALID, "unknown sampler");
    P.halton = R.sampler == LR_SAMPLER_HALTON ? 1u : 0u;
    P.hgens = P.halton ? R.sampler_generations : 0u;
    if (P.hgens * 10u > kHaltonPrimes) return lr::fail(LR_E_INVALID, "Halton: at most 28 sampler generations");
    P.hoff = P.halton ? R.sample_offset : 0u;
    P.perm = nullptr;
    P.perm_off = nullptr;
    return LR_OK;
}

// Device view of a scene (its observer table must be built for the call's domain).
static DevScene dev_scene(const lr_scene* sc) {
    DevScene S;
    S.nodes = sc->bvh->d_nodes;
    S.prims = sc->bvh->d_prims;
    S.prim_mat = sc->d_prim_mat;
    S.mats = sc->d_mats;
    S.spectra = sc->d_spectra;
    S.table = sc->d_table;
    S.lights = sc->d_lights;
    S.light_tris = sc->d_light_tris;
    S.n_lights = sc->n_lights;
    S.n_triangles = sc->n_triangles;
    S.obs.n = sc->obs_n;
    S.obs.w = sc->d_obs;
    S.obs.cdf = sc->d_obs + sc->obs_n;
    S.obs.xyz = sc->d_obs + 2 * sc->obs_n;
    S.obs.dxyz = sc->d_obs + 5 * sc->obs_n;
    return S;
}

// Launches a captured graph through a cached executable: updated in place
// when the topology allows it, re-instantiated otherwise.
static int launch_graph(hipGraphExec_t& exec, hipGraph_t graph, hipStream_t s) {
    if (exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(exec, graph, &err_node, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            LR_HIP_CHECK(hipGraphExecDestroy(exec));
            exec = nullptr;
        }
    }
    if (!exec) LR_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    LR_HIP_CHECK(hipGraphLaunch(exec, s));
    return LR_OK;
}

static int render_impl(lr_scene* sc, const lr_camera* cam, const lr_render_params* prm, lr_guide* guide,
                       const lr_guide_params* gparams, float* film_xyz, float* film_w, lr_render_stats* stats,
                       void* stream) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!sc || !cam || !prm) return lr::fail(LR_E_INVALID, "NULL scene, camera or params");
    if (!film_xyz || !film_w) return lr::fail(LR_E_INVALID, "NULL film buffer");
    const lr_render_params& R = *prm;
    if (!R.width || !R.height || !R.spp || !R.wavelengths)
        return lr::fail(LR_E_INVALID, "width, height, spp and wavelengths must be > 0");
    if (int rc = lr::bind(sc->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (int rc = build_observer(sc, R.lambda_min, R.lambda_max)) return rc;

    DevCamera C;
    if (int rc = make_camera(cam, C)) return rc;
    Params P;
    if (int rc = make_params(R, P)) return rc;
    if (P.halton) {
        if (guide) return lr::fail(LR_E_UNSUPPORTED, "the Halton sampler is for plain path tracing (lr_render)");
        if (int rc = ensure_halton(sc)) return rc;
        P.perm = sc->d_halton_perm;
        P.perm_off = sc->d_halton_off;
    }
    const uint64_t ppp = (uint64_t)R.spp * R.wavelengths;
    const uint64_t n_pix = (uint64_t)R.width * R.height;
    // default queue tile: 32M paths (every bounce launch's tail is paid once
    // per chunk; 3.3 GB of queues), 4M for guided renders that train, whose
    // chunk is also the training cadence; never more than the render's own
    // paths.  Without training the tile changes no result (whole pixels per
    // chunk, random numbers keyed by the global path id).
    const bool trains = guide && gparams && gparams->train;
    const uint64_t chunk_req = R.chunk_paths ? R.chunk_paths : (trains ? (1u << 22) : (1u << 25));
    const uint64_t pix_per_chunk = std::min<uint64_t>(std::max<uint64_t>(1, chunk_req / ppp), std::max<uint64_t>(1, n_pix));
    const uint64_t cap64 = pix_per_chunk * ppp;
    if (cap64 >= (1ull << 31)) return lr::fail(LR_E_INVALID, "spp * wavelengths too large for one chunk");
    const uint32_t cap = (uint32_t)cap64;
    if (int rc = ensure_work(sc, cap, s)) return rc;
    const bool guided = guide != nullptr;
    GuideArgs GA;
    std::memset(&GA, 0, sizeof(GA));
    float* g_in = nullptr;
    float* g_mix = nullptr;
    float* r_w = nullptr;
    uint32_t rec_cap = 0, batch = 0, guide_depth = 0;
    if (guided) {
        if (!gparams) return lr::fail(LR_E_INVALID, "guided render needs lr_guide_params");
        if (!(gparams->selection >= 0.0f && gparams->selection < 1.0f))
            return lr::fail(LR_E_INVALID, "guide selection must be in [0, 1)");
        rec_cap = gparams->max_records ? gparams->max_records : (1u << 22);
        batch = gparams->train_batch ? gparams->train_batch : 65536u;
        if (int rc = ensure_guide_work(sc, cap, rec_cap, s)) return rc;
        g_in = sc->d_gwork;
        g_mix = g_in + 8 * (size_t)cap;
        GA.in = g_in;
        GA.mix = g_mix;
        GA.mix_ld = cap;
        GA.sel = gparams->selection;
        GA.light_share = gparams->light_share;
        if (!(GA.light_share >= 0.0f && GA.light_share < 1.0f))
            return lr::fail(LR_E_INVALID, "guide light_share must be in [0, 1)");
        GA.learned = gparams->learned_selection ? 1u : 0u;
        GA.balance = gparams->balance_heuristic ? 1u : 0u;
        GA.image_weight = gparams->image_weighting ? 1u : 0u;
        GA.record_depth = gparams->record_depth ? gparams->record_depth : 0xFFFFFFFFu;
        guide_depth = gparams->guide_depth ? gparams->guide_depth : 0xFFFFFFFFu;
        const lr_qbvh_info& I = sc->bvh->info;
        for (int k = 0; k < 3; ++k) {
            GA.lo[k] = I.scene_lo[k];
            const float ext = I.scene_hi[k] - I.scene_lo[k];
            GA.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
        }
        GA.lam_lo = R.lambda_min;
        GA.inv_lam_ext = 1.0f / (R.lambda_max - R.lambda_min);
        GA.record = gparams->train ? 1u : 0u;
        GA.cap = rec_cap;
        GA.count = sc->d_counts + 8;
        GA.r_in = g_mix + 48 * (size_t)cap;
        GA.r_dir = GA.r_in + 8 * (size_t)rec_cap;
        GA.r_aux = GA.r_dir + 3 * (size_t)rec_cap;
        r_w = GA.r_aux + 4 * (size_t)rec_cap;
        GA.est = r_w + rec_cap;
        GA.boost = GA.est + cap;
        GA.rr = gparams->learned_rr ? 1u : 0u;
        GA.rr_scale = gparams->rr_scale > 0.0f ? gparams->rr_scale : 1.0f;
        if (!(GA.rr_scale < INFINITY)) return lr::fail(LR_E_INVALID, "rr_scale must be finite");
        // least survival probability: a survivor's weight is at most 1 / rr_min
        GA.rr_min = gparams->rr_min_survival > 0.0f ? gparams->rr_min_survival : 0.05f;
        GA.rr_start = gparams->rr_start;
        if (!(GA.rr_min <= 1.0f)) return lr::fail(LR_E_INVALID, "rr_min_survival must be in (0, 1]");
    }

    const DevScene S = dev_scene(sc);

    float* w = sc->d_work;
    PathState ps{w, w + cap, w + 2 * (size_t)cap, w + 3 * (size_t)cap};
    Queue q[2];
    for (int b = 0; b < 2; ++b) {
        float* base = w + (4 + 9 * b) * (size_t)cap;
        for (int k = 0; k < 8; ++k) q[b].r[k] = base + k * (size_t)cap;
        q[b].pid = reinterpret_cast<uint32_t*>(base + 8 * (size_t)cap);
        q[b].root = nullptr;
    }
    float* hb = w + 22 * (size_t)cap;
    Hits H{reinterpret_cast<uint32_t*>(hb), hb + cap, hb + 2 * (size_t)cap, hb + 3 * (size_t)cap};
    uint32_t* counts = sc->d_counts;
    uint64_t* rays_total = reinterpret_cast<uint64_t*>(sc->d_counts + 4);
    BatchArgs NB;
    std::memset(&NB, 0, sizeof(NB));
    const bool training = guided && gparams->train && gparams->train_steps > 0;
    // the mixture is read only by the guide strategy, learned selection and
    // the roulette: without them (a call that samples with the plain
    // strategies and only records, say) the inference is skipped
    const bool infer = guided && (GA.learned || GA.sel > 0.0f || GA.rr);

    // graph mode: every launch below is captured on the scene's own stream
    // (the caller's may be the legacy null stream, which cannot capture)
    const hipStream_t caller = s;
    if (R.graph) {
        if (!sc->cstream) {
            LR_HIP_CHECK(hipStreamCreateWithFlags(&sc->cstream, hipStreamNonBlocking));
            LR_HIP_CHECK(hipEventCreateWithFlags(&sc->ev_in, hipEventDisableTiming));
            LR_HIP_CHECK(hipEventCreateWithFlags(&sc->ev_out, hipEventDisableTiming));
        }
        s = sc->cstream;
        // everything the captured launches use is allocated and queried first
        if (sc->bvh->info.n_nodes)
            if (int rc = lr::trace_reserve(sc->bvh, cap, s)) return rc;
        LR_HIP_CHECK(hipEventRecord(sc->ev_in, caller));
        LR_HIP_CHECK(hipStreamWaitEvent(s, sc->ev_in, 0));
        LR_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    }

    uint32_t chunks = 0;
    auto body = [&]() -> int {
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counts, 16u);
    for (uint64_t pix0 = 0; pix0 < n_pix; pix0 += pix_per_chunk) {
        const uint32_t npix = (uint32_t)std::min<uint64_t>(pix_per_chunk, n_pix - pix0);
        const uint32_t n = (uint32_t)(npix * ppp);
        const uint64_t path0 = pix0 * ppp;
        const dim3 blk(256), grd((n + 255) / 256);
        const dim3 ablk(kAppendBlock), agrd((n + kAppendBlock - 1) / kAppendBlock);
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counts, 1u);
        if (guided) hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counts + 8, 1u);
        hipLaunchKernelGGL(k_camera<false>, agrd, ablk, 0, s, S, C, P, path0, n, ps, q[0], counts, NB);
        for (uint32_t gen = 0; gen < P.max_gen; ++gen) {
            const int cur = gen & 1, nxt = cur ^ 1;
            hipLaunchKernelGGL(k_next_gen, dim3(1), dim3(64), 0, s, counts, (uint32_t)cur, (uint32_t)nxt, rays_total);
            lr_rays_soa rays{q[cur].r[0], q[cur].r[1], q[cur].r[2], q[cur].r[3],
                             q[cur].r[4], q[cur].r[5], q[cur].r[6], q[cur].r[7]};
            lr_hits_soa hits{H.prim, H.t, H.u, H.v};
            if (sc->bvh->info.n_nodes == 0) {
                hipLaunchKernelGGL(k_fill_miss_queue, grd, blk, 0, s, counts + cur, H);
            } else if (int rc = lr::trace_closest_queue(sc->bvh, &rays, n, counts + cur, &hits, s)) {
                return rc;
            }
            if (guided && gen < guide_depth) {
                hipLaunchKernelGGL(k_guide_inputs, grd, blk, 0, s, S, q[cur], counts + cur, H, ps, GA, g_in);
                if (infer)
                    if (int rc = lr_guide_infer_columns(guide, g_in, n, counts + cur, g_mix, cap, s)) return rc;
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<true, false>), dim3((n + kAppendBlockWide - 1) / kAppendBlockWide),
                                   dim3(kAppendBlockWide), 0, s, S, P, gen, path0, q[cur], counts + cur, H,
                                   geom_args(sc), ps, q[nxt], counts + nxt, GA, NB);
            } else {
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<false, false>), agrd, ablk, 0, s, S, P, gen, path0, q[cur], counts + cur, H,
                                   geom_args(sc), ps, q[nxt], counts + nxt, GA, NB);
            }
        }
        if (training) {
            // recorded vertices -> training weights -> Adam steps on batches drawn
            // from all of them; the record count stays on the device
            const uint32_t* nrec = counts + 8;
            hipLaunchKernelGGL(k_guide_targets, dim3(1024), blk, 0, s, rec_cap, nrec, GA.r_aux, ps.L, r_w);
            if (int rc = lr_guide_prepare_records(guide, r_w, nrec, rec_cap, gparams->weight_clamp, s)) return rc;
            for (uint32_t st = 0; st < gparams->train_steps; ++st)
                if (int rc = lr_guide_train_records(guide, GA.r_in, GA.r_dir, r_w, nrec, rec_cap, R.seed, chunks, st,
                                                    batch, s))
                    return rc;
        }
        if (P.halton)
            hipLaunchKernelGGL(k_splat_scatter, grd, blk, 0, s, S, P, path0, n, ps, film_xyz, film_w);
        else
            hipLaunchKernelGGL(k_splat_tiles, dim3(splat_blocks(npix, (uint32_t)ppp)), blk, 0, s, S, P, pix0, npix,
                               ps, film_xyz, film_w);
        LR_HIP_CHECK(hipGetLastError());
        ++chunks;
    }
    return LR_OK;
    };
    int rc = body();
    if (R.graph) {
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (!rc && ec != hipSuccess) rc = lr::fail(LR_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
        if (!rc) rc = launch_graph(sc->gexec, graph, s);
        if (graph) (void)hipGraphDestroy(graph);
        if (!rc) {
            LR_HIP_CHECK(hipEventRecord(sc->ev_out, s));
            LR_HIP_CHECK(hipStreamWaitEvent(caller, sc->ev_out, 0));
        }
        s = caller;
    }
    if (rc) return rc;
    if (stats) {
        uint64_t rays = 0;
        LR_HIP_CHECK(hipMemcpyAsync(&rays, rays_total, sizeof(rays), hipMemcpyDeviceToHost, s));
        LR_HIP_CHECK(hipStreamSynchronize(s));
        stats->paths = n_pix * ppp;
        stats->rays = rays;
        stats->chunks = chunks;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return LR_OK;
}

// -- batches: many scenes in one wavefront -------------------------------------

__global__ void k_stamp(uint64_t* out) {
    if (threadIdx.x == 0) out[0] = wall_clock64();
}

struct lr_batch {
    lr_context* ctx = nullptr;
    std::vector<lr_scene*> scenes;
    lr_qbvh tree;                 // every scene's nodes / prims in one array, indices rebased
    std::vector<uint32_t> roots;  // root node of each scene
    uint32_t* d_roots = nullptr;
    // per-call tables (grow-only)
    DevScene* d_scenes = nullptr;
    Params* d_params = nullptr;
    DevCamera* d_cams = nullptr;
    GeomArgs* d_geoms = nullptr;
    SceneFilm* d_films = nullptr;
    Segment* d_seg = nullptr;
    size_t seg_cap = 0;
    SplatTile* d_tiles = nullptr;
    size_t tile_cap = 0;
    uint64_t* d_stamps = nullptr;
    size_t stamp_cap = 0;
    // wavefront workspace (grow-only): path state 4, two queues 2 x 10, hits 4
    size_t cap = 0;
    float* d_work = nullptr;
    uint32_t* d_counts = nullptr;
    hipStream_t cstream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace {

void free_batch(lr_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->cstream) {
        (void)hipStreamSynchronize(b->cstream);
        b->ctx->release(b->cstream);
    }
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->cstream) (void)hipStreamDestroy(b->cstream);
    void* ptrs[] = {b->tree.d_nodes, b->tree.d_prims, b->d_roots, b->d_scenes, b->d_params, b->d_cams,
                    (void*)b->d_geoms, b->d_films, b->d_seg, b->d_tiles, b->d_stamps, b->d_work, b->d_counts};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

template <class T>
int grow(T** p, size_t& cap, size_t n) {
    if (cap >= n) return LR_OK;
    if (*p) LR_HIP_CHECK(hipFree((void*)*p));
    *p = nullptr;
    LR_HIP_CHECK(hipMalloc((void**)p, sizeof(T) * n));
    cap = n;
    return LR_OK;
}

struct ChunkPlan {
    uint32_t seg_begin, n_seg, n_paths, n_pix, max_gen;
};

}  // namespace

extern "C" {

int lr_batch_create(lr_context* ctx, lr_scene* const* scenes, uint32_t n, lr_batch** out) {
    if (!ctx || !scenes || !out || n == 0) return lr::fail(LR_E_INVALID, "need a context, scenes and out");
    *out = nullptr;
    uint64_t nodes = 0, prims = 0;
    uint32_t max_stack = 1, max_depth = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const lr_scene* sc = scenes[k];
        if (!sc) return lr::fail(LR_E_INVALID, "scene " + std::to_string(k) + " is NULL");
        if (sc->ctx != ctx) return lr::fail(LR_E_INVALID, "scene " + std::to_string(k) + " belongs to another context");
        if (sc->bvh->host.nodes.empty())
            return lr::fail(LR_E_INVALID, "scene " + std::to_string(k) + " has no geometry (batches need a tree per scene)");
        nodes += sc->bvh->host.nodes.size();
        prims += sc->bvh->host.prims.size();
        max_stack = std::max(max_stack, sc->bvh->info.max_stack);
        max_depth = std::max(max_depth, sc->bvh->info.max_depth);
    }
    if (nodes >= (1ull << 31) || prims >= (1ull << 29))
        return lr::fail(LR_E_INVALID, "batch too large for the node encoding (2^31 nodes, 2^29 primitives)");
    if (int rc = lr::bind(ctx)) return rc;
    lr_batch* b = new (std::nothrow) lr_batch();
    if (!b) return lr::fail(LR_E_NOMEM, "out of host memory");
    b->ctx = ctx;
    b->scenes.assign(scenes, scenes + n);
    // concatenate the trees: inner children += node base, leaf first prim += prim base
    std::vector<lr_node> hn;
    std::vector<lr_prim> hp;
    hn.reserve(nodes);
    hp.reserve(prims);
    for (uint32_t k = 0; k < n; ++k) {
        const lr::BuiltQbvh& H = scenes[k]->bvh->host;
        const uint32_t nb = (uint32_t)hn.size(), pb = (uint32_t)hp.size();
        b->roots.push_back(nb);
        for (const lr_node& nd : H.nodes) {
            lr_node m = nd;
            for (int c = 0; c < 4; ++c) {
                const uint32_t e = nd.child[c];
                if (e == LR_CHILD_EMPTY) continue;
                if (e & 0x80000000u) {
                    const uint32_t first = ((e >> 2) & 0x1FFFFFFFu) + pb;
                    m.child[c] = 0x80000000u | (first << 2) | (e & 3u);
                } else {
                    m.child[c] = e + nb;
                }
            }
            hn.push_back(m);
        }
        hp.insert(hp.end(), H.prims.begin(), H.prims.end());
    }
    lr_qbvh& T = b->tree;
    T.ctx = ctx;
    T.info.n_nodes = (uint32_t)hn.size();
    T.info.n_prims = (uint32_t)hp.size();
    T.info.max_depth = max_depth;
    T.info.max_stack = max_stack;
    const std::vector<lr_node_q8> hq = lr::quantize_nodes(hn);
    hipError_t e = hipMalloc(&T.d_nodes, sizeof(lr_node_q8) * hq.size());
    if (e == hipSuccess) e = hipMalloc(&T.d_prims, sizeof(lr_prim) * hp.size());
    if (e == hipSuccess) e = hipMalloc(&b->d_roots, sizeof(uint32_t) * n);
    if (e == hipSuccess) e = hipMemcpy(T.d_nodes, hq.data(), sizeof(lr_node_q8) * hq.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(T.d_prims, hp.data(), sizeof(lr_prim) * hp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_roots, b->roots.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&b->d_scenes, sizeof(DevScene) * n);
    if (e == hipSuccess) e = hipMalloc(&b->d_params, sizeof(Params) * n);
    if (e == hipSuccess) e = hipMalloc(&b->d_cams, sizeof(DevCamera) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_geoms, sizeof(GeomArgs) * n);
    if (e == hipSuccess) e = hipMalloc(&b->d_films, sizeof(SceneFilm) * n);
    if (e == hipSuccess) e = hipMalloc(&b->d_counts, 64 * sizeof(uint32_t));
    if (e != hipSuccess) {
        free_batch(b);
        return lr::fail(LR_E_DEVICE, std::string("lr_batch_create: ") + hipGetErrorString(e));
    }
    *out = b;
    return LR_OK;
}

int lr_batch_destroy(lr_batch* b) {
    free_batch(b);
    return LR_OK;
}

int lr_render_batch(lr_batch* b, const lr_batch_item* items, uint32_t flags, float* scene_ms, lr_render_stats* stats,
                    void* stream) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!b || !items) return lr::fail(LR_E_INVALID, "NULL batch or items");
    if (int rc = lr::bind(b->ctx)) return rc;
    const uint32_t n = (uint32_t)b->scenes.size();
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<DevScene> hs(n);
    std::vector<Params> hprm(n);
    std::vector<DevCamera> hcam(n);
    std::vector<GeomArgs> hgeo(n);
    std::vector<SceneFilm> hfilm(n);
    std::vector<uint64_t> paths(n);
    for (uint32_t k = 0; k < n; ++k) {
        const lr_batch_item& it = items[k];
        if (!it.film_xyz || !it.film_weight) return lr::fail(LR_E_INVALID, "item " + std::to_string(k) + ": NULL film");
        if (int rc = make_camera(&it.camera, hcam[k])) return rc;
        if (int rc = make_params(it.params, hprm[k])) return rc;
        if (hprm[k].halton) return lr::fail(LR_E_UNSUPPORTED, "batches render with the Stratifier only");
        lr_scene* sc = b->scenes[k];
        if (int rc = build_observer(sc, it.params.lambda_min, it.params.lambda_max)) return rc;
        hs[k] = dev_scene(sc);
        hgeo[k] = geom_args(sc);
        hfilm[k] = SceneFilm{it.film_xyz, it.film_weight};
        paths[k] = (uint64_t)it.params.width * it.params.height * it.params.spp * it.params.wavelengths;
    }
    // shortest scene first: the mean and median completion times are least
    std::vector<uint32_t> order(n);
    for (uint32_t k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return paths[a] < paths[c]; });
    // chunks of whole pixels, spanning scenes: the smallest nonzero
    // chunk_paths of the items, 4M paths when none sets one
    uint64_t cap = 0;
    for (uint32_t k = 0; k < n; ++k)
        if (items[k].params.chunk_paths)
            cap = cap ? std::min<uint64_t>(cap, items[k].params.chunk_paths) : items[k].params.chunk_paths;
    if (!cap) cap = 1u << 22;
    std::vector<Segment> segs;
    std::vector<ChunkPlan> chunks;
    std::vector<uint32_t> last_chunk(n, 0);
    ChunkPlan cur{0, 0, 0, 0, 0};
    for (uint32_t k : order) {
        const Params& P = hprm[k];
        const uint64_t ppp = (uint64_t)P.spp * P.nl;
        if (ppp > cap) return lr::fail(LR_E_INVALID, "item " + std::to_string(k) + ": spp * wavelengths exceeds a chunk");
        const uint64_t npix = (uint64_t)P.width * P.height;
        uint64_t done = 0;
        while (done < npix) {
            uint64_t fit = (cap - cur.n_paths) / ppp;
            if (fit == 0) {
                chunks.push_back(cur);
                cur = ChunkPlan{(uint32_t)segs.size(), 0, 0, 0, 0};
                fit = cap / ppp;
            }
            const uint64_t take = std::min(fit, npix - done);
            segs.push_back(Segment{k, cur.n_paths, cur.n_pix, (uint32_t)done});
            cur.n_seg += 1;
            cur.n_paths += (uint32_t)(take * ppp);
            cur.n_pix += (uint32_t)take;
            cur.max_gen = std::max(cur.max_gen, P.max_gen);
            done += take;
            last_chunk[k] = (uint32_t)chunks.size();
        }
    }
    if (cur.n_seg) chunks.push_back(cur);
    // splat tiles: per chunk, each segment's pixels in groups of max(1, 256 / ppp)
    std::vector<SplatTile> tiles;
    std::vector<uint32_t> tile0(chunks.size() + 1, 0);
    for (size_t c = 0; c < chunks.size(); ++c) {
        const ChunkPlan& ch = chunks[c];
        tile0[c] = (uint32_t)tiles.size();
        for (uint32_t j = 0; j < ch.n_seg; ++j) {
            const Segment& sg = segs[ch.seg_begin + j];
            const uint32_t end = j + 1 < ch.n_seg ? segs[ch.seg_begin + j + 1].pix_begin : ch.n_pix;
            const uint32_t ppp = hprm[sg.scene].spp * hprm[sg.scene].nl;
            const uint32_t pb = ppp >= 256u ? 1u : 256u / ppp;
            for (uint32_t p = sg.pix_begin; p < end; p += pb) tiles.push_back(SplatTile{j, p, std::min(pb, end - p)});
        }
    }
    tile0[chunks.size()] = (uint32_t)tiles.size();
    // the previous call may still read the tables: wait for it before rewriting them
    if (b->cstream) LR_HIP_CHECK(hipStreamSynchronize(b->cstream));
    LR_HIP_CHECK(hipStreamSynchronize(s));
    if (int rc = grow(&b->d_seg, b->seg_cap, segs.size())) return rc;
    if (int rc = grow(&b->d_tiles, b->tile_cap, tiles.size())) return rc;
    LR_HIP_CHECK(hipMemcpy(b->d_tiles, tiles.data(), sizeof(SplatTile) * tiles.size(), hipMemcpyHostToDevice));
    if (int rc = grow(&b->d_stamps, b->stamp_cap, chunks.size() + 1)) return rc;
    LR_HIP_CHECK(hipMemcpy(b->d_seg, segs.data(), sizeof(Segment) * segs.size(), hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(b->d_scenes, hs.data(), sizeof(DevScene) * n, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(b->d_params, hprm.data(), sizeof(Params) * n, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(b->d_cams, hcam.data(), sizeof(DevCamera) * n, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy((void*)b->d_geoms, hgeo.data(), sizeof(GeomArgs) * n, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(b->d_films, hfilm.data(), sizeof(SceneFilm) * n, hipMemcpyHostToDevice));
    if (b->cap < cap) {
        if (b->d_work) LR_HIP_CHECK(hipFree(b->d_work));
        b->d_work = nullptr;
        b->cap = 0;  // a failed allocation below leaves no stale capacity behind
        LR_HIP_CHECK(hipMalloc(&b->d_work, sizeof(float) * 28 * cap));
        b->cap = cap;
    }
    float* w = b->d_work;
    PathState ps{w, w + cap, w + 2 * cap, w + 3 * cap};
    Queue q[2];
    for (int k = 0; k < 2; ++k) {
        float* base = w + (4 + 10 * k) * cap;
        for (int r = 0; r < 8; ++r) q[k].r[r] = base + r * cap;
        q[k].pid = reinterpret_cast<uint32_t*>(base + 8 * cap);
        q[k].root = reinterpret_cast<uint32_t*>(base + 9 * cap);
    }
    float* hb = w + 24 * cap;
    Hits H{reinterpret_cast<uint32_t*>(hb), hb + cap, hb + 2 * cap, hb + 3 * cap};
    uint32_t* counts = b->d_counts;
    uint64_t* rays_total = reinterpret_cast<uint64_t*>(b->d_counts + 4);
    BatchArgs B;
    B.scenes = b->d_scenes;
    B.params = b->d_params;
    B.cams = b->d_cams;
    B.geoms = b->d_geoms;
    B.roots = b->d_roots;
    B.films = b->d_films;
    GuideArgs NG;
    std::memset(&NG, 0, sizeof(NG));
    const DevScene S0 = hs[0];
    const Params P0 = hprm[0];
    const DevCamera C0 = hcam[0];

    const bool graph = (flags & LR_BATCH_GRAPH) != 0;
    const hipStream_t caller = s;
    if (graph) {
        if (!b->cstream) {
            LR_HIP_CHECK(hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking));
            LR_HIP_CHECK(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
            LR_HIP_CHECK(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
        }
        s = b->cstream;
    }
    if (int rc = lr::trace_reserve(&b->tree, (uint32_t)cap, s)) return rc;
    if (graph) {
        LR_HIP_CHECK(hipEventRecord(b->ev_in, caller));
        LR_HIP_CHECK(hipStreamWaitEvent(s, b->ev_in, 0));
        LR_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    }
    auto body = [&]() -> int {
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counts, 16u);
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, b->d_stamps);
        for (size_t c = 0; c < chunks.size(); ++c) {
            const ChunkPlan& ch = chunks[c];
            B.seg = b->d_seg + ch.seg_begin;
            B.n_seg = ch.n_seg;
            const dim3 blk(256);
            const dim3 ablk(kAppendBlock), agrd((ch.n_paths + kAppendBlock - 1) / kAppendBlock);
            hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counts, 1u);
            hipLaunchKernelGGL(k_camera<true>, agrd, ablk, 0, s, S0, C0, P0, (uint64_t)0, ch.n_paths, ps, q[0], counts,
                               B);
            for (uint32_t gen = 0; gen < ch.max_gen; ++gen) {
                const int cu = gen & 1, nx = cu ^ 1;
                hipLaunchKernelGGL(k_next_gen, dim3(1), dim3(64), 0, s, counts, (uint32_t)cu, (uint32_t)nx, rays_total);
                lr_rays_soa rays{q[cu].r[0], q[cu].r[1], q[cu].r[2], q[cu].r[3],
                                 q[cu].r[4], q[cu].r[5], q[cu].r[6], q[cu].r[7]};
                lr_hits_soa hits{H.prim, H.t, H.u, H.v};
                if (int rc = lr::trace_closest_queue(&b->tree, &rays, ch.n_paths, counts + cu, &hits, s, q[cu].root))
                    return rc;
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<false, true>),
                                   dim3((ch.n_paths + kAppendBlockWide - 1) / kAppendBlockWide), dim3(kAppendBlockWide), 0,
                                   s, S0, P0, gen, (uint64_t)0,
                                   q[cu], counts + cu, H, hgeo[0], ps, q[nx], counts + nx, NG, B);
            }
            hipLaunchKernelGGL(k_splat_tiles_batch, dim3(tile0[c + 1] - tile0[c]), blk, 0, s, B, b->d_tiles + tile0[c],
                               ps);
            hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, b->d_stamps + c + 1);
            LR_HIP_CHECK(hipGetLastError());
        }
        return LR_OK;
    };
    int rc = body();
    if (graph) {
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (!rc && ec != hipSuccess) rc = lr::fail(LR_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
        if (!rc) rc = launch_graph(b->gexec, g, s);
        if (g) (void)hipGraphDestroy(g);
        if (!rc) {
            LR_HIP_CHECK(hipEventRecord(b->ev_out, s));
            LR_HIP_CHECK(hipStreamWaitEvent(caller, b->ev_out, 0));
        }
        s = caller;
    }
    if (rc) return rc;
    if (scene_ms || stats) {
        std::vector<uint64_t> st(chunks.size() + 1);
        uint64_t rays = 0;
        LR_HIP_CHECK(hipMemcpyAsync(st.data(), b->d_stamps, sizeof(uint64_t) * st.size(), hipMemcpyDeviceToHost, s));
        LR_HIP_CHECK(hipMemcpyAsync(&rays, rays_total, sizeof(rays), hipMemcpyDeviceToHost, s));
        LR_HIP_CHECK(hipStreamSynchronize(s));
        int khz = 0;
        LR_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->ctx->device));
        if (scene_ms)
            for (uint32_t k = 0; k < n; ++k)
                scene_ms[k] = khz > 0 ? (float)((double)(st[last_chunk[k] + 1] - st[0]) / (double)khz) : 0.0f;
        if (stats) {
            uint64_t total = 0;
            for (uint64_t p : paths) total += p;
            stats->paths = total;
            stats->rays = rays;
            stats->chunks = (uint32_t)chunks.size();
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    return LR_OK;
}

int lr_render(lr_scene* sc, const lr_camera* cam, const lr_render_params* prm, float* film_xyz, float* film_w,
              lr_render_stats* stats, void* stream) {
    return render_impl(sc, cam, prm, nullptr, nullptr, film_xyz, film_w, stats, stream);
}

int lr_render_guided(lr_scene* sc, const lr_camera* cam, const lr_render_params* prm, lr_guide* guide,
                     const lr_guide_params* gparams, float* film_xyz, float* film_w, lr_render_stats* stats,
                     void* stream) {
    if (!guide) return lr::fail(LR_E_INVALID, "NULL guide");
    return render_impl(sc, cam, prm, guide, gparams, film_xyz, film_w, stats, stream);
}

}  // extern "C"


// ============================================================================
// DirectLighting (src/liar/tracers/direct_lighting.cpp:119-377) as a
// wavefront.  Every surface hit adds its emission, a light estimate
// (RayTracer::estimateLightContribution, src/liar/kernel/ray_tracer.cpp:188-268)
// whose samples become shadow rays, and continues along its specular (and
// glossy) reflection and transmission directions.  A dielectric continues
// both ways, so a camera sample's rays form a tree: every queue entry carries
// the weight of its branch, and radiance reaches the camera sample's
// accumulator by atomics (so films are reproducible up to the rounding of
// the order of those additions).
// ============================================================================
namespace {

struct DLQueue {
    float* r[8];     // ox oy oz dx dy dz tnear tfar
    uint32_t* pid;   // camera path (chunk-local)
    float* w;        // branch weight: product of value |cos| / pdf of its continuations
    uint32_t* code;  // bit 31: high quality; bits 0..22: one bit per continuation (0 reflection, 1 transmission)
};

struct ShadowQueue {
    float* r[8];
    uint32_t* pid;
    float* c;  // radiance added to the camera path when the ray is unoccluded
};

struct DLArgs {
    uint32_t max_gen;    // generations 0 (camera) .. max_gen are cast
    uint32_t secondary;  // light samples of a hit that is not high quality
    uint32_t glossy;     // continue along glossy reflections
    uint32_t slots;      // shadow-ray candidates per hit (uniform loop bound)
    uint32_t qcap, scap;
    uint32_t* qover;  // branch entries that did not fit (the host grows the queues and repeats the chunk)
    uint32_t* sover;  // shadow rays that did not fit
    unsigned long long* lq;  // per camera path: radiance in fixed point (kRadianceFx)
};

constexpr uint32_t kHQ = 0x80000000u;

// A camera path's DirectLighting radiance arrives from many threads (its ray
// tree's emission and shadow rays) in scheduling order.  It is summed in
// 64-bit fixed point, 2^-32 per unit, with integer atomics, which are exact
// and order-free, so films are bit-reproducible; a term is quantised to
// 2.3e-10 absolute and a path's sum holds up to 2^31.  Terms are >= 0; a
// non-finite or larger one saturates at 2^31.
constexpr double kRadianceFx = 4294967296.0;
__device__ inline unsigned long long radiance_fx(float c) {
    const double v = (double)c * kRadianceFx;
    return v >= 0.0 && v < 9.2e18 ? (unsigned long long)(v + 0.5) : 0x7FFFFFFFFFFFFFFFull;
}

// Chunk start: every camera path's fixed-point radiance at zero.
__global__ void k_dl_zero(uint32_t n, unsigned long long* lq) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lq[i] = 0ull;
}

// Chunk end: the fixed-point sums into the path state the splat reads.
__global__ void k_dl_fold(uint32_t n, const unsigned long long* lq, PathState ps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ps.L[i] = (float)((double)lq[i] * (1.0 / kRadianceFx));
}

// Power heuristic a^2 / (a^2 + b^2) for a > 0, written so that large pdfs
// do not overflow in FP32 (ray_tracer.cpp:180-183).
__device__ inline float power_heuristic(float a, float b) {
    const float r = b / a;
    return 1.0f / (1.0f + r * r);
}

__device__ inline bool bsdf_is_specular(uint32_t t) {
    return t == LR_MAT_MIRROR || t == LR_MAT_CONDUCTOR || t == LR_MAT_DIELECTRIC;
}

// Random numbers of a tree node: (generation, branch) replaces the generation
// of the path tracer's key, so every node of every tree has its own streams.
__device__ inline lr::u4 dl_rng(const Params& P, uint64_t gp, uint32_t gen, uint32_t br, uint32_t purpose) {
    return rng(P, gp, gen | (br << 8), purpose);
}

// Fifth uniform of a 128-bit block whose four words also give four: the low
// bytes of x, y, z as a 24-bit fraction (u01 reads the top 24 bits).
__device__ inline uint32_t dl_pair_fifth(const lr::u4& r) {
    return ((r.x & 0xffu) | ((r.y & 0xffu) << 8) | ((r.z & 0xffu) << 16)) << 8;
}

struct DLHit {
    BsdfParams B;
    V3 wi, tt, bb, nf, ngl, start;
    float w, lam;
    uint64_t gp;
    uint32_t gen, br;
    bool hq;
};

// Shadow-ray candidate j of a hit.  High quality: light k's 2 n_k
// candidates follow those of lights 0..k-1, even ones light samples, odd
// ones BSDF samples (n_k = lr_light.samples).  Otherwise candidates 2s and
// 2s+1 are the light and BSDF sample of secondary sample s, whose light is
// picked by power and whose estimate is divided by (samples x pick pdf)
// (direct_lighting.cpp:259-310).  Returns false when the candidate makes no
// ray; c is its radiance if unoccluded, [kTolerance, far) its extent.
// le_k / le_v cache the radiance of the last light a candidate of this hit
// used (candidates of one light are consecutive slots): a sampled spectrum
// costs a binary search through global memory.
// The two candidates of a pair (2s, 2s+1: same light, same pick) share their
// random numbers: the even one draws the pick (purpose 10 + s, secondary
// samples only) and one 128-bit block (purpose 100 + s) into rs / rp, the
// odd one reads them.  Whatever ends the even candidate before a draw ends
// the odd one too, so the odd one never reads a stale block.
__device__ inline bool dl_candidate(const DevScene& S, const Params& P, const DLArgs& A, const DLHit& H, uint32_t j,
                                    float& c, V3& dir, float& far, uint32_t& le_k, float& le_v, lr::u4& rs,
                                    lr::u4& rp) {
    uint32_t k = 0, local = 0;
    float nl = 1.0f, scale = 1.0f;
    if (H.hq) {
        uint32_t base = 0;
        bool found = false;
        for (uint32_t m = 0; m < S.n_lights && !found; ++m) {
            const uint32_t n2 = 2u * S.lights[m].samples;
            if (j < base + n2) {
                k = m;
                local = j - base;
                found = true;
            }
            base += n2;
        }
        if (!found) return false;
        nl = (float)S.lights[k].samples;
    } else {
        if (j >= 2u * A.secondary || S.n_lights == 0) return false;
        if ((j & 1u) == 0u) rs = dl_rng(P, H.gp, H.gen, H.br, 10u + (j >> 1));
        k = light_pick(S, lr::u01(rs.x));
        const float psel = S.lights[k].sel_pdf;
        if (!(psel > 0.0f)) return false;
        local = j & 1u;
        scale = (float)A.secondary * psel;
    }
    const DevLight& Lk = S.lights[k];
    if (k != le_k) {
        le_v = spectrum_eval(S, Lk.radiance, H.lam, false);
        le_k = k;
    }
    const float Le = le_v;
    if (Le == 0.0f) return false;
    if ((j & 1u) == 0u) rp = dl_rng(P, H.gp, H.gen, H.br, 100u + (j >> 1));
    float f = 0.0f, pb = 0.0f, pl = 0.0f;
    float dist = INFINITY;  // to the mesh light's surface along dir
    if ((local & 1u) == 0u) {
        // light sample, weighted against BSDF sampling (ray_tracer.cpp:211-231)
        dir = light_sample(Lk, S.light_tris, H.start, lr::u01(rp.x), lr::u01(rp.y), pl, dist);
        if (!(pl > 0.0f)) return false;
        // sampleEmission with the target's normal rejects lights behind the
        // shading normal (light_area.cpp:242-252)
        if (!(dot(dir, H.nf) > 0.0f)) return false;
        const V3 wo = v3(dot(dir, H.tt), dot(dir, H.bb), dot(dir, H.nf));
        if (!bsdf_evaluate(H.B, H.wi, wo, H.ngl, f, pb) || f == 0.0f) return false;
        c = H.w * f * Le * power_heuristic(nl * pl, nl * pb) * fabsf(wo.z) / (nl * pl * scale);
    } else {
        // BSDF sample that reaches the light (ray_tracer.cpp:234-264)
        V3 wo;
        if (!bsdf_sample(H.B, H.wi, lr::u01(rp.z), lr::u01(rp.w), lr::u01(dl_pair_fifth(rp)), wo, f, pb) ||
            !(pb > 0.0f) || f == 0.0f)
            return false;
        dir = normalize(wo.x * H.tt + wo.y * H.bb + wo.z * H.nf);
        pl = light_pdf(Lk, S.light_tris, H.start, dir);
        if (!(pl > 0.0f)) return false;
        if (Lk.type == LR_LIGHT_MESH) {
            V3 nl_;
            dist = mesh_light_hit(Lk, S.light_tris, H.start, dir, nl_);
        }
        const float wt = bsdf_is_specular(H.B.type) ? 1.0f : power_heuristic(nl * pb, nl * pl);
        c = H.w * f * Le * wt * fabsf(wo.z) / (nl * pb * scale);
    }
    far = INFINITY;
    if (Lk.type == LR_LIGHT_PARALLELOGRAM) {
        // ends 1e-4 (relative) short of the light's plane, so that the
        // light's own triangles never occlude it in FP32
        const V3 n = ld3(Lk.n);
        const float t = dot(ld3(Lk.o) - H.start, n) / dot(dir, n);
        if (!(t > 0.0f)) return false;
        far = t * (1.0f - 1e-4f);
    } else if (Lk.type == LR_LIGHT_MESH) {
        // likewise short of the sampled point / the surface the ray reaches
        if (!(dist > 0.0f) || !isfinite(dist)) return false;
        far = dist * (1.0f - 1e-4f);
    }
    return c != 0.0f && isfinite(c);
}

// Camera entries: the path's throughput, high quality, branch 0.
__global__ void k_dl_init(const uint32_t* count, PathState ps, DLQueue d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    d.w[i] = ps.thr[d.pid[i]];
    d.code[i] = kHQ;
}

// Start of a generation: queue `cur` is about to be traced, the next queue
// and the shadow queue start empty.
__global__ void k_dl_next(uint32_t* counts, uint32_t cur, uint32_t nxt, uint32_t qcap, uint64_t* rays) {
    if (threadIdx.x == 0) {
        *rays += min(counts[cur], qcap);
        counts[nxt] = 0u;
        counts[2] = 0u;
    }
}

__global__ void k_dl_fill_miss(const uint32_t* count, uint32_t cap, Hits h) {
    const uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        h.prim[i] = LR_MISS;
        h.t[i] = INFINITY;
    }
}

// One generation's hits: emission, shadow-ray candidates, continuations.
// Grid-stride over the queue in whole workgroups (every thread of a block
// runs the same number of appends).
// Shadow rays of a block are staged in LDS and flushed to the shadow queue
// with one atomic and coalesced copies per flush (every 512-768 rays),
// instead of one atomic per block and candidate slot: the queue's single
// counter serialises its atomics (~10 ns each), which set the kernel's time.
constexpr uint32_t kStage = 768;
struct ShadowStage {
    float r[8][kStage];
    uint32_t pid[kStage];
    float c[kStage];
};

// Copies the block's `fill` staged rays to the shadow queue (all threads).
__device__ void stage_flush(ShadowStage& st, uint32_t fill, ShadowQueue sq, uint32_t* scount, uint32_t scap,
                            uint32_t* sover) {
    __shared__ uint32_t s_gbase;
    __syncthreads();  // staged writes done
    if (threadIdx.x == 0) {
        s_gbase = atomicAdd(scount, fill);
        if (s_gbase + (uint64_t)fill > scap) atomicAdd(sover, (uint32_t)(s_gbase + (uint64_t)fill - max(s_gbase, scap)));
    }
    __syncthreads();
    const uint32_t gb = s_gbase;
    for (uint32_t k = threadIdx.x; k < fill; k += blockDim.x) {
        const uint32_t g = gb + k;
        if (g < scap) {
#pragma unroll
            for (int w = 0; w < 8; ++w) sq.r[w][g] = st.r[w][k];
            sq.pid[g] = st.pid[k];
            sq.c[g] = st.c[k];
        }
    }
    __syncthreads();  // the stage is reused
}

__global__ __launch_bounds__(256) void k_dl_shade(DevScene S, Params P, uint32_t gen, Hits h,
                                                  GeomArgs Gm, DLQueue q,
                                                  const uint32_t* count, PathState ps, DLQueue qn,
                                                  uint32_t* count_next, ShadowQueue sq, uint32_t* scount,
                                                  uint64_t path0, DLArgs A) {
    const uint32_t n = min(*count, A.qcap);
    __shared__ ShadowStage stage;
    __shared__ uint32_t s_wcnt[2][16];  // per-wave counts, double-buffered by slot parity
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t fill = 0, parity = 0;  // block-uniform
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        DLHit H;
        bool surf = false;
        uint32_t pid = 0;
        V3 x = v3(0, 0, 0);
        if (i < n && h.prim[i] != LR_MISS) {  // escaped rays add nothing (alpha = 0)
            pid = q.pid[i];
            const uint32_t prim = h.prim[i];
            const V3 o = v3(q.r[0][i], q.r[1][i], q.r[2][i]);
            const V3 d = v3(q.r[3][i], q.r[4][i], q.r[5][i]);
            const float tnear = q.r[6][i];
            H.w = q.w[i];
            const uint32_t code = q.code[i];
            H.hq = (code & kHQ) != 0u;
            H.br = code & ~kHQ;
            H.gen = gen;
            H.gp = path0 + pid;
            H.lam = ps.lam[pid];
            const HitFrame F = hit_frame(Gm, S.n_triangles, prim, o, d, tnear, h.t[i], h.u[i], h.v[i]);
            x = F.x;
            const bool leaving = F.leaving;
            H.nf = F.ns;
            H.tt = F.tt;
            H.bb = F.bb;
            H.ngl = local_geometric_normal(F);
            const DevMaterial m = S.mats[S.prim_mat[prim]];
            // shader emission: camera rays and objects that are not lights
            // (direct_lighting.cpp:248-255)
            if ((gen == 0 || m.light == LR_NONE) && m.emission != LR_NONE) {
                const float Le = spectrum_eval(S, m.emission, H.lam, false);
                if (Le != 0.0f) atomicAdd(&A.lq[pid], radiance_fx(H.w * Le));
            }
            if (m.type != LR_MAT_NONE) {
                surf = true;
                H.B.type = m.type;
                H.B.refl = spectrum_eval(S, m.refl, H.lam, true);
                H.B.trans = spectrum_eval(S, m.trans, H.lam, true);
                H.B.eta = fmaxf(spectrum_eval(S, m.eta, H.lam, false), 1e-9f);
                H.B.kappa = fmaxf(spectrum_eval(S, m.kappa, H.lam, false), 1e-9f);
                H.B.ior = leaving ? H.B.eta / H.B.kappa : H.B.kappa / H.B.eta;
                const float ru = fmaxf(m.roughness, 1e-3f), rv = fmaxf(m.roughness_v, 1e-3f);
                H.B.alpha_u = ru * ru;
                H.B.alpha_v = rv * rv;
                H.B.fuzz = m.fuzz;
                const V3 wiw = -d;
                H.wi = v3(dot(wiw, H.tt), dot(wiw, H.bb), fmaxf(dot(wiw, H.nf), 1.1920929e-7f));
                H.start = x + (10.0f * kTolerance) * H.nf;
            }
        }
        uint32_t le_k = ~0u;
        float le_v = 0.0f;
        lr::u4 rs{0u, 0u, 0u, 0u}, rp{0u, 0u, 0u, 0u};
        for (uint32_t j = 0; j < A.slots; ++j) {
            float c = 0.0f, far = 0.0f;
            V3 dir = v3(0, 0, 1);
            const bool alive = surf && dl_candidate(S, P, A, H, j, c, dir, far, le_k, le_v, rs, rp);
            // block-local compaction into the stage: one barrier per slot
            const uint64_t mask = __ballot(alive);
            if (lane == 0) s_wcnt[parity][wave] = (uint32_t)__popcll(mask);
            __syncthreads();
            uint32_t off = fill, tot = 0;
            for (uint32_t w = 0; w < nw; ++w) {
                const uint32_t cw = s_wcnt[parity][w];
                if (w < wave) off += cw;
                tot += cw;
            }
            parity ^= 1u;
            if (alive) {
                const uint32_t k = off + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                stage.r[0][k] = H.start.x;
                stage.r[1][k] = H.start.y;
                stage.r[2][k] = H.start.z;
                stage.r[3][k] = dir.x;
                stage.r[4][k] = dir.y;
                stage.r[5][k] = dir.z;
                stage.r[6][k] = kTolerance;
                stage.r[7][k] = far;
                stage.pid[k] = pid;
                stage.c[k] = c;
            }
            fill += tot;
            if (fill + blockDim.x > kStage) {
                stage_flush(stage, fill, sq, scount, A.scap, A.sover);
                fill = 0;
            }
        }
        // specular / glossy continuations (direct_lighting.cpp:314-377):
        // reflection with the reflection caps, transmission (dielectric only)
        // with the transmission caps; high quality only along specular ones
        const bool glossy = surf && (H.B.type == LR_MAT_COOK_TORRANCE || H.B.type == LR_MAT_COOK_TORRANCE_BECKMANN);
        const bool cont = surf && gen < A.max_gen && (bsdf_is_specular(H.B.type) || (glossy && A.glossy));
        for (uint32_t side = 0; side < 2; ++side) {
            bool alive = false;
            V3 dir = v3(0, 0, 1), org = v3(0, 0, 0);
            float wc = 0.0f;
            if (cont && (side == 0 || H.B.type == LR_MAT_DIELECTRIC)) {
                const lr::u4 r = dl_rng(P, H.gp, gen, H.br, 5u + side);
                V3 wo;
                float f = 0.0f, pdf = 0.0f;
                if (bsdf_sample(H.B, H.wi, lr::u01(r.x), lr::u01(r.y), lr::u01(r.z), wo, f, pdf, side ? 2u : 1u) &&
                    pdf > 0.0f && f != 0.0f && (side ? wo.z < 0.0f : wo.z > 0.0f)) {
                    wc = H.w * f * fabsf(wo.z) / pdf;
                    if (wc > 0.0f && isfinite(wc)) {
                        alive = true;
                        dir = normalize(wo.x * H.tt + wo.y * H.bb + wo.z * H.nf);
                        org = x + ((side ? -10.0f : 10.0f) * kTolerance) * H.nf;
                    }
                }
            }
            const uint32_t pos = block_append(alive, count_next);
            if (alive) {
                if (pos < A.qcap) {
                    qn.r[0][pos] = org.x;
                    qn.r[1][pos] = org.y;
                    qn.r[2][pos] = org.z;
                    qn.r[3][pos] = dir.x;
                    qn.r[4][pos] = dir.y;
                    qn.r[5][pos] = dir.z;
                    qn.r[6][pos] = kTolerance;
                    qn.r[7][pos] = INFINITY;
                    qn.pid[pos] = pid;
                    qn.w[pos] = wc;
                    qn.code[pos] = ((H.hq && bsdf_is_specular(H.B.type)) ? kHQ : 0u) | ((H.br << 1) | side);
                } else {
                    atomicAdd(A.qover, 1u);
                }
            }
        }
    }
    if (fill) stage_flush(stage, fill, sq, scount, A.scap, A.sover);
}

// Unoccluded shadow rays add their radiance to their camera path.
__global__ void k_dl_resolve(const uint32_t* scount, uint32_t scap, ShadowQueue sq, const uint8_t* occluded,
                             unsigned long long* lq, uint64_t* rays) {
    const uint32_t n = min(*scount, scap);
    if (blockIdx.x == 0 && threadIdx.x == 0) *rays += n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (!occluded || !occluded[i]) atomicAdd(&lq[sq.pid[i]], radiance_fx(sq.c[i]));
}

// Grows the DirectLighting workspace of a scene: two branch queues (11 words
// per entry), their hits (4 words) and the shadow queue (10 words per ray +
// an occlusion byte).
int ensure_dl_work(lr_scene* sc, size_t qcap, size_t scap, hipStream_t s) {
    if (sc->d_dlwork && sc->d_dlocc && sc->dl_qcap >= qcap && sc->dl_scap >= scap) return LR_OK;
    LR_HIP_CHECK(hipStreamSynchronize(s));
    qcap = std::max(qcap, sc->dl_qcap);
    scap = std::max(scap, sc->dl_scap);
    if (sc->d_dlwork) LR_HIP_CHECK(hipFree(sc->d_dlwork));
    if (sc->d_dlocc) LR_HIP_CHECK(hipFree(sc->d_dlocc));
    sc->d_dlwork = nullptr;
    sc->d_dlocc = nullptr;
    sc->dl_qcap = sc->dl_scap = 0;
    LR_HIP_CHECK(hipMalloc(&sc->d_dlwork, sizeof(float) * (26 * qcap + 10 * scap)));
    LR_HIP_CHECK(hipMalloc(&sc->d_dlocc, scap));
    sc->dl_qcap = qcap;
    sc->dl_scap = scap;
    return LR_OK;
}

}  // namespace

extern "C" {

int lr_render_direct(lr_scene* sc, const lr_camera* cam, const lr_render_params* prm, const lr_direct_params* dprm,
                     float* film_xyz, float* film_w, lr_render_stats* stats, void* stream) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!sc || !cam || !prm) return lr::fail(LR_E_INVALID, "NULL scene, camera or params");
    if (!film_xyz || !film_w) return lr::fail(LR_E_INVALID, "NULL film buffer");
    if (int rc = lr::bind(sc->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const lr_render_params& R = *prm;
    DevCamera C;
    if (int rc = make_camera(cam, C)) return rc;
    Params P;
    if (int rc = make_params(R, P)) return rc;
    if (P.halton) return lr::fail(LR_E_UNSUPPORTED, "DirectLighting renders with the Stratifier only");
    P.max_gen = 1;              // the camera kernel makes generation 0 only
    P.rr_depth = 0xFFFFFFFFu;   // no roulette
    DLArgs A;
    std::memset(&A, 0, sizeof(A));
    A.max_gen = dprm ? dprm->max_generation : 8u;
    A.secondary = dprm && dprm->secondary_light_samples ? dprm->secondary_light_samples : 1u;
    A.glossy = dprm ? (dprm->trace_glossy ? 1u : 0u) : 1u;
    if (A.max_gen > 23) return lr::fail(LR_E_INVALID, "lr_render_direct: max_generation must be <= 23");
    if (A.secondary > 8) return lr::fail(LR_E_INVALID, "lr_render_direct: secondary_light_samples must be <= 8");
    if (int rc = build_observer(sc, R.lambda_min, R.lambda_max)) return rc;
    const DevScene S = dev_scene(sc);
    A.slots = sc->n_lights ? std::max(2 * sc->light_samples, 2 * A.secondary) : 0u;
    if (A.slots > 256) return lr::fail(LR_E_INVALID, "lr_render_direct: more than 128 light samples per hit");

    const uint64_t ppp = (uint64_t)R.spp * R.wavelengths;
    const uint64_t n_pix = (uint64_t)R.width * R.height;
    const uint64_t chunk_req = R.chunk_paths ? R.chunk_paths : (1u << 22);
    // The shadow queue holds slots rays per camera path and its 32-bit
    // counters must not wrap: a chunk's slots * paths stays within 2^30.
    const uint64_t kQueueMax = 1ull << 30;
    const uint64_t slots_eff = std::max(A.slots, 1u);
    if (slots_eff * ppp > kQueueMax)
        return lr::fail(LR_E_INVALID, "lr_render_direct: light samples * spp * wavelengths exceed 2^30 rays per pixel");
    const uint64_t pix_per_chunk = std::min<uint64_t>(
        std::min<uint64_t>(std::max<uint64_t>(1, chunk_req / ppp), kQueueMax / (slots_eff * ppp)),
        std::max<uint64_t>(1, n_pix));  // never more than the render's own pixels
    const uint64_t cap64 = pix_per_chunk * ppp;
    if (cap64 >= (1ull << 26)) return lr::fail(LR_E_INVALID, "spp * wavelengths too large for one chunk");
    const uint32_t cap = (uint32_t)cap64;
    if (int rc = ensure_work(sc, cap, s)) return rc;  // camera path state
    size_t qcap = std::max<size_t>(sc->dl_qcap, 2 * (size_t)cap);
    size_t scap = std::max<size_t>(sc->dl_scap, (size_t)std::max(A.slots, 1u) * cap);
    uint32_t* counts = sc->d_counts;  // [0..1] branch queues, [2] shadow rays, [3] qover, [4..5] rays, [6] sover
    uint64_t* rays_dev = reinterpret_cast<uint64_t*>(counts + 4);
    float* pw = sc->d_work;
    const PathState ps{pw, pw + cap, pw + 2 * (size_t)cap, pw + 3 * (size_t)cap};
    // the camera paths' fixed-point radiance: words 4..5 of the path workspace
    // (16-byte aligned), which the path tracer's queues use and this renderer does not
    A.lq = reinterpret_cast<unsigned long long*>(pw + 4 * (size_t)cap);
    BatchArgs NB;
    std::memset(&NB, 0, sizeof(NB));
    const dim3 blk(256);
    const uint32_t grid_max = (uint32_t)std::max(1, sc->ctx->n_cu) * 16u;
    uint64_t rays = 0;
    uint32_t chunks = 0;
    for (uint64_t pix0 = 0; pix0 < n_pix; pix0 += pix_per_chunk) {
        const uint32_t npix = (uint32_t)std::min<uint64_t>(pix_per_chunk, n_pix - pix0);
        const uint32_t n = (uint32_t)(npix * ppp);
        const uint64_t path0 = pix0 * ppp;
        for (;;) {  // repeated with larger queues when a branch or shadow queue overflowed
            if (int rc = ensure_dl_work(sc, qcap, scap, s)) return rc;
            qcap = sc->dl_qcap;
            scap = sc->dl_scap;
            A.qcap = (uint32_t)qcap;
            A.scap = (uint32_t)scap;
            A.qover = counts + 3;
            A.sover = counts + 6;
            DLQueue q[2];
            float* b = sc->d_dlwork;
            for (int k = 0; k < 2; ++k) {
                float* base = b + (size_t)k * 11 * qcap;
                for (int r = 0; r < 8; ++r) q[k].r[r] = base + (size_t)r * qcap;
                q[k].pid = reinterpret_cast<uint32_t*>(base + 8 * qcap);
                q[k].w = base + 9 * qcap;
                q[k].code = reinterpret_cast<uint32_t*>(base + 10 * qcap);
            }
            float* hb = b + 22 * qcap;
            const Hits H{reinterpret_cast<uint32_t*>(hb), hb + qcap, hb + 2 * qcap, hb + 3 * qcap};
            ShadowQueue sq;
            {
                float* base = hb + 4 * qcap;
                for (int r = 0; r < 8; ++r) sq.r[r] = base + (size_t)r * scap;
                sq.pid = reinterpret_cast<uint32_t*>(base + 8 * scap);
                sq.c = base + 9 * scap;
            }
            Queue cq;  // the camera kernel writes the first branch queue's rays
            for (int r = 0; r < 8; ++r) cq.r[r] = q[0].r[r];
            cq.pid = q[0].pid;
            cq.root = nullptr;
            hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counts, 8u);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_camera<false>), dim3((n + kAppendBlock - 1) / kAppendBlock),
                               dim3(kAppendBlock), 0, s, S, C, P, path0, n,
                               ps, cq, counts, NB);
            hipLaunchKernelGGL(k_dl_init, dim3((n + 255) / 256), blk, 0, s, counts, ps, q[0]);
            hipLaunchKernelGGL(k_dl_zero, dim3((n + 255) / 256), blk, 0, s, n, A.lq);
            // no specular (or traced glossy) material: no hit continues, so
            // generations after the camera's are empty and not launched
            const uint32_t last_gen = (sc->has_specular || (A.glossy && sc->has_glossy)) ? A.max_gen : 0u;
            for (uint32_t gen = 0; gen <= last_gen; ++gen) {
                const int cur = gen & 1, nxt = cur ^ 1;
                const uint32_t glimit = gen == 0 ? n : A.qcap;  // entries this generation can hold
                const uint32_t grid = std::min<uint32_t>((glimit + 255) / 256, grid_max);
                hipLaunchKernelGGL(k_dl_next, dim3(1), dim3(64), 0, s, counts, (uint32_t)cur, (uint32_t)nxt, A.qcap,
                                   rays_dev);
                lr_rays_soa rs{q[cur].r[0], q[cur].r[1], q[cur].r[2], q[cur].r[3],
                               q[cur].r[4], q[cur].r[5], q[cur].r[6], q[cur].r[7]};
                lr_hits_soa hs{H.prim, H.t, H.u, H.v};
                if (sc->bvh->info.n_nodes == 0) {
                    hipLaunchKernelGGL(k_dl_fill_miss, dim3(grid), blk, 0, s, counts + cur, A.qcap, H);
                } else if (int rc = lr::trace_closest_queue(sc->bvh, &rs, glimit, counts + cur, &hs, s)) {
                    return rc;
                }
                hipLaunchKernelGGL(k_dl_shade, dim3(grid), blk, 0, s, S, P, gen, H, geom_args(sc), q[cur],
                                   counts + cur, ps, q[nxt], counts + nxt, sq, counts + 2, path0, A);
                const uint8_t* occ = nullptr;
                if (A.slots && sc->bvh->info.n_nodes) {
                    lr_rays_soa ss{sq.r[0], sq.r[1], sq.r[2], sq.r[3], sq.r[4], sq.r[5], sq.r[6], sq.r[7]};
                    if (int rc = lr::trace_any_queue(sc->bvh, &ss, A.scap, counts + 2, sc->d_dlocc, s)) return rc;
                    occ = sc->d_dlocc;
                }
                if (A.slots)
                    hipLaunchKernelGGL(k_dl_resolve, dim3(grid_max), blk, 0, s, counts + 2, A.scap, sq, occ, A.lq,
                                       rays_dev);
            }
            LR_HIP_CHECK(hipGetLastError());
            uint32_t host[8];
            LR_HIP_CHECK(hipMemcpyAsync(host, counts, sizeof(host), hipMemcpyDeviceToHost, s));
            LR_HIP_CHECK(hipStreamSynchronize(s));
            if (host[3] == 0 && host[6] == 0) {
                uint64_t r64;
                std::memcpy(&r64, host + 4, sizeof(r64));
                rays += r64;
                break;
            }
            // the chunk's radiance is discarded (the camera kernel resets it) and the chunk repeated
            if (host[3]) qcap *= 2;
            if (host[6]) scap *= 2;
            if (qcap > (1ull << 30) || scap > (1ull << 30))
                return lr::fail(LR_E_INVALID, "lr_render_direct: ray tree too large; use a smaller chunk_paths");
        }
        hipLaunchKernelGGL(k_dl_fold, dim3((n + 255) / 256), blk, 0, s, n, A.lq, ps);
        hipLaunchKernelGGL(k_splat_tiles, dim3(splat_blocks(npix, P.spp * P.nl)), blk, 0, s, S, P, pix0, npix, ps,
                           film_xyz, film_w);
        LR_HIP_CHECK(hipGetLastError());
        ++chunks;
    }
    if (stats) {
        LR_HIP_CHECK(hipStreamSynchronize(s));
        stats->paths = n_pix * ppp;
        stats->rays = rays;
        stats->chunks = chunks;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return LR_OK;
}

}  // extern "C"

extern "C" {

int lr_halton_permutation(uint32_t index, uint32_t* prime, uint32_t* out, uint32_t capacity) {
    if (!prime) return lr::fail(LR_E_INVALID, "NULL prime");
    if (index >= kHaltonPrimes) return lr::fail(LR_E_INVALID, "Halton prime index must be < 284");
    const HaltonTables& T = halton_tables();
    const uint32_t p = T.off[index + 1] - T.off[index];
    *prime = p;
    if (out && capacity >= p)
        for (uint32_t k = 0; k < p; ++k) out[k] = T.perm[T.off[index] + k];
    return LR_OK;
}

}  // extern "C"
