// guide.hip -- path-guiding network on gfx950 MFMA: a 6 -> 64 (frequency
// encoding) -> 5 x 64 (ReLU) -> 40 MLP whose 40 outputs parameterise an
// 8-lobe von Mises-Fisher mixture over incident directions.  BF16 weights and
// activations, FP32 accumulation (v_mfma_f32_16x16x16_bf16), FP32 master
// weights with Adam.
//
// Kernels
//   k_infer       fully fused forward: weights in LDS, each wave runs 16-row
//                 tiles through all layers, activations stay in LDS; outputs
//                 the processed mixture (weights, unit means, concentrations)
//   k_train_fb    fused forward + loss + backward data path for 16-row tiles;
//                 writes activations A_l and gradients dZ_l feature-major
//                 (bf16) for the weight-gradient GEMMs
//   k_wgrad       dW_l = dZ_l^T A_{l-1} on MFMA, split over row chunks
//                 (partials), plus bias gradients
//   k_adam        fixed-order reduction of the partials + Adam + bf16 copy
//
// MFMA 16x16x16 bf16 lane layout (wave64, lane = i + 16 j): A operand lane
// holds A[i][4j..4j+3], B operand lane holds B[4j..4j+3][i], accumulator lane
// holds D[4j..4j+3][i].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lr_geom.h"
#include "lr_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 6;        // raw inputs
constexpr int kW = 64;        // width (encoding and hidden)
constexpr int kHidden = 5;    // hidden layers
constexpr int kOut = 40;      // 8 lobes x (logit, mean xyz, concentration)
constexpr int kOutPad = 48;   // padded to whole 16-column tiles
constexpr int kLobes = 8;
constexpr int kLayers = kHidden + 1;  // weight matrices
constexpr int kFreq = 4;      // frequencies per input
constexpr int kRowPad = 8;    // LDS row padding (bf16 elements)
constexpr int kWS = kW + kRowPad;  // LDS row stride of Wt and activations

// Parameter layout (FP32 master, and the bf16 copy): per layer l, Wt_l
// [nout][64] row-major (nout = 64, or 48 for the output layer), then biases.
constexpr int kWtElems = kHidden * kW * kW + kOutPad * kW;       // 23552
constexpr int kBiasElems = kHidden * kW + kOutPad;               // 368
constexpr int kParams = kWtElems + kBiasElems;                   // 23920

__host__ __device__ constexpr int layer_nout(int l) { return l < kHidden ? kW : kOutPad; }
__host__ __device__ constexpr int wt_offset(int l) { return l * kW * kW; }
__host__ __device__ constexpr int bias_offset(int l) { return kWtElems + l * kW; }

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Orders a wave's LDS writes before its later reads by other lanes of the
// same wave (memory fence at wavefront scope + no code motion across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float softplus(float x) { return x > 20.0f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }

// Frequency encoding of input row v[6] (each in [0, 1]) into 64 features:
// f < 6 raw (2v - 1); 6 <= f < 54 sin / cos(2^k pi v_d), k < 4; f = 54 is 1;
// the rest 0.
__device__ __forceinline__ float encode(const float* v, int f) {
    if (f < kIn) return 2.0f * v[f] - 1.0f;
    if (f < kIn + kIn * 2 * kFreq) {
        const int idx = f - kIn;
        const int d = idx / (2 * kFreq);
        const int k = (idx % (2 * kFreq)) >> 1;
        const float a = 3.14159265358979f * (float)(1 << k) * v[d];
        return (idx & 1) ? cosf(a) : sinf(a);
    }
    return f == kIn + kIn * 2 * kFreq ? 1.0f : 0.0f;
}

// Loads every layer's bf16 Wt into LDS (padded rows) and biases (fp32).
__device__ void load_weights(const uint16_t* __restrict__ wbf, const float* __restrict__ bias,
                             uint16_t* s_w, float* s_b) {
    for (int l = 0; l < kLayers; ++l) {
        const int nout = layer_nout(l);
        for (int e = threadIdx.x; e < nout * kW; e += blockDim.x) {
            const int n = e / kW, k = e % kW;
            s_w[l * kW * kWS + n * kWS + k] = wbf[wt_offset(l) + e];
        }
    }
    for (int e = threadIdx.x; e < kBiasElems; e += blockDim.x) s_b[e] = bias[e];
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {
    const uint2 u = *reinterpret_cast<const uint2*>(p);  // 8-byte aligned
    bf16x4 r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}

// out tile (16 rows x 16 cols at n0) = act[16 x 64] * W_l[:, n0..n0+15]
__device__ __forceinline__ f32x4 tile_fwd(const uint16_t* act, const uint16_t* wl, int n0, int i, int j) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < kW / 16; ++kb) {
        const bf16x4 a = ld4(act + i * kWS + kb * 16 + 4 * j);
        const bf16x4 b = ld4(wl + (n0 + i) * kWS + kb * 16 + 4 * j);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Runs the hidden layers for one 16-row tile.  act_in holds the encoding;
// acts[l] (l = 0..4) receive post-ReLU hidden activations (bf16).
__device__ __forceinline__ void forward_hidden(const uint16_t* s_w, const float* s_b, uint16_t* const* acts,
                                               int i, int j) {
    for (int l = 0; l < kHidden; ++l) {
        const uint16_t* in = acts[l];
        uint16_t* out = acts[l + 1];
        const uint16_t* wl = s_w + l * kW * kWS;
        f32x4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = tile_fwd(in, wl, 16 * t, i, j);
        wave_sync();
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float z = acc[t][r] + s_b[l * kW + 16 * t + i];
                out[(4 * j + r) * kWS + 16 * t + i] = f2bf(fmaxf(z, 0.0f));
            }
        wave_sync();
    }
}

// Mixture processing of one row's 40 raw outputs (fp32) -> weights pi,
// unit means, concentrations kappa = softplus(s) + 0.1.
struct Lobe {
    float pi, mx, my, mz, kappa;
};

__device__ inline void process_row(const float* o, Lobe* L) {
    float mx = -INFINITY;
    for (int k = 0; k < kLobes; ++k) mx = fmaxf(mx, o[5 * k]);
    float sum = 0.0f;
    for (int k = 0; k < kLobes; ++k) {
        L[k].pi = expf(o[5 * k] - mx);
        sum += L[k].pi;
    }
    for (int k = 0; k < kLobes; ++k) {
        L[k].pi /= sum;
        const float x = o[5 * k + 1], y = o[5 * k + 2], z = o[5 * k + 3];
        const float n = sqrtf(x * x + y * y + z * z) + 1e-12f;
        L[k].mx = x / n;
        L[k].my = y / n;
        L[k].mz = z / n;
        L[k].kappa = softplus(o[5 * k + 4]) + 0.1f;
    }
}

// log vMF density at cosine c = mu . w:  log(k / (2 pi (1 - e^{-2k}))) + k (c - 1)
__device__ inline float log_vmf(float kappa, float c) {
    return logf(kappa) - logf(6.283185307179586f * (-expm1f(-2.0f * kappa))) + kappa * (c - 1.0f);
}

constexpr int kTileRows = 16;
constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;

// ---------------------------------------------------------------------------
// inference
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_infer(const uint16_t* __restrict__ wbf, const float* __restrict__ bias,
                                                  const float* __restrict__ in, uint32_t n,
                                                  const uint32_t* __restrict__ n_dev, float* __restrict__ mix) {
    extern __shared__ uint8_t smem[];
    uint16_t* s_w = reinterpret_cast<uint16_t*>(smem);
    float* s_b = reinterpret_cast<float*>(smem + sizeof(uint16_t) * kLayers * kW * kWS);
    uint8_t* wave_base = smem + sizeof(uint16_t) * kLayers * kW * kWS + sizeof(float) * kBiasElems;
    if (n_dev) n = *n_dev;
    load_weights(wbf, bias, s_w, s_b);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    // per wave: two ping-pong activation buffers (16 x kWS bf16) + outputs (16 x 48 fp32)
    const size_t act_bytes = sizeof(uint16_t) * kTileRows * kWS;
    uint8_t* wb = wave_base + wave * (2 * act_bytes + sizeof(float) * kTileRows * kOutPad);
    uint16_t* bufA = reinterpret_cast<uint16_t*>(wb);
    uint16_t* bufB = reinterpret_cast<uint16_t*>(wb + act_bytes);
    float* s_out = reinterpret_cast<float*>(wb + 2 * act_bytes);
    const uint32_t tiles = (n + kTileRows - 1) / kTileRows;
    for (uint32_t tile = blockIdx.x * kWaves + wave; tile < tiles; tile += gridDim.x * kWaves) {
        const uint32_t row0 = tile * kTileRows;
        // encoding: lane (i, j) -> row i, features 16 j .. 16 j + 15
        {
            const uint32_t row = row0 + i;
            float v[kIn];
            for (int d = 0; d < kIn; ++d) v[d] = row < n ? in[(size_t)row * 8 + d] : 0.0f;
            for (int f = 16 * j; f < 16 * j + 16; ++f) bufA[i * kWS + f] = f2bf(encode(v, f));
        }
        wave_sync();
        uint16_t* cur = bufA;
        uint16_t* nxt = bufB;
        for (int l = 0; l < kHidden; ++l) {
            const uint16_t* wl = s_w + l * kW * kWS;
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = tile_fwd(cur, wl, 16 * t, i, j);
            wave_sync();
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    nxt[(4 * j + r) * kWS + 16 * t + i] = f2bf(fmaxf(acc[t][r] + s_b[l * kW + 16 * t + i], 0.0f));
            wave_sync();
            uint16_t* tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
        {
            const uint16_t* wl = s_w + kHidden * kW * kWS;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const f32x4 acc = tile_fwd(cur, wl, 16 * t, i, j);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    s_out[(4 * j + r) * kOutPad + 16 * t + i] = acc[r] + s_b[kHidden * kW + 16 * t + i];
            }
        }
        wave_sync();
        // lanes 0..15 turn their row's outputs into mixture parameters
        if (j == 0) {
            const uint32_t row = row0 + i;
            if (row < n) {
                Lobe L[kLobes];
                process_row(s_out + i * kOutPad, L);
                float* m = mix + (size_t)row * kOut;
                for (int k = 0; k < kLobes; ++k) {
                    m[5 * k + 0] = L[k].pi;
                    m[5 * k + 1] = L[k].mx;
                    m[5 * k + 2] = L[k].my;
                    m[5 * k + 3] = L[k].mz;
                    m[5 * k + 4] = L[k].kappa;
                }
            }
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// training: fused forward + loss + backward data path
// ---------------------------------------------------------------------------
// Global scratch (feature-major, bf16): A[l][64][n_pad] for l = 0..5 (0 is
// the encoding, 1..5 hidden outputs) and G[l][nout_l][n_pad] for l = 0..5
// (dZ of layer l).  n_pad is a multiple of 16.
__global__ __launch_bounds__(kBlock) void k_train_fb(const uint16_t* __restrict__ wbf, const float* __restrict__ bias,
                                                     const float* __restrict__ in, const float* __restrict__ dirs,
                                                     const float* __restrict__ weight, uint32_t n, uint32_t n_pad,
                                                     const float* __restrict__ inv_wsum, uint16_t* __restrict__ A,
                                                     uint16_t* __restrict__ G, float* __restrict__ row_loss) {
    extern __shared__ uint8_t smem[];
    uint16_t* s_w = reinterpret_cast<uint16_t*>(smem);
    float* s_b = reinterpret_cast<float*>(smem + sizeof(uint16_t) * kLayers * kW * kWS);
    uint8_t* wave_base = smem + sizeof(uint16_t) * kLayers * kW * kWS + sizeof(float) * kBiasElems;
    load_weights(wbf, bias, s_w, s_b);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    const size_t act_bytes = sizeof(uint16_t) * kTileRows * kWS;
    // per wave: 6 activation buffers, 2 gradient buffers, fp32 outputs
    uint8_t* wb = wave_base + wave * (8 * act_bytes + sizeof(float) * kTileRows * kOutPad);
    uint16_t* acts[kLayers];
    for (int l = 0; l < kLayers; ++l) acts[l] = reinterpret_cast<uint16_t*>(wb + l * act_bytes);
    uint16_t* gbuf[2] = {reinterpret_cast<uint16_t*>(wb + 6 * act_bytes),
                         reinterpret_cast<uint16_t*>(wb + 7 * act_bytes)};
    float* s_out = reinterpret_cast<float*>(wb + 8 * act_bytes);
    const float wnorm = *inv_wsum;
    const uint32_t tiles = n_pad / kTileRows;
    for (uint32_t tile = blockIdx.x * kWaves + wave; tile < tiles; tile += gridDim.x * kWaves) {
        const uint32_t row0 = tile * kTileRows;
        {
            const uint32_t row = row0 + i;
            float v[kIn];
            for (int d = 0; d < kIn; ++d) v[d] = row < n ? in[(size_t)row * 8 + d] : 0.0f;
            for (int f = 16 * j; f < 16 * j + 16; ++f) acts[0][i * kWS + f] = f2bf(encode(v, f));
        }
        wave_sync();
        forward_hidden(s_w, s_b, acts, i, j);
        {
            const uint16_t* wl = s_w + kHidden * kW * kWS;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const f32x4 acc = tile_fwd(acts[kHidden], wl, 16 * t, i, j);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    s_out[(4 * j + r) * kOutPad + 16 * t + i] = acc[r] + s_b[kHidden * kW + 16 * t + i];
            }
        }
        wave_sync();
        // loss and its gradient w.r.t. the 40 raw outputs (lanes 0..15, one row each)
        if (j == 0) {
            const uint32_t row = row0 + i;
            float g[kOutPad];
            for (int c = 0; c < kOutPad; ++c) g[c] = 0.0f;
            float loss = 0.0f;
            if (row < n) {
                const float* o = s_out + i * kOutPad;
                Lobe L[kLobes];
                process_row(o, L);
                const float wx = dirs[3 * (size_t)row], wy = dirs[3 * (size_t)row + 1], wz = dirs[3 * (size_t)row + 2];
                float lp[kLobes], lmax = -INFINITY;
                for (int k = 0; k < kLobes; ++k) {
                    const float c = L[k].mx * wx + L[k].my * wy + L[k].mz * wz;
                    lp[k] = logf(L[k].pi) + log_vmf(L[k].kappa, c);
                    lmax = fmaxf(lmax, lp[k]);
                }
                float s = 0.0f;
                for (int k = 0; k < kLobes; ++k) s += expf(lp[k] - lmax);
                const float logq = lmax + logf(s);
                const float w = weight[row] * wnorm;
                loss = -w * logq;
                for (int k = 0; k < kLobes; ++k) {
                    const float resp = expf(lp[k] - logq);  // responsibility
                    const float c = L[k].mx * wx + L[k].my * wy + L[k].mz * wz;
                    const float kap = L[k].kappa;
                    // d(-log q)/d logit_k
                    g[5 * k + 0] = -w * (resp - L[k].pi);
                    // d(-log q)/d s_k:  d log vmf / d kappa = 1/k - 2/(e^{2k} - 1) + (c - 1)
                    const float e2 = expm1f(2.0f * kap);
                    const float dlk = 1.0f / kap - 2.0f / e2 + (c - 1.0f);
                    g[5 * k + 4] = -w * resp * dlk * sigmoid(o[5 * k + 4]);
                    // d(-log q)/d m_k = -resp * kappa * (I - mu mu^T) w / |m|
                    const float x = o[5 * k + 1], y = o[5 * k + 2], z = o[5 * k + 3];
                    const float nm = sqrtf(x * x + y * y + z * z) + 1e-12f;
                    const float px = wx - c * L[k].mx, py = wy - c * L[k].my, pz = wz - c * L[k].mz;
                    const float f = -w * resp * kap / nm;
                    g[5 * k + 1] = f * px;
                    g[5 * k + 2] = f * py;
                    g[5 * k + 3] = f * pz;
                }
            }
            if (row < n) row_loss[row] = loss;
            for (int c = 0; c < kOutPad; ++c) gbuf[0][i * kWS + c] = f2bf(g[c]);
        }
        wave_sync();
        // dZ of the output layer -> global G[5] (feature-major)
        {
            uint16_t* Gl = G + (size_t)kHidden * kW * n_pad;
            for (int c = j; c < kOutPad; c += 4) {
                Gl[(size_t)c * n_pad + row0 + i] = gbuf[0][i * kWS + c];
            }
        }
        // backward data path: dA_{l} = dZ_{l+1} W_{l+1}^T, dZ_l = dA_l * (A_{l+1} > 0)
        int gcur = 0;
        for (int l = kHidden - 1; l >= 0; --l) {
            const uint16_t* wl = s_w + (l + 1) * kW * kWS;  // Wt_{l+1} [nout][64]
            const int kin = layer_nout(l + 1);              // contraction over layer l+1 outputs
            uint16_t* gin = gbuf[gcur];
            uint16_t* gout = gbuf[gcur ^ 1];
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int kb = 0; kb < kin / 16; ++kb) {
                    const bf16x4 a = ld4(gin + i * kWS + kb * 16 + 4 * j);
                    bf16x4 b;
#pragma unroll
                    for (int r = 0; r < 4; ++r) b[r] = (short)wl[(kb * 16 + 4 * j + r) * kWS + 16 * t + i];
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
                }
            }
            wave_sync();
            uint16_t* Gl = G + (size_t)l * kW * n_pad;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint16_t h[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 4 * j + r, col = 16 * t + i;
                    const bool on = bf2f(acts[l + 1][row * kWS + col]) > 0.0f;
                    h[r] = f2bf(on ? acc[t][r] : 0.0f);
                    gout[row * kWS + col] = h[r];
                }
                uint2 pk;
                pk.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
                pk.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
                *reinterpret_cast<uint2*>(Gl + (size_t)(16 * t + i) * n_pad + row0 + 4 * j) = pk;
            }
            wave_sync();
            gcur ^= 1;
        }
        // activations A_0..A_5 -> global (feature-major)
        for (int l = 0; l < kLayers; ++l) {
            uint16_t* Al = A + (size_t)l * kW * n_pad;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int col = 16 * t + i;
                uint2 pk;
                pk.x = (uint32_t)acts[l][(4 * j + 0) * kWS + col] | ((uint32_t)acts[l][(4 * j + 1) * kWS + col] << 16);
                pk.y = (uint32_t)acts[l][(4 * j + 2) * kWS + col] | ((uint32_t)acts[l][(4 * j + 3) * kWS + col] << 16);
                *reinterpret_cast<uint2*>(Al + (size_t)col * n_pad + row0 + 4 * j) = pk;
            }
        }
        wave_sync();
    }
}

// dWt_l[n][k] = sum_rows G_l[n][row] * A_l[k][row] over this block's row chunk;
// bias gradient db_l[n] = sum_rows G_l[n][row].  One block per (layer, chunk);
// wave w computes output rows n in [16 w, 16 w + 16) for all 4 k tiles.
__global__ __launch_bounds__(kBlock) void k_wgrad(const uint16_t* __restrict__ A, const uint16_t* __restrict__ G,
                                                  uint32_t n_pad, uint32_t chunk_rows, float* __restrict__ partial) {
    const int l = blockIdx.y;
    const uint32_t chunk = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    const int nout = layer_nout(l);
    const uint16_t* Gl = G + (size_t)l * kW * n_pad;
    const uint16_t* Al = A + (size_t)l * kW * n_pad;
    float* out = partial + (size_t)chunk * kParams;
    const uint32_t r0 = chunk * chunk_rows;
    const uint32_t r1 = min(n_pad, r0 + chunk_rows);
    if (16 * wave < nout) {
        f32x4 acc[4];
        for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        float bsum = 0.0f;
        for (uint32_t row = r0; row < r1; row += 16) {
            const bf16x4 a = ld4(Gl + (size_t)(16 * wave + i) * n_pad + row + 4 * j);  // A_mfma[n][row]
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x4 b = ld4(Al + (size_t)(16 * t + i) * n_pad + row + 4 * j);  // B_mfma[row][k]
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
            if (j == 0)
                for (int r = 0; r < 16; ++r) bsum += bf2f(Gl[(size_t)(16 * wave + i) * n_pad + row + r]);
        }
        // acc[t][r] = dWt[16 wave + 4 j + r][16 t + i]
        for (int t = 0; t < 4; ++t)
            for (int r = 0; r < 4; ++r)
                out[wt_offset(l) + (16 * wave + 4 * j + r) * kW + 16 * t + i] = acc[t][r];
        if (j == 0) out[bias_offset(l) + 16 * wave + i] = bsum;
    }
}

// Fixed-order sum of the chunk partials, Adam, bf16 copy.
__global__ void k_adam(const float* __restrict__ partial, uint32_t chunks, float* __restrict__ w,
                       float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ wbf, float lr, float b1,
                       float b2, float eps, float c1, float c2, float* __restrict__ grad_out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kParams) return;
    float g = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) g += partial[(size_t)c * kParams + p];
    if (grad_out) grad_out[p] = g;
    if (lr == 0.0f) return;  // gradient only
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float nw = w[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    w[p] = nw;
    if (p < (uint32_t)kWtElems) wbf[p] = f2bf(nw);
}

// 1 / sum(weights), fixed-order two-level reduction in one block.
__global__ void k_wsum(const float* __restrict__ weight, uint32_t n, float* __restrict__ inv_wsum) {
    __shared__ float s[256];
    float acc = 0.0f;
    for (uint32_t r = threadIdx.x; r < n; r += 256) acc += weight[r];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *inv_wsum = s[0] > 0.0f ? 1.0f / s[0] : 0.0f;
}

__global__ void k_sum(const float* __restrict__ x, uint32_t n, float* __restrict__ out) {
    __shared__ float s[256];
    float acc = 0.0f;
    for (uint32_t r = threadIdx.x; r < n; r += 256) acc += x[r];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

constexpr size_t kInferSmem = sizeof(uint16_t) * kLayers * kW * kWS + sizeof(float) * kBiasElems +
                              kWaves * (2 * sizeof(uint16_t) * kTileRows * kWS + sizeof(float) * kTileRows * kOutPad);
constexpr size_t kTrainSmem = sizeof(uint16_t) * kLayers * kW * kWS + sizeof(float) * kBiasElems +
                              kWaves * (8 * sizeof(uint16_t) * kTileRows * kWS + sizeof(float) * kTileRows * kOutPad);

}  // namespace

struct lr_guide {
    lr_context* ctx = nullptr;
    float* w = nullptr;       // fp32 master params
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wbf = nullptr;  // bf16 copy of the Wt part
    uint32_t step = 0;
    float lr = 1e-3f;
    // training scratch (grow-only)
    size_t cap_rows = 0;
    uint16_t* A = nullptr;
    uint16_t* G = nullptr;
    float* partial = nullptr;
    uint32_t cap_chunks = 0;
    float* row_loss = nullptr;
    float* scalars = nullptr;  // [0] 1/sum w, [1] loss sum
};

namespace {

void free_guide(lr_guide* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    void* ptrs[] = {g->w, g->m, g->v, g->wbf, g->A, g->G, g->partial, g->row_loss, g->scalars};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete g;
}

uint16_t host_f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int upload_params(lr_guide* g, const float* host) {
    std::vector<uint16_t> bf(kWtElems);
    for (int p = 0; p < kWtElems; ++p) bf[p] = host_f2bf(host[p]);
    LR_HIP_CHECK(hipMemcpy(g->w, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(g->wbf, bf.data(), sizeof(uint16_t) * kWtElems, hipMemcpyHostToDevice));
    return LR_OK;
}

}  // namespace

extern "C" {

int lr_guide_param_count(void) { return kParams; }

int lr_guide_create(lr_context* ctx, uint64_t seed, float lr, lr_guide** out) {
    if (!ctx || !out) return lr::fail(LR_E_INVALID, "NULL context or out");
    *out = nullptr;
    if (!(lr >= 0.0f)) return lr::fail(LR_E_INVALID, "learning rate must be >= 0");
    if (int rc = lr::bind(ctx)) return rc;
    LR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_infer),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kInferSmem));
    LR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_train_fb),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrainSmem));
    lr_guide* g = new (std::nothrow) lr_guide();
    if (!g) return lr::fail(LR_E_NOMEM, "out of host memory");
    g->ctx = ctx;
    g->lr = lr;
    hipError_t e = hipMalloc(&g->w, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&g->m, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&g->v, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&g->wbf, sizeof(uint16_t) * kWtElems);
    if (e == hipSuccess) e = hipMalloc(&g->scalars, sizeof(float) * 16);
    if (e == hipSuccess) e = hipMemset(g->m, 0, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMemset(g->v, 0, sizeof(float) * kParams);
    if (e != hipSuccess) {
        free_guide(g);
        return lr::fail(LR_E_DEVICE, std::string("lr_guide_create: ") + hipGetErrorString(e));
    }
    // He-uniform hidden weights, small output layer (near-uniform mixture), zero biases
    std::vector<float> h(kParams, 0.0f);
    uint32_t ctr = 0;
    auto uni = [&](void) {
        const lr::u4 r = lr::philox4x32((uint32_t)seed, (uint32_t)(seed >> 32), lr::u4{ctr++, 0x6a1du, 0u, 0u});
        return lr::u01(r.x) * 2.0f - 1.0f;
    };
    for (int l = 0; l < kLayers; ++l) {
        const int nout = layer_nout(l);
        const float lim = l < kHidden ? std::sqrt(6.0f / kW) : 0.1f * std::sqrt(6.0f / kW);
        for (int n = 0; n < nout; ++n)
            for (int k = 0; k < kW; ++k) h[wt_offset(l) + n * kW + k] = (l == kLayers - 1 && n >= kOut) ? 0.0f : lim * uni();
    }
    if (int rc = upload_params(g, h.data())) {
        free_guide(g);
        return rc;
    }
    *out = g;
    return LR_OK;
}

int lr_guide_destroy(lr_guide* g) {
    free_guide(g);
    return LR_OK;
}

int lr_guide_get_params(const lr_guide* g, float* host) {
    if (!g || !host) return lr::fail(LR_E_INVALID, "NULL argument");
    if (int rc = lr::bind(g->ctx)) return rc;
    LR_HIP_CHECK(hipMemcpy(host, g->w, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    return LR_OK;
}

int lr_guide_set_params(lr_guide* g, const float* host) {
    if (!g || !host) return lr::fail(LR_E_INVALID, "NULL argument");
    if (int rc = lr::bind(g->ctx)) return rc;
    LR_HIP_CHECK(hipMemset(g->m, 0, sizeof(float) * kParams));
    LR_HIP_CHECK(hipMemset(g->v, 0, sizeof(float) * kParams));
    g->step = 0;
    return upload_params(g, host);
}

int lr_guide_infer(lr_guide* g, const float* inputs, uint32_t n, const uint32_t* n_dev, float* mixture,
                   void* stream) {
    if (!g || (n && (!inputs || !mixture))) return lr::fail(LR_E_INVALID, "NULL argument");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(g->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t tiles = (n + kTileRows - 1) / kTileRows;
    const uint32_t grid = std::min<uint32_t>((tiles + kWaves - 1) / kWaves, (uint32_t)g->ctx->n_cu * 2);
    hipLaunchKernelGGL(k_infer, dim3(std::max(1u, grid)), dim3(kBlock), kInferSmem, s, g->wbf, g->w + kWtElems,
                       inputs, n, n_dev, mixture);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

int lr_guide_train_step(lr_guide* g, const float* inputs, const float* dirs, const float* weights, uint32_t n,
                        float* loss_out, float* grad_out, void* stream) {
    if (!g || (n && (!inputs || !dirs || !weights))) return lr::fail(LR_E_INVALID, "NULL argument");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(g->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t n_pad = (n + 15) & ~15u;
    const uint32_t chunk_rows = 1024;
    const uint32_t chunks = (n_pad + chunk_rows - 1) / chunk_rows;
    if (g->cap_rows < n_pad || g->cap_chunks < chunks) {
        LR_HIP_CHECK(hipStreamSynchronize(s));
        if (g->A) LR_HIP_CHECK(hipFree(g->A));
        if (g->G) LR_HIP_CHECK(hipFree(g->G));
        if (g->partial) LR_HIP_CHECK(hipFree(g->partial));
        if (g->row_loss) LR_HIP_CHECK(hipFree(g->row_loss));
        g->A = g->G = nullptr;
        g->partial = g->row_loss = nullptr;
        LR_HIP_CHECK(hipMalloc(&g->A, sizeof(uint16_t) * kLayers * kW * (size_t)n_pad));
        LR_HIP_CHECK(hipMalloc(&g->G, sizeof(uint16_t) * kLayers * kW * (size_t)n_pad));
        LR_HIP_CHECK(hipMalloc(&g->partial, sizeof(float) * kParams * (size_t)chunks));
        LR_HIP_CHECK(hipMalloc(&g->row_loss, sizeof(float) * (size_t)n_pad));
        g->cap_rows = n_pad;
        g->cap_chunks = chunks;
    }
    hipLaunchKernelGGL(k_wsum, dim3(1), dim3(256), 0, s, weights, n, g->scalars);
    const uint32_t tiles = n_pad / kTileRows;
    const uint32_t grid = std::min<uint32_t>((tiles + kWaves - 1) / kWaves, (uint32_t)g->ctx->n_cu * 2);
    hipLaunchKernelGGL(k_train_fb, dim3(std::max(1u, grid)), dim3(kBlock), kTrainSmem, s, g->wbf, g->w + kWtElems,
                       inputs, dirs, weights, n, n_pad, g->scalars, g->A, g->G, g->row_loss);
    // rows n .. n_pad-1 are written by k_train_fb with zero gradients, so
    // they add nothing to the weight gradients
    hipLaunchKernelGGL(k_wgrad, dim3(chunks, kLayers), dim3(kBlock), 0, s, g->A, g->G, n_pad, chunk_rows, g->partial);
    g->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)g->step), c2 = 1.0f - std::pow(b2, (float)g->step);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, g->partial, chunks, g->w, g->m, g->v,
                       g->wbf, g->lr, b1, b2, 1e-8f, c1, c2, grad_out);
    if (loss_out) hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, g->row_loss, n, loss_out);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

}  // extern "C"
