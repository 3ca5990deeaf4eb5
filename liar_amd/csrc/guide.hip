// guide.hip -- path-guiding network on gfx950 MFMA: a 6 -> 64 (frequency
// encoding) -> 5 x 64 (ReLU) -> 40 MLP whose 40 outputs parameterise an
// 8-lobe von Mises-Fisher mixture over incident directions.  BF16 weights and
// activations, FP32 accumulation (v_mfma_f32_16x16x16_bf16), FP32 master
// weights with Adam.
//
// Kernels
//   k_infer       fully fused forward on v_mfma_f32_16x16x32_bf16 in the
//                 transposed form Z^T = W^T-tiles x X^T: the accumulator of
//                 one layer (lane = row, 4 consecutive features) is, after
//                 bias + ReLU + bf16 packing, exactly the B operand of the
//                 next, so activations never leave registers; weights sit in
//                 LDS in per-lane fragment order; each lane ends holding two
//                 whole lobes and the softmax needs two cross-lane exchanges
//   k_train_fb    fused forward + loss + backward data path for 16-row tiles;
//                 writes activations A_l and gradients dZ_l feature-major
//                 (bf16) for the weight-gradient GEMMs
//   k_wgrad       dW_l = dZ_l^T A_{l-1} on MFMA, split over row chunks
//                 (partials), plus bias gradients
//   k_adam        fixed-order reduction of the partials + Adam + bf16 copy
//
// MFMA 16x16x16 bf16 lane layout (wave64, lane = i + 16 j): A operand lane
// holds A[i][4j..4j+3], B operand lane holds B[4j..4j+3][i], accumulator lane
// holds D[4j..4j+3][i].  16x16x32 is the same with 8 K values per lane
// (A[i][8j..8j+7], B[8j..8j+7][i]) and the same accumulator layout; the
// contraction is order-free, so A and B may place any K permutation in those
// slots as long as both use the same one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lr_geom.h"
#include "lr_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 6;        // raw inputs
constexpr int kW = 64;        // width (encoding and hidden)
constexpr int kHidden = 5;    // hidden layers
constexpr int kOut = 40;      // 8 lobes x (logit, mean xyz, concentration)
constexpr int kSel = 2;       // strategy-selection logits (outputs 40, 41)
constexpr int kOutPad = 48;   // padded to whole 16-column tiles
constexpr int kMix = 48;      // floats per inferred row: 40 lobe values, c_light, c_guide, c_bsdf, 0 x 5
constexpr float kSelFloor = 0.025f;  // least probability of each strategy
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

// Fragment order of the bf16 weights for k_infer.  Fragment (l, t, kb) is the
// 16x32 A operand of layer l, output tile t (16 rows), K block kb (32 inputs):
// 64 lanes x 8 bf16, lane (i, j) element e holding W_l[out 16t+i][in kslot].
// K slot (kb, j, e) is input 32kb + 4j + e for e < 4 and 32kb + 16 + 4j + e - 4
// otherwise: the features an accumulator pair (tiles 2kb, 2kb+1) leaves in
// lane j.  The output layer's physical rows are permuted so that lane j of a
// row ends with lobes 2j and 2j+1: physical row 16t + m holds logical output
// 10(m / 4) + 4t + m % 4 when 4t + m % 4 < 10, and 40 + 2(m / 4) + m % 4 - 2
// otherwise (so lane 0 also holds the selection logits 40 and 41).
constexpr int kFrags = kHidden * 4 * 2 + 3 * 2;  // 46 fragments of 512 bf16
static_assert(kFrags * 512 == kWtElems, "fragment count");
__host__ __device__ constexpr int frag_id(int l, int t, int kb) { return l < kHidden ? (l * 4 + t) * 2 + kb : 40 + t * 2 + kb; }
__host__ __device__ constexpr int kslot_feature(int kb, int j, int e) { return 32 * kb + (e < 4 ? 4 * j + e : 16 + 4 * j + e - 4); }
// logical output of the output layer's physical row p
__host__ __device__ constexpr int out_logical(int p) {
    return (4 * (p >> 4) + (p & 3)) < 10 ? 10 * ((p & 15) >> 2) + 4 * (p >> 4) + (p & 3)
                                         : kOut + 2 * ((p & 15) >> 2) + (p & 3) - 2;
}
// fragment-order index of logical weight Wt_l[n][k]
__host__ __device__ inline int frag_index(int l, int n, int k) {
    const int kb = k >> 5, c = k & 31;
    const int j = (c & 15) >> 2, e = (c & 3) + (c >= 16 ? 4 : 0);
    int t, i;
    if (l < kHidden) {
        t = n >> 4;
        i = n & 15;
    } else {
        // lane group and slot of this output
        const int jj = n < kOut ? n / 10 : (n - kOut) >> 1, sl = n < kOut ? n % 10 : 10 + ((n - kOut) & 1);
        t = sl >> 2;
        i = 4 * jj + (sl & 3);
    }
    return frag_id(l, t, kb) * 512 + (i + 16 * j) * 8 + e;
}

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

// Lobe concentration from its raw output s: kappa = 0.1 + e^min(s, 7.6)
// (0.1 .. ~2000).  The log scale lets a lobe sharpen to the ~0.01 sr targets
// of small bright sources in few steps; the cap keeps e^{kappa (c - 1)}
// consistent between sampling and evaluation in FP32.
constexpr float kKappaLogMax = 7.6f;
__device__ __forceinline__ float kappa_of(float s) { return 0.1f + expf(fminf(s, kKappaLogMax)); }
__device__ __forceinline__ float dkappa_ds(float s) { return s < kKappaLogMax ? expf(s) : 0.0f; }

// Frequency encoding of input row v[6] (each in [0, 1]) into 64 features:
// f < 6 raw (2v - 1); 6 <= f < 54 sin / cos(2^k pi v_d), k < 4; f = 54 is 1;
// the rest 0.
__device__ __forceinline__ float encode(const float* v, int f) {
    if (f < kIn) return 2.0f * v[f] - 1.0f;
    if (f < kIn + kIn * 2 * kFreq) {
        const int idx = f - kIn;
        const int d = idx / (2 * kFreq);
        const int k = (idx % (2 * kFreq)) >> 1;
        // sin / cos(2^k pi v) = v_sin / v_cos of 2^(k-1) v revolutions (|arg| <= 4)
        const float rev = 0.5f * (float)(1 << k) * v[d];
        return (idx & 1) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
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
// unit means, concentrations kappa = 0.1 + e^min(s, 7.6).
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
        L[k].kappa = kappa_of(o[5 * k + 4]);
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
constexpr int kInferWaves = 4;
constexpr int kInferBlock = 64 * kInferWaves;
constexpr int kInferRT = 2;  // 16-row tiles per wave iteration (share each weight fragment)
constexpr int kInferRows = 16 * kInferRT;
constexpr size_t kInferSmem = sizeof(uint4) * kFrags * 64 + sizeof(float) * kBiasElems;

__device__ __forceinline__ bf16x8 pack_relu(const f32x4& lo, const f32x4& hi) {
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        r[e] = (__bf16)fmaxf(lo[e], 0.0f);
        r[e + 4] = (__bf16)fmaxf(hi[e], 0.0f);
    }
    return r;
}

__device__ __forceinline__ f32x4 bias4(const float* s_b, int idx) {
    const float4 b = *reinterpret_cast<const float4*>(s_b + idx);
    return f32x4{b.x, b.y, b.z, b.w};
}

// Strategy probabilities from the two selection logits: r_l = sigmoid(a0),
// r_g = sigmoid(a1); (light, guide, BSDF) = floor + (1 - 3 floor) *
// (r_l, (1 - r_l) r_g, (1 - r_l)(1 - r_g)).
__host__ __device__ inline void selection_probs(float a0, float a1, float* c) {
    const float rl = 1.0f / (1.0f + expf(-a0)), rg = 1.0f / (1.0f + expf(-a1));
    const float sc = 1.0f - 3.0f * kSelFloor;
    c[0] = kSelFloor + sc * rl;
    c[1] = kSelFloor + sc * (1.0f - rl) * rg;
    c[2] = kSelFloor + sc * (1.0f - rl) * (1.0f - rg);
}

// Per lane: the two lobes (logit, mean xyz, raw concentration) of one row in
// o[0..9] (lane 0: selection logits in o[10..11]); the softmax over the row's
// 8 logits spans lanes i, i+16, i+32, i+48.
__device__ __forceinline__ void write_lobes(const float* o, bool valid, bool sel_lane, float* __restrict__ m) {
    float mx = fmaxf(o[0], o[5]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float ea = __expf(o[0] - mx), eb = __expf(o[5] - mx);
    float sum = ea + eb;
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    if (!valid) return;
    const float inv = 1.0f / sum;
    float r[10];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float* q = o + 5 * h;
        const float nrm = 1.0f / (sqrtf(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f);
        r[5 * h + 0] = (h ? eb : ea) * inv;
        r[5 * h + 1] = q[1] * nrm;
        r[5 * h + 2] = q[2] * nrm;
        r[5 * h + 3] = q[3] * nrm;
        r[5 * h + 4] = kappa_of(q[4]);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) reinterpret_cast<float2*>(m)[c] = make_float2(r[2 * c], r[2 * c + 1]);
    if (sel_lane) {  // lane 0 of the row: floats 40..47 of the row
        float c[3];
        selection_probs(o[10], o[11], c);
        float4* s4 = reinterpret_cast<float4*>(m + kOut);
        s4[0] = make_float4(c[0], c[1], c[2], 0.0f);
        s4[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// mix[row][48] = lobes (pi, mu xyz, kappa) x 8, then (c_light, c_guide,
// c_bsdf, 0 x 5), for rows < n (n from n_dev when given).  frag: the fragment-order bf16 weights; bias: FP32 biases in the
// master layout (the output layer's are permuted here).
__global__ __launch_bounds__(kInferBlock) void k_infer(const uint16_t* __restrict__ frag, const float* __restrict__ bias,
                                                       const float* __restrict__ in, uint32_t n,
                                                       const uint32_t* __restrict__ n_dev, float* __restrict__ mix) {
    extern __shared__ uint8_t smem[];
    uint4* s_f = reinterpret_cast<uint4*>(smem);
    float* s_b = reinterpret_cast<float*>(smem + sizeof(uint4) * kFrags * 64);
    if (n_dev) n = *n_dev;
    if (blockIdx.x * (uint32_t)(kInferWaves * kInferRows) >= n) return;  // no tile for this block
    for (int e = threadIdx.x; e < kFrags * 64; e += kInferBlock) s_f[e] = reinterpret_cast<const uint4*>(frag)[e];
    for (int e = threadIdx.x; e < kBiasElems; e += kInferBlock) {
        const int p = e - kHidden * kW;
        s_b[e] = p < 0 ? bias[e] : bias[kHidden * kW + out_logical(p)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    const uint32_t tiles = (n + kInferRows - 1) / kInferRows;
    for (uint32_t tile = blockIdx.x * kInferWaves + wave; tile < tiles; tile += gridDim.x * kInferWaves) {
        // keeps the fragment reads inside the loop: hoisted, the 46 fragments
        // would pin 184 registers and cut occupancy to one wave per SIMD
        asm volatile("" ::: "memory");
        bf16x8 x[kInferRT][2];
#pragma unroll
        for (int rt = 0; rt < kInferRT; ++rt) {
            const uint32_t row = tile * kInferRows + 16 * rt + i;
            float v[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (row < n) {
                const float4 a = reinterpret_cast<const float4*>(in)[2 * (size_t)row];
                const float4 b = reinterpret_cast<const float4*>(in)[2 * (size_t)row + 1];
                v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
            }
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int e = 0; e < 8; ++e) x[rt][kb][e] = (__bf16)encode(v, kslot_feature(kb, j, e));
        }
        f32x4 acc[kInferRT][4];
#pragma unroll
        for (int l = 0; l < kHidden; ++l) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const f32x4 b = bias4(s_b, l * kW + 16 * t + 4 * j);
#pragma unroll
                for (int rt = 0; rt < kInferRT; ++rt) acc[rt][t] = b;
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    const bf16x8 w = __builtin_bit_cast(bf16x8, s_f[frag_id(l, t, kb) * 64 + lane]);
#pragma unroll
                    for (int rt = 0; rt < kInferRT; ++rt)
                        acc[rt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, x[rt][kb], acc[rt][t], 0, 0, 0);
                }
            }
#pragma unroll
            for (int rt = 0; rt < kInferRT; ++rt) {
                x[rt][0] = pack_relu(acc[rt][0], acc[rt][1]);
                x[rt][1] = pack_relu(acc[rt][2], acc[rt][3]);
            }
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const f32x4 b = bias4(s_b, kHidden * kW + 16 * t + 4 * j);
#pragma unroll
            for (int rt = 0; rt < kInferRT; ++rt) acc[rt][t] = b;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const bf16x8 w = __builtin_bit_cast(bf16x8, s_f[frag_id(kHidden, t, kb) * 64 + lane]);
#pragma unroll
                for (int rt = 0; rt < kInferRT; ++rt)
                    acc[rt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, x[rt][kb], acc[rt][t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int rt = 0; rt < kInferRT; ++rt) {
            const uint32_t row = tile * kInferRows + 16 * rt + i;
            float o[12];
#pragma unroll
            for (int sl = 0; sl < 12; ++sl) o[sl] = acc[rt][sl >> 2][sl & 3];
            write_lobes(o, row < n, j == 0, mix + (size_t)row * kMix + 10 * j);
        }
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
                // strategy selection: -w log(c_l p_l + c_g p_g + c_b p_b) with the
                // record's light / BSDF pdfs (inputs 6, 7) and the guide pdf held
                // fixed; only when the record carries them (p_l + p_b > 0)
                const float pl = in[(size_t)row * 8 + 6], pb = in[(size_t)row * 8 + 7];
                if (pl + pb > 0.0f) {
                    const float pg = expf(logq);
                    float c[3];
                    selection_probs(o[kOut], o[kOut + 1], c);
                    const float pmix = c[0] * pl + c[1] * pg + c[2] * pb;
                    const float rl = 1.0f / (1.0f + expf(-o[kOut])), rg = 1.0f / (1.0f + expf(-o[kOut + 1]));
                    const float f = -w * (1.0f - 3.0f * kSelFloor) / pmix;
                    loss -= w * logf(pmix);
                    g[kOut] = f * rl * (1.0f - rl) * (pl - rg * pg - (1.0f - rg) * pb);
                    g[kOut + 1] = f * (1.0f - rl) * rg * (1.0f - rg) * (pg - pb);
                }
                for (int k = 0; k < kLobes; ++k) {
                    const float resp = expf(lp[k] - logq);  // responsibility
                    const float c = L[k].mx * wx + L[k].my * wy + L[k].mz * wz;
                    const float kap = L[k].kappa;
                    // d(-log q)/d logit_k
                    g[5 * k + 0] = -w * (resp - L[k].pi);
                    // d(-log q)/d s_k:  d log vmf / d kappa = 1/k - 2/(e^{2k} - 1) + (c - 1)
                    const float e2 = expm1f(2.0f * kap);
                    const float dlk = 1.0f / kap - 2.0f / e2 + (c - 1.0f);
                    g[5 * k + 4] = -w * resp * dlk * dkappa_ds(o[5 * k + 4]);
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
// wave w computes output rows n in [16 w, 16 w + 16) for all 4 k tiles.  The
// bias sums come from the G values each lane already holds as its MFMA A
// operand, reduced over the 4 lanes of a row at the end.
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
#pragma unroll
            for (int e = 0; e < 4; ++e) bsum += bf2f((uint16_t)a[e]);
        }
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        // acc[t][r] = dWt[16 wave + 4 j + r][16 t + i]
        for (int t = 0; t < 4; ++t)
            for (int r = 0; r < 4; ++r)
                out[wt_offset(l) + (16 * wave + 4 * j + r) * kW + 16 * t + i] = acc[t][r];
        if (j == 0) out[bias_offset(l) + 16 * wave + i] = bsum;
    }
}

// Fixed-order sum of the chunk partials, Adam, bf16 copy.
__global__ void k_adam(const float* __restrict__ partial, uint32_t chunks, float* __restrict__ w,
                       float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ wbf,
                       uint16_t* __restrict__ frag, float lr, float b1, float b2, float eps, float c1, float c2,
                       float* __restrict__ grad_out) {
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
    if (p < (uint32_t)kWtElems) {
        const uint16_t h = f2bf(nw);
        wbf[p] = h;
        const int l = min((int)p / (kW * kW), kHidden), r = (int)p - wt_offset(l);
        frag[frag_index(l, r / kW, r % kW)] = h;
    }
}

// Fixed-order sum of x[0..n) by one block of 1024 threads: each thread sums
// a strided set of 4-wide groups (4 loads in flight), then a tree in LDS.
// Elements before the first 16-byte boundary are summed on their own.
__device__ float block_sum(const float* __restrict__ x, uint32_t n) {
    __shared__ float s[1024];
    const uint32_t t = threadIdx.x;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    const uint32_t head = min(n, (uint32_t)((4u - ((uint32_t)(reinterpret_cast<uintptr_t>(x) >> 2) & 3u)) & 3u));
    if (t < head) a2 = x[t];
    x += head;
    n -= head;
    const uint32_t n4 = n >> 2;
    const float4* x4 = reinterpret_cast<const float4*>(x);
    uint32_t g = t;
    for (; g + 3 * 1024 < n4; g += 4 * 1024) {
        const float4 p = x4[g], q = x4[g + 1024], r = x4[g + 2048], u = x4[g + 3072];
        a0 += (p.x + p.y) + (p.z + p.w);
        a1 += (q.x + q.y) + (q.z + q.w);
        a2 += (r.x + r.y) + (r.z + r.w);
        a3 += (u.x + u.y) + (u.z + u.w);
    }
    for (; g < n4; g += 1024) {
        const float4 p = x4[g];
        a0 += (p.x + p.y) + (p.z + p.w);
    }
    if (t < (n & 3u)) a1 += x[4 * n4 + t];
    s[t] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)t < w) s[t] += s[t + w];
        __syncthreads();
    }
    return s[0];
}

// 1 / sum(weights)
__global__ __launch_bounds__(1024) void k_wsum(const float* __restrict__ weight, uint32_t n, float* __restrict__ inv_wsum) {
    const float total = block_sum(weight, n);
    if (threadIdx.x == 0) *inv_wsum = total > 0.0f ? 1.0f / total : 0.0f;
}

__global__ __launch_bounds__(1024) void k_sum(const float* __restrict__ x, uint32_t n, float* __restrict__ out) {
    const float total = block_sum(x, n);
    if (threadIdx.x == 0) *out = total;
}

constexpr size_t kTrainSmem = sizeof(uint16_t) * kLayers * kW * kWS + sizeof(float) * kBiasElems +
                              kWaves * (8 * sizeof(uint16_t) * kTileRows * kWS + sizeof(float) * kTileRows * kOutPad);

}  // namespace

struct lr_guide {
    lr_context* ctx = nullptr;
    float* w = nullptr;       // fp32 master params
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wbf = nullptr;  // bf16 copy of the Wt part (master layout)
    uint16_t* frag = nullptr; // bf16 copy of the Wt part in fragment order
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
    void* ptrs[] = {g->w, g->m, g->v, g->wbf, g->frag, g->A, g->G, g->partial, g->row_loss, g->scalars};
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
    std::vector<uint16_t> bf(kWtElems), fr(kWtElems, 0);
    for (int p = 0; p < kWtElems; ++p) bf[p] = host_f2bf(host[p]);
    for (int l = 0; l < kLayers; ++l)
        for (int n = 0; n < layer_nout(l); ++n)
            for (int k = 0; k < kW; ++k) fr[frag_index(l, n, k)] = bf[wt_offset(l) + n * kW + k];
    LR_HIP_CHECK(hipMemcpy(g->w, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(g->wbf, bf.data(), sizeof(uint16_t) * kWtElems, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(g->frag, fr.data(), sizeof(uint16_t) * kWtElems, hipMemcpyHostToDevice));
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
    if (e == hipSuccess) e = hipMalloc(&g->frag, sizeof(uint16_t) * kWtElems);
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
            for (int k = 0; k < kW; ++k)
                h[wt_offset(l) + n * kW + k] = (l == kLayers - 1 && n >= kOut + kSel) ? 0.0f : lim * uni();
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
    const uint32_t tiles = (n + kInferRows - 1) / kInferRows;
    const uint32_t grid = std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves, (uint32_t)g->ctx->n_cu * 3);
    hipLaunchKernelGGL(k_infer, dim3(std::max(1u, grid)), dim3(kInferBlock), kInferSmem, s, g->frag,
                       g->w + kWtElems, inputs, n, n_dev, mixture);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

}  // extern "C"

namespace {

// Row chunking of the weight-gradient pass: 16 .. 64 chunks of >= 256 rows,
// enough blocks to spread a small batch over the chip, few partials for the
// Adam reduction.
void wgrad_chunks(uint32_t n_pad, uint32_t& chunk_rows, uint32_t& chunks) {
    const uint32_t want = std::min<uint32_t>(64, std::max<uint32_t>(16, n_pad / 1024));
    chunk_rows = std::max<uint32_t>(256, ((n_pad + want - 1) / want + 15) & ~15u);
    chunks = (n_pad + chunk_rows - 1) / chunk_rows;
}

int grow_training(lr_guide* g, uint32_t n_pad, uint32_t chunks, hipStream_t s) {
    if (g->cap_rows >= n_pad && g->cap_chunks >= chunks) return LR_OK;
    n_pad = std::max<uint32_t>(n_pad, (uint32_t)g->cap_rows);
    chunks = std::max(chunks, g->cap_chunks);
    {
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
    return LR_OK;
}

}  // namespace

int lr_guide_reserve(lr_guide* g, uint32_t n, hipStream_t s) {
    if (!g || n == 0) return LR_OK;
    const uint32_t n_pad = (n + 15) & ~15u;
    uint32_t chunk_rows, chunks;
    wgrad_chunks(n_pad, chunk_rows, chunks);
    return grow_training(g, n_pad, chunks, s);
}

extern "C" {

int lr_guide_train_step(lr_guide* g, const float* inputs, const float* dirs, const float* weights,
                                   uint32_t n, float* loss_out, float* grad_out, void* stream) {
    if (!g || (n && (!inputs || !dirs || !weights))) return lr::fail(LR_E_INVALID, "NULL argument");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(g->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t n_pad = (n + 15) & ~15u;
    uint32_t chunk_rows, chunks;
    wgrad_chunks(n_pad, chunk_rows, chunks);
    if (int rc = grow_training(g, n_pad, chunks, s)) return rc;
    hipLaunchKernelGGL(k_wsum, dim3(1), dim3(1024), 0, s, weights, n, g->scalars);
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
                       g->wbf, g->frag, g->lr, b1, b2, 1e-8f, c1, c2, grad_out);
    if (loss_out) hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, g->row_loss, n, loss_out);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

}  // extern "C"
