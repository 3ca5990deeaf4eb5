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
//   k_train_fb    fused forward + loss + backward data path, register
//                 resident like k_infer (backward on transposed fragments);
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
// physical row of the output layer's logical output n (inverse of out_logical)
__host__ __device__ inline int out_physical(int n) {
    const int jj = n < kOut ? n / 10 : (n - kOut) >> 1, sl = n < kOut ? n % 10 : 10 + ((n - kOut) & 1);
    return 16 * (sl >> 2) + 4 * jj + (sl & 3);
}
// K slot (kb, j, e) of feature / physical row c: the inverse of kslot_feature
__host__ __device__ inline void kslot_of(int c, int& kb, int& j, int& e) {
    kb = c >> 5;
    const int r = c & 31;
    j = (r & 15) >> 2;
    e = (r & 3) + (r >= 16 ? 4 : 0);
}
// fragment-order index of logical weight Wt_l[n][k]
__host__ __device__ inline int frag_index(int l, int n, int k) {
    int kb, j, e;
    kslot_of(k, kb, j, e);
    const int p = l < kHidden ? n : out_physical(n);
    return frag_id(l, p >> 4, kb) * 512 + ((p & 15) + 16 * j) * 8 + e;
}

// Transposed fragments for the backward pass.  Fragment (l, t, kb), l = 1..5,
// is the A operand of dH_l^T = W_l^T-tiles x dZ_l^T: lane (i, j) element e holds
// Wt_l[n][k] with k = 16 t + i (an input of layer l) and n the (physical, for
// the output layer) output in K slot (kb, j, e); outputs past 48 are zeros.
constexpr int kFragsT = kHidden * 4 * 2;  // 40 fragments of 512 bf16
__host__ __device__ constexpr int fragT_id(int l, int t, int kb) { return ((l - 1) * 4 + t) * 2 + kb; }
__host__ __device__ inline int fragT_index(int l, int n, int k) {
    int kb, j, e;
    kslot_of(l < kHidden ? n : out_physical(n), kb, j, e);
    return fragT_id(l, k >> 4, kb) * 512 + ((k & 15) + 16 * j) * 8 + e;
}

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

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

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {
    const uint2 u = *reinterpret_cast<const uint2*>(p);  // 8-byte aligned
    bf16x4 r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
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
// (dZ of layer l, the output layer's rows in logical order).  n_pad is a
// multiple of 16.
//
// Same register-resident form as k_infer, one 16-row tile per wave
// iteration: the forward keeps every layer's packed activations in
// registers; each lane computes the loss gradient of its two lobes (lane 0
// of a row also that of the selection logits); the backward pass runs
// dH^T = W^T-tiles x dZ^T on the transposed fragments, and the ReLU mask
// comes from the activation registers.  A and dZ go to global memory for the
// weight-gradient GEMMs.
constexpr int kTrainWaves = 4;
constexpr int kTrainBlock = 64 * kTrainWaves;
constexpr size_t kTrainSmem = sizeof(uint4) * (kFrags + kFragsT) * 64 + sizeof(float) * kBiasElems;

// Bits of element e of a bf16 vector (by a whole-vector bit cast: a bit cast
// of one element of the __bf16 vector stored the wrong elements).
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ uint16_t bf_bits(const bf16x8& v, int e) { return __builtin_bit_cast(u16x8, v)[e]; }

__device__ __forceinline__ bf16x8 pack8(const float* v) {
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
    return r;
}

__global__ __launch_bounds__(kTrainBlock) void k_train_fb(const uint16_t* __restrict__ frag,
                                                          const uint16_t* __restrict__ fragT,
                                                          const float* __restrict__ bias, const float* __restrict__ in,
                                                          const float* __restrict__ dirs,
                                                          const float* __restrict__ weight, uint32_t n, uint32_t n_pad,
                                                          const float* __restrict__ inv_wsum, uint16_t* __restrict__ A,
                                                          uint16_t* __restrict__ G, float* __restrict__ row_loss) {
    extern __shared__ uint8_t smem[];
    uint4* s_f = reinterpret_cast<uint4*>(smem);
    uint4* s_t = s_f + kFrags * 64;
    float* s_b = reinterpret_cast<float*>(s_t + kFragsT * 64);
    for (int e = threadIdx.x; e < kFrags * 64; e += kTrainBlock) s_f[e] = reinterpret_cast<const uint4*>(frag)[e];
    for (int e = threadIdx.x; e < kFragsT * 64; e += kTrainBlock) s_t[e] = reinterpret_cast<const uint4*>(fragT)[e];
    for (int e = threadIdx.x; e < kBiasElems; e += kTrainBlock) {
        const int p = e - kHidden * kW;
        s_b[e] = p < 0 ? bias[e] : bias[kHidden * kW + out_logical(p)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    const float wnorm = *inv_wsum;
    const uint32_t tiles = n_pad / kTileRows;
    for (uint32_t tile = blockIdx.x * kTrainWaves + wave; tile < tiles; tile += gridDim.x * kTrainWaves) {
        asm volatile("" ::: "memory");  // fragment reads stay inside the loop (as in k_infer)
        const uint32_t row = tile * kTileRows + i;
        const bool valid = row < n;
        float v[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (valid) {
            const float4 a = reinterpret_cast<const float4*>(in)[2 * (size_t)row];
            const float4 b = reinterpret_cast<const float4*>(in)[2 * (size_t)row + 1];
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        }
        // forward: x[0] the encoding, x[l] the output of hidden layer l - 1
        bf16x8 x[kLayers][2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) x[0][kb][e] = (__bf16)encode(v, kslot_feature(kb, j, e));
        f32x4 acc[4];
#pragma unroll
        for (int l = 0; l < kHidden; ++l) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[t] = bias4(s_b, l * kW + 16 * t + 4 * j);
#pragma unroll
                for (int kb = 0; kb < 2; ++kb)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, s_f[frag_id(l, t, kb) * 64 + lane]), x[l][kb], acc[t], 0, 0, 0);
            }
            x[l + 1][0] = pack_relu(acc[0], acc[1]);
            x[l + 1][1] = pack_relu(acc[2], acc[3]);
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            acc[t] = bias4(s_b, kHidden * kW + 16 * t + 4 * j);
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, s_f[frag_id(kHidden, t, kb) * 64 + lane]), x[kHidden][kb], acc[t], 0,
                    0, 0);
        }
        float o[12];
#pragma unroll
        for (int sl = 0; sl < 12; ++sl) o[sl] = acc[sl >> 2][sl & 3];

        // loss and its gradient: this lane's lobes 2j, 2j + 1 (slots 0..4, 5..9)
        const float wx = valid ? dirs[3 * (size_t)row] : 0.0f;
        const float wy = valid ? dirs[3 * (size_t)row + 1] : 0.0f;
        const float wz = valid ? dirs[3 * (size_t)row + 2] : 1.0f;
        float mx = fmaxf(o[0], o[5]);
        mx = fmaxf(mx, __shfl_xor(mx, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 32));
        const float ea = expf(o[0] - mx), eb = expf(o[5] - mx);
        float esum = ea + eb;
        esum += __shfl_xor(esum, 16);
        esum += __shfl_xor(esum, 32);
        float pi[2] = {ea / esum, eb / esum}, kap[2], cc[2], nm[2], mu[2][3], lp[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float* q = o + 5 * h;
            nm[h] = sqrtf(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f;
            mu[h][0] = q[1] / nm[h];
            mu[h][1] = q[2] / nm[h];
            mu[h][2] = q[3] / nm[h];
            kap[h] = kappa_of(q[4]);
            cc[h] = mu[h][0] * wx + mu[h][1] * wy + mu[h][2] * wz;
            lp[h] = logf(pi[h]) + log_vmf(kap[h], cc[h]);
        }
        float lmax = fmaxf(lp[0], lp[1]);
        lmax = fmaxf(lmax, __shfl_xor(lmax, 16));
        lmax = fmaxf(lmax, __shfl_xor(lmax, 32));
        float ls = expf(lp[0] - lmax) + expf(lp[1] - lmax);
        ls += __shfl_xor(ls, 16);
        ls += __shfl_xor(ls, 32);
        const float logq = lmax + logf(ls);
        const float w = valid ? weight[row] * wnorm : 0.0f;
        float g[12];
#pragma unroll
        for (int sl = 0; sl < 12; ++sl) g[sl] = 0.0f;
        float loss = -w * logq;
        if (valid) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float resp = expf(lp[h] - logq);  // responsibility
                g[5 * h + 0] = -w * (resp - pi[h]);
                // d log vmf / d kappa = 1/k - 2/(e^{2k} - 1) + (c - 1)
                const float dlk = 1.0f / kap[h] - 2.0f / expm1f(2.0f * kap[h]) + (cc[h] - 1.0f);
                g[5 * h + 4] = -w * resp * dlk * dkappa_ds(o[5 * h + 4]);
                // d(-log q)/d m = -resp * kappa * (I - mu mu^T) w / |m|
                const float f = -w * resp * kap[h] / nm[h];
                g[5 * h + 1] = f * (wx - cc[h] * mu[h][0]);
                g[5 * h + 2] = f * (wy - cc[h] * mu[h][1]);
                g[5 * h + 3] = f * (wz - cc[h] * mu[h][2]);
            }
            // strategy selection (lane 0 of the row): -w log(c_l p_l + c_g q + c_b p_b),
            // q held fixed, only when the record carries p_l, p_b (inputs 6, 7)
            const float pl = v[6], pb = v[7];
            if (j == 0 && pl + pb > 0.0f) {
                const float pg = expf(logq);
                float c[3];
                selection_probs(o[10], o[11], c);
                const float pmix = c[0] * pl + c[1] * pg + c[2] * pb;
                const float rl = 1.0f / (1.0f + expf(-o[10])), rg = 1.0f / (1.0f + expf(-o[11]));
                const float f = -w * (1.0f - 3.0f * kSelFloor) / pmix;
                loss -= w * logf(pmix);
                g[10] = f * rl * (1.0f - rl) * (pl - rg * pg - (1.0f - rg) * pb);
                g[11] = f * (1.0f - rl) * rg * (1.0f - rg) * (pg - pb);
            }
            if (j == 0) row_loss[row] = loss;
        }
        // activations -> A[l] (feature-major)
#pragma unroll
        for (int l = 0; l < kLayers; ++l) {
            uint16_t* Al = A + (size_t)l * kW * n_pad + row;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    Al[(size_t)kslot_feature(kb, j, e) * n_pad] = bf_bits(x[l][kb], e);
        }
        // output-layer dZ (bf16) -> G[5] in logical row order, and as the B operand
        float gb[16];
#pragma unroll
        for (int sl = 0; sl < 12; ++sl) gb[sl] = (float)(__bf16)g[sl];
#pragma unroll
        for (int sl = 12; sl < 16; ++sl) gb[sl] = 0.0f;
        {
            uint16_t* G5 = G + (size_t)kHidden * kW * n_pad + row;
#pragma unroll
            for (int sl = 0; sl < 12; ++sl)
                G5[(size_t)out_logical(16 * (sl >> 2) + 4 * j + (sl & 3)) * n_pad] =
                    __builtin_bit_cast(uint16_t, (__bf16)gb[sl]);
        }
        bf16x8 dz[2] = {pack8(gb), pack8(gb + 8)};
        // backward: dH_l = W_l^T dZ_l, dZ_{l-1} = dH_l * (x[l] > 0)
#pragma unroll
        for (int l = kLayers - 1; l >= 1; --l) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < 2; ++kb)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, s_t[fragT_id(l, t, kb) * 64 + lane]), dz[kb], acc[t], 0, 0, 0);
            }
            float d[16];
            uint16_t* Gl = G + (size_t)(l - 1) * kW * n_pad + row;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float hv = (float)x[l][t >> 1][(t & 1) * 4 + r];
                    const __bf16 b = (__bf16)(hv > 0.0f ? acc[t][r] : 0.0f);
                    d[4 * t + r] = (float)b;
                    Gl[(size_t)(16 * t + 4 * j + r) * n_pad] = __builtin_bit_cast(uint16_t, b);
                }
            dz[0] = pack8(d);
            dz[1] = pack8(d + 8);
        }
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
                       float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ frag, uint16_t* __restrict__ fragT, float lr, float b1, float b2,
                       float eps, float c1, float c2, float* __restrict__ grad_out) {
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
        const int l = min((int)p / (kW * kW), kHidden), r = (int)p - wt_offset(l);
        frag[frag_index(l, r / kW, r % kW)] = h;
        if (l > 0) fragT[fragT_index(l, r / kW, r % kW)] = h;
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


}  // namespace

struct lr_guide {
    lr_context* ctx = nullptr;
    float* w = nullptr;       // fp32 master params
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* frag = nullptr; // bf16 copy of the Wt part in fragment order
    uint16_t* fragT = nullptr; // bf16 copy of layers 1..5 in transposed fragment order
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
    void* ptrs[] = {g->w, g->m, g->v, g->frag, g->fragT, g->A, g->G, g->partial, g->row_loss, g->scalars};
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
    std::vector<uint16_t> bf(kWtElems), fr(kWtElems, 0), ft(kFragsT * 512, 0);
    for (int p = 0; p < kWtElems; ++p) bf[p] = host_f2bf(host[p]);
    for (int l = 0; l < kLayers; ++l)
        for (int n = 0; n < layer_nout(l); ++n)
            for (int k = 0; k < kW; ++k) {
                fr[frag_index(l, n, k)] = bf[wt_offset(l) + n * kW + k];
                if (l > 0) ft[fragT_index(l, n, k)] = bf[wt_offset(l) + n * kW + k];
            }
    LR_HIP_CHECK(hipMemcpy(g->w, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(g->frag, fr.data(), sizeof(uint16_t) * kWtElems, hipMemcpyHostToDevice));
    LR_HIP_CHECK(hipMemcpy(g->fragT, ft.data(), sizeof(uint16_t) * ft.size(), hipMemcpyHostToDevice));
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
    if (e == hipSuccess) e = hipMalloc(&g->frag, sizeof(uint16_t) * kWtElems);
    if (e == hipSuccess) e = hipMalloc(&g->fragT, sizeof(uint16_t) * kFragsT * 512);
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
    if (g->A && g->G && g->partial && g->row_loss && g->cap_rows >= n_pad && g->cap_chunks >= chunks)
        return LR_OK;
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
        // capacities drop with the buffers: a failed allocation below leaves
        // a state that the next call grows again
        g->cap_rows = 0;
        g->cap_chunks = 0;
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
    // one workgroup per CU fits the 88 KB of fragments in LDS
    const uint32_t grid = std::min<uint32_t>((tiles + kTrainWaves - 1) / kTrainWaves, (uint32_t)g->ctx->n_cu);
    hipLaunchKernelGGL(k_train_fb, dim3(std::max(1u, grid)), dim3(kTrainBlock), kTrainSmem, s, g->frag, g->fragT,
                       g->w + kWtElems, inputs, dirs, weights, n, n_pad, g->scalars, g->A, g->G, g->row_loss);
    // rows n .. n_pad-1 are written by k_train_fb with zero gradients, so
    // they add nothing to the weight gradients
    hipLaunchKernelGGL(k_wgrad, dim3(chunks, kLayers), dim3(kBlock), 0, s, g->A, g->G, n_pad, chunk_rows, g->partial);
    g->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)g->step), c2 = 1.0f - std::pow(b2, (float)g->step);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, g->partial, chunks, g->w, g->m, g->v,
                       g->frag, g->fragT, g->lr, b1, b2, 1e-8f, c1, c2, grad_out);
    if (loss_out) hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, g->row_loss, n, loss_out);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

}  // extern "C"
