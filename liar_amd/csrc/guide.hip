// guide.hip -- path-guiding network on gfx950 MFMA: a 6 -> 64 (frequency
// encoding) -> 5 x 64 (ReLU) -> 48 MLP whose outputs parameterise an 8-lobe
// von Mises-Fisher mixture over incident directions, the strategy
// selection of one-sample MIS and an estimate of the radiance a vertex
// reflects towards the previous one (for Russian roulette).  BF16 weights and activations, FP32
// accumulation (v_mfma_f32_16x16x32_bf16), FP32 master weights with Adam.
//
// Kernels
//   k_infer       fully fused forward in the transposed form Z^T = W^T-tiles
//                 x X^T: the accumulator of one layer (lane = row, 4
//                 consecutive features) is, after bias + ReLU + bf16 packing,
//                 exactly the B operand of the next, so activations never
//                 leave registers; weights sit in LDS in per-lane fragment
//                 order; each lane ends holding two whole lobes and the
//                 softmax needs two cross-lane exchanges
//   k_train       one Adam step's rows (given, or gathered from the tracer's
//                 records): forward + loss + backward register resident like
//                 k_infer, and the weight gradients dW_l = dZ_l^T X_l on MFMA
//                 from a per-layer LDS stage, accumulated per workgroup
//   k_adam        fixed-order reduction of the workgroup partials, Adam, bf16
//                 copies
//
// MFMA 16x16x32 bf16 lane layout (wave64, lane = i + 16 j): A operand lane
// holds A[i][8j..8j+7], B operand lane holds B[8j..8j+7][i], accumulator lane
// holds D[4j..4j+3][i].  The contraction is order-free, so A and B may place
// any K permutation in those slots as long as both use the same one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lr_geom.h"
#include "lr_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 6;        // raw inputs
constexpr int kW = 64;        // width (encoding and hidden)
constexpr int kHidden = 5;    // hidden layers
constexpr int kOut = 40;      // 8 lobes x (logit, mean xyz, concentration)
constexpr int kSel = 2;       // strategy-selection logits (outputs 40, 41)
constexpr int kRad = 42;      // log of the reflected-radiance estimate
constexpr int kOutPad = 48;   // padded to whole 16-column tiles
constexpr int kMix = 48;      // floats per inferred row: 40 lobe values, c_light, c_guide, c_bsdf, radiance, 0 x 4
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

// Lobe concentration from its raw output s: kappa = 0.1 + e^min(s, 7.6)
// (0.1 .. ~2000).  The log scale lets a lobe sharpen to the ~0.01 sr targets
// of small bright sources in few steps; the cap keeps e^{kappa (c - 1)}
// consistent between sampling and evaluation in FP32.
constexpr float kKappaLogMax = 7.6f;
__device__ __forceinline__ float kappa_of(float s) { return 0.1f + expf(fminf(s, kKappaLogMax)); }
__device__ __forceinline__ float dkappa_ds(float s) { return s < kKappaLogMax ? expf(s) : 0.0f; }

// Reflected-radiance estimate from its raw output r: e^clamp(r, -20, 20).
// Trained on the records' one-sample estimates y = f |cos| L_i / p_mix with
// the relative loss (e^r - y')^2 / e^{2r}, the denominator held fixed and
// y' = min(y, 50 e^r): the gradient 2 (1 - min(y e^-r, 50)) is scale free and
// its zero in expectation is e^r = E[y].
constexpr float kRadLogMax = 20.0f;
constexpr float kRadRatioMax = 50.0f;
__device__ __forceinline__ float rad_of(float r) { return expf(fminf(fmaxf(r, -kRadLogMax), kRadLogMax)); }

// Frequency encoding of input row v[6] (each in [0, 1]) into 64 features,
// laid out so that every lane computes its 16 K slots with no indexing by
// lane: slot (kb, j, e) of the first layer's B operand (logical feature
// kslot_feature(kb, j, e)), s = 8 kb + e:
//   s < 12        sin (s even) / cos (s odd) of 2^j pi v_{s/2}
//   s >= 12, j=0  2 v_{s-12} - 1  (inputs 0..3)
//   s >= 12, j=1  2 v_4 - 1, 2 v_5 - 1, 1, 0
//   s >= 12, j>1  0
// (6 raw inputs, 48 sines / cosines, a constant, 9 zeros).  rev_scale =
// 2^(j-1): v_sin / v_cos take revolutions, |argument| <= 4.
__device__ __forceinline__ float encode_slot(const float* v, float rev_scale, int j, int kb, int e) {
    const int s = 8 * kb + e;
    if (s < 12) {
        const float rev = rev_scale * v[s >> 1];
        return (s & 1) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
    }
    const float first = 2.0f * v[s - 12] - 1.0f;
    const float second = s < 14 ? 2.0f * v[s - 8] - 1.0f : (s == 14 ? 1.0f : 0.0f);
    return j == 0 ? first : (j == 1 ? second : 0.0f);
}

// log vMF density at cosine c = mu . w:  log(k / (2 pi (1 - e^{-2k}))) + k (c - 1)
__device__ inline float log_vmf(float kappa, float c) {
    return logf(kappa) - logf(6.283185307179586f * (-expm1f(-2.0f * kappa))) + kappa * (c - 1.0f);
}

// Copies N16 16-byte words from global memory to LDS with the direct
// global-to-LDS load (global_load_lds_dwordx4: lane l of a wave writes LDS
// word base + l, no registers), every load of the block in flight at once;
// the caller's __syncthreads() waits for them (s_waitcnt vmcnt(0) precedes
// its barrier).  N16 is a multiple of 64, so every wave's words are whole.
// (A copy through registers held 22 words per thread for the training
// kernel, which the compiler kept in scratch.)
template <int BLOCK, int N16>
__device__ __forceinline__ void copy_to_lds(uint4* __restrict__ dst, const uint4* __restrict__ src) {
    static_assert(N16 % 64 == 0 && BLOCK % 64 == 0, "whole waves");
    const int wbase = (int)(threadIdx.x & ~63u), lane = (int)(threadIdx.x & 63u);
#pragma unroll
    for (int k = 0; k < (N16 + BLOCK - 1) / BLOCK; ++k) {
        const int e0 = wbase + k * BLOCK;  // wave-uniform
        if (e0 < N16) __builtin_amdgcn_global_load_lds(src + e0 + lane, dst + e0, 16, 0, 0);
    }
}

// ---------------------------------------------------------------------------
// inference
// ---------------------------------------------------------------------------
constexpr int kInferWaves = 4;
constexpr int kInferBlock = 64 * kInferWaves;
constexpr int kInferRT = 2;  // 16-row tiles per wave iteration (share each weight fragment)
constexpr int kInferRows = 16 * kInferRT;
constexpr size_t kInferSmem = sizeof(uint4) * kFrags * 64 + sizeof(float) * kBiasElems;

// bf16(max(x, 0)) as max(bf16(x), 0) on the bits: a negative bf16 or -0 is
// a negative int16, so one packed integer max clamps a converted pair (the
// same bits for every non-NaN x; fmaxf would also quieten signalling NaNs,
// two instructions per value).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_pair(float a, float b) {
    const bf16x2 r = __builtin_convertvector((f32x2){a, b}, bf16x2);  // one v_cvt_pk_bf16_f32
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, r), (s16x2)0));
}
__device__ __forceinline__ bf16x8 pack_relu(const f32x4& lo, const f32x4& hi) {
    return __builtin_bit_cast(bf16x8, make_uint4(relu_pair(lo[0], lo[1]), relu_pair(lo[2], lo[3]),
                                                 relu_pair(hi[0], hi[1]), relu_pair(hi[2], hi[3])));
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
// o[0..9] (lane 0: selection logits in o[10..11], lane 16: the radiance
// output in o[10]); the softmax over the row's 8 logits spans lanes i, i+16,
// i+32, i+48.
//
// The mixture goes to row-major rows of 48 floats (SOA false: float c of row r
// at mix[48 r + c]) or to 48 columns of ld floats (SOA true: at mix[c ld + r],
// so that a shading kernel's lanes read one value of consecutive rows
// together).  The head's reciprocals and square root are the hardware's
// approximate ones (1 ulp): sampling and evaluation read the same stored
// values, so they stay consistent.
template <bool SOA>
__device__ __forceinline__ void write_lobes(const float* o, bool valid, bool sel_lane, float* __restrict__ mix,
                                            uint32_t row, size_t ld) {
    const int j = (int)((threadIdx.x & 63u) >> 4);
    const float rad = __shfl(o[10], (int)(threadIdx.x & 15u) + 16);
    float mx = fmaxf(o[0], o[5]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float ea = __expf(o[0] - mx), eb = __expf(o[5] - mx);
    float sum = ea + eb;
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    if (!valid) return;
    const float inv = __builtin_amdgcn_rcpf(sum);
    float r[10];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float* q = o + 5 * h;
        const float nrm = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f);
        r[5 * h + 0] = (h ? eb : ea) * inv;
        r[5 * h + 1] = q[1] * nrm;
        r[5 * h + 2] = q[2] * nrm;
        r[5 * h + 3] = q[3] * nrm;
        r[5 * h + 4] = kappa_of(q[4]);
    }
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (sel_lane) {  // lane 0 of the row: floats 40..43 of the row (44..47 are 0)
        selection_probs(o[10], o[11], c);
        c[3] = rad_of(rad);
    }
    if (SOA) {
        float* m = mix + row;
#pragma unroll
        for (int k = 0; k < 10; ++k) m[(size_t)(10 * j + k) * ld] = r[k];
        if (sel_lane)
#pragma unroll
            for (int k = 0; k < 8; ++k) m[(size_t)(kOut + k) * ld] = k < 4 ? c[k] : 0.0f;
    } else {
        float* m = mix + (size_t)row * kMix;
#pragma unroll
        for (int k = 0; k < 5; ++k) reinterpret_cast<float2*>(m + 10 * j)[k] = make_float2(r[2 * k], r[2 * k + 1]);
        if (sel_lane) {
            float4* s4 = reinterpret_cast<float4*>(m + kOut);
            s4[0] = make_float4(c[0], c[1], c[2], c[3]);
            s4[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

// Per row: lobes (pi, mu xyz, kappa) x 8, then (c_light, c_guide, c_bsdf,
// reflected-radiance estimate, 0 x 4), for rows < n (n from n_dev when
// given), in the layout of write_lobes<SOA> (ld: the column length).  frag:
// the fragment-order bf16 weights; bias: FP32 biases in the master layout
// (the output layer's are permuted here).
template <bool SOA>
__global__ __launch_bounds__(kInferBlock) void k_infer(const uint16_t* __restrict__ frag, const float* __restrict__ bias,
                                                       const float* __restrict__ in, uint32_t n,
                                                       const uint32_t* __restrict__ n_dev, float* __restrict__ mix,
                                                       size_t ld) {
    extern __shared__ uint8_t smem[];
    uint4* s_f = reinterpret_cast<uint4*>(smem);
    float* s_b = reinterpret_cast<float*>(smem + sizeof(uint4) * kFrags * 64);
    if (n_dev) n = *n_dev;
    if (blockIdx.x * (uint32_t)(kInferWaves * kInferRows) >= n) return;  // no tile for this block
    copy_to_lds<kInferBlock, kFrags * 64>(s_f, reinterpret_cast<const uint4*>(frag));
    for (int e = threadIdx.x; e < kBiasElems; e += kInferBlock) {
        const int p = e - kHidden * kW;
        s_b[e] = p < 0 ? bias[e] : bias[kHidden * kW + out_logical(p)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    const float rev_scale = 0.5f * (float)(1 << j);
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
                for (int e = 0; e < 8; ++e) x[rt][kb][e] = (__bf16)encode_slot(v, rev_scale, j, kb, e);
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
            write_lobes<SOA>(o, row < n, j == 0, mix, row, ld);
        }
    }
}

// ---------------------------------------------------------------------------
// training: one fused launch per Adam step (rows -> forward -> loss ->
// backward -> weight gradients), then k_adam
// ---------------------------------------------------------------------------
// A workgroup of WAVES waves takes kGroupRows = 16 WAVES rows per iteration, 16
// per wave, in the register-resident form of k_infer: the forward keeps each
// layer's packed activations x[l] (the next layer's B operand); each lane
// computes the loss gradient of its two lobes (lane 0 of a row also that of
// the selection logits); the backward runs dH^T = W^T-tiles x dZ^T on the
// transposed fragments, the ReLU mask from the activation registers.
//
// Weight gradients dWt_l = dZ_l^T X_l contract over rows, which the
// registers hold one per lane; so per layer every wave stages its x[l] and
// dZ_l feature-major in LDS (a 16 x 16 transpose per fragment), and after one
// barrier wave w forms output tile w of dWt_l for all of the group's rows
// with v_mfma_f32_16x16x32_bf16 (A = dZ^T rows of LDS, B = X rows of LDS).
// The stage is double buffered, so one barrier per layer separates a
// buffer's reads from its next writes.  Each wave keeps its tile's
// accumulators for all six layers across iterations; the workgroup writes
// one partial (weight and bias gradients, sum of weights, sum of losses) at
// the end.  Nothing per row goes to global memory.
//
// The gradient is that of the weighted loss sum_r w_r l_r; k_adam divides it
// by sum_r w_r (Adam's step is nearly invariant to that scale, the reported
// gradient and loss are those of the normalised weights).
//
// Two shapes: 4 waves (64 rows per iteration, 1 wave per SIMD) finish a
// step of up to 4,096 rows sooner; 8 waves (128 rows, 2 waves per SIMD, each
// wave half the dWt tiles, half the partials) run larger steps faster.
constexpr size_t kTrainFragBytes = sizeof(uint4) * (kFrags + kFragsT) * 64 + sizeof(float) * kBiasElems;
template <int WAVES>
struct TrainShape {
    static constexpr int kBlock = 64 * WAVES;
    static constexpr int kGroupRows = 16 * WAVES;          // rows per workgroup iteration
    static constexpr int kKBlocks = kGroupRows / 32;       // 32-row MFMA K blocks per stage
    static constexpr int kWTiles = 16 / WAVES;             // dWt output tiles (of 16 per layer) per wave
    static constexpr int kStageStride = kGroupRows + 8;    // bf16 per feature row (+16 B spreads LDS banks)
    static constexpr int kStageOperand = kW * kStageStride;
    static constexpr size_t kSmem = kTrainFragBytes + 2 * 2 * sizeof(uint16_t) * kStageOperand;
    static_assert(kTrainFragBytes % 16 == 0 && (kStageStride * 2) % 16 == 0, "16-byte LDS reads");
    static_assert(kSmem <= 160 * 1024, "LDS");
};
constexpr uint32_t kTrainWideRows = 4096;  // steps above this many rows use 8 waves (8,192 rows: 23.3 against 25.0 us)
constexpr int kPartial = kParams + 16;  // per workgroup: gradients, then sum w, sum w * loss
constexpr int kMaxChunks = 1024;         // partial slots (workgroups of one step)

// Rows of one step: the rows themselves, or (GATHER) rows drawn uniformly
// with replacement from min(*count, cap) records by Philox keyed by the
// seed, counter (row, step, chunk), so that every bounce's records are
// represented whatever order they were appended in.
struct TrainRows {
    const float* in;    // 8 floats per row / record
    const float* dirs;  // 3
    const float* w;     // 1
    uint32_t n;         // rows in the step
    const uint32_t* count;
    uint32_t cap, seed_lo, seed_hi, chunk, step;
    // weight of each row's radiance loss (device scalar: the guide's radiance
    // weight times the rows' or records' mean weight); null: no radiance loss
    const float* rad_scale;
};

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

template <bool GATHER, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_train(const uint16_t* __restrict__ frag,
                                                       const uint16_t* __restrict__ fragT,
                                                       const float* __restrict__ bias, TrainRows R,
                                                       float* __restrict__ partial) {
    using T = TrainShape<WAVES>;
    constexpr int kTrainBlock = T::kBlock, kGroupRows = T::kGroupRows, kKBlocks = T::kKBlocks,
                  kWTiles = T::kWTiles, kStageStride = T::kStageStride, kStageOperand = T::kStageOperand;
    extern __shared__ uint8_t smem[];
    __shared__ float s_red[2 * WAVES];
    uint4* s_f = reinterpret_cast<uint4*>(smem);
    uint4* s_t = s_f + kFrags * 64;
    float* s_b = reinterpret_cast<float*>(s_t + kFragsT * 64);
    uint16_t* s_stage = reinterpret_cast<uint16_t*>(smem + kTrainFragBytes);
    copy_to_lds<kTrainBlock, kFrags * 64>(s_f, reinterpret_cast<const uint4*>(frag));
    copy_to_lds<kTrainBlock, kFragsT * 64>(s_t, reinterpret_cast<const uint4*>(fragT));
    for (int e = threadIdx.x; e < kBiasElems; e += kTrainBlock) {
        const int p = e - kHidden * kW;
        s_b[e] = p < 0 ? bias[e] : bias[kHidden * kW + out_logical(p)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
    const float rev_scale = 0.5f * (float)(1 << j);
    const uint32_t nrec = GATHER ? min(*R.count, R.cap) : 0u;
    const uint32_t groups = (R.n + kGroupRows - 1) / kGroupRows;
    // wave w forms output tiles (n tile w % 4, k tiles kWTiles (w / 4) + u) of
    // every layer's dWt; the waves of n tile w % 4 with w < 4 also its bias
    const int wn = wave & 3, wk = kWTiles * (wave >> 2);
    // per-lane stage bases: st_w writes row 16 wave + i of feature 4 j (+ a
    // constant); st_a reads the A operand (feature 16 wn + i, rows 8 j ..),
    // st_b the B operand (feature 16 wk + i, rows 8 j ..)
    uint16_t* const st_w = s_stage + 4 * j * kStageStride + 16 * wave + i;
    const uint16_t* const st_a = s_stage + (16 * wn + i) * kStageStride + 8 * j;
    const uint16_t* const st_b = s_stage + (16 * wk + i) * kStageStride + 8 * j;
    f32x4 accW[kLayers][kWTiles];
    float accB[kLayers];
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        accB[l] = 0.0f;
#pragma unroll
        for (int u = 0; u < kWTiles; ++u) accW[l][u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float sum_w = 0.0f, sum_loss = 0.0f;
    const float rad_s = R.rad_scale ? *R.rad_scale : 0.0f;
    for (uint32_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {  // uniform over the workgroup
        asm volatile("" ::: "memory");  // fragment reads stay inside the loop (as in k_infer)
        const uint32_t row = grp * kGroupRows + 16 * wave + i;
        const bool valid = row < R.n && (!GATHER || nrec > 0u);
        uint32_t src = row;
        if (GATHER && valid) {
            const lr::u4 h = lr::philox4x32(R.seed_lo ^ 0x5bd1e995u, R.seed_hi, lr::u4{row, R.step, R.chunk, 0x6e1du});
            src = (uint32_t)(((uint64_t)h.x * nrec) >> 32);
        }
        float v[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        float wx = 0.0f, wy = 0.0f, wz = 1.0f, w = 0.0f;
        if (valid) {
            const float4 a = reinterpret_cast<const float4*>(R.in)[2 * (size_t)src];
            const float4 b = reinterpret_cast<const float4*>(R.in)[2 * (size_t)src + 1];
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
            wx = R.dirs[3 * (size_t)src];
            wy = R.dirs[3 * (size_t)src + 1];
            wz = R.dirs[3 * (size_t)src + 2];
            w = R.w[src];
        }
        // forward: x[0] the encoding, x[l] the output of hidden layer l - 1
        bf16x8 x[kLayers][2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) x[0][kb][e] = (__bf16)encode_slot(v, rev_scale, j, kb, e);
        f32x4 acc[4];
#pragma unroll
        for (int l = 0; l < kHidden; ++l) {
            // each layer reads its own fragments: scheduled early, all 86 would
            // need 344 registers
            asm volatile("" ::: "memory");
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
        asm volatile("" ::: "memory");
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
        float g[12];
#pragma unroll
        for (int sl = 0; sl < 12; ++sl) g[sl] = 0.0f;
        if (valid) {
            float loss = -w * logq;
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
            if (j == 0) {
                sum_w += w;
                sum_loss += loss;
            }
            // reflected radiance (lane 16 of the row, output 42): rad_s (1 -
            // min(w / e^r, 50))^2, the record weight w being the estimate y;
            // trains output row 42 only (see dz_shared below)
            if (j == 1 && rad_s > 0.0f) {
                const float r = o[10];
                const float ratio = fminf(w / rad_of(r), kRadRatioMax);
                const float d = 1.0f - ratio;
                g[10] = (r > -kRadLogMax && r < kRadLogMax) ? rad_s * 2.0f * d : 0.0f;
                sum_loss += rad_s * d * d;
            }
        }
        // output-layer dZ in bf16: dz[kb][e] is physical output kslot_feature(kb, j, e)
        float gb[16];
#pragma unroll
        for (int sl = 0; sl < 12; ++sl) gb[sl] = g[sl];
#pragma unroll
        for (int sl = 12; sl < 16; ++sl) gb[sl] = 0.0f;
        bf16x8 dz[2] = {pack8(gb), pack8(gb + 8)};
        // the radiance estimate is a linear readout of the last hidden layer:
        // its loss trains output row 42 alone, not the layers the mixture and
        // the selection share (the output layer's backward drops it)
        if (j == 1) gb[10] = 0.0f;
        const bf16x8 dz_shared = pack8(gb + 8);
        // per layer, from the output layer down: stage X_l and dZ_l, form the
        // weight gradients, then dZ_{l-1} = (W_l^T dZ_l) * (x[l] > 0)
#pragma unroll
        for (int l = kLayers - 1; l >= 0; --l) {
            // stage buffer l & 1: X_l, then dZ_l; every offset below past the
            // per-lane bases is a compile-time constant
            const int buf = (l & 1) * 2 * kStageOperand;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                // element pair (2h, 2h + 1) of a packed vector is one dword:
                // low half to feature kslot_feature(kb, j, 2h), high half to the next
                const uint4 ux = __builtin_bit_cast(uint4, x[l][kb]), ud = __builtin_bit_cast(uint4, dz[kb]);
                const uint32_t wx4[4] = {ux.x, ux.y, ux.z, ux.w}, wd4[4] = {ud.x, ud.y, ud.z, ud.w};
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const int f0 = kslot_feature(kb, 0, 2 * h);  // feature less its 4 j
                    uint16_t* px = st_w + buf + f0 * kStageStride;
                    px[0] = (uint16_t)wx4[h];
                    px[kStageStride] = (uint16_t)(wx4[h] >> 16);
                    px[kStageOperand] = (uint16_t)wd4[h];
                    px[kStageOperand + kStageStride] = (uint16_t)(wd4[h] >> 16);
                }
            }
            __syncthreads();
            if (16 * wn < layer_nout(l)) {
#pragma unroll
                for (int rb = 0; rb < kKBlocks; ++rb) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(st_a + buf + kStageOperand + 32 * rb);
#pragma unroll
                    for (int u = 0; u < kWTiles; ++u) {
                        const bf16x8 b =
                            *reinterpret_cast<const bf16x8*>(st_b + buf + 16 * u * kStageStride + 32 * rb);
                        accW[l][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, accW[l][u], 0, 0, 0);
                    }
                    if (wk == 0) {
#pragma unroll
                        for (int e = 0; e < 8; ++e) accB[l] += (float)a[e];
                    }
                }
            }
            if (l >= 1) {
                if (l == kLayers - 1) dz[1] = dz_shared;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, s_t[fragT_id(l, t, kb) * 64 + lane]), dz[kb], acc[t], 0, 0, 0);
                }
                float d[16];
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float hv = (float)x[l][t >> 1][(t & 1) * 4 + q];
                        d[4 * t + q] = hv > 0.0f ? acc[t][q] : 0.0f;
                    }
                dz[0] = pack8(d);
                dz[1] = pack8(d + 8);
            }
        }
    }
    // partial of this workgroup: accW[l][u][q] = dWt_l[physical 16 wn + 4 j + q][16 (wk + u) + i]
    float* out = partial + (size_t)blockIdx.x * kPartial;
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        float bsum = accB[l];
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (16 * wn < layer_nout(l)) {
#pragma unroll
            for (int u = 0; u < kWTiles; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int p = 16 * wn + 4 * j + q;
                    const int n = l < kHidden ? p : out_logical(p);
                    out[wt_offset(l) + n * kW + 16 * (wk + u) + i] = accW[l][u][q];
                }
            if (j == 0 && wk == 0) {
                const int p = 16 * wn + i;
                out[bias_offset(l) + (l < kHidden ? p : out_logical(p))] = bsum;
            }
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        sum_w += __shfl_xor(sum_w, m);
        sum_loss += __shfl_xor(sum_loss, m);
    }
    if (lane == 0) {
        s_red[wave] = sum_w;
        s_red[WAVES + wave] = sum_loss;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f, b = 0.0f;
        for (int k = 0; k < WAVES; ++k) {
            a += s_red[k];
            b += s_red[WAVES + k];
        }
        out[kParams] = a;
        out[kParams + 1] = b;
    }
}

// Fixed-order sum of the workgroup partials, normalised by the batch's
// weight sum; Adam; both bf16 copies.  A block takes kAdamParams parameters;
// each parameter's kAdamLanes threads sum every kAdamLanes-th partial (four
// loads in flight, so a step's 128 partials take two rounds of loads), then
// the lane sums are added in a fixed tree.  loss_out (block 0): the
// normalised loss.
constexpr int kAdamParams = 16;
constexpr int kAdamLanes = 256 / kAdamParams;
__global__ __launch_bounds__(256) void k_adam(const float* __restrict__ partial, uint32_t chunks,
                                              float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                              uint16_t* __restrict__ frag, uint16_t* __restrict__ fragT, float lr,
                                              float b1, float b2, float eps, float c1, float c2,
                                              float* __restrict__ grad_out, float* __restrict__ loss_out) {
    __shared__ float s_ws[256], s_ls[256], s_g[256];
    float a = 0.0f, b = 0.0f;
    for (uint32_t c = threadIdx.x; c < chunks; c += 256) {
        a += partial[(size_t)c * kPartial + kParams];
        b += partial[(size_t)c * kPartial + kParams + 1];
    }
    s_ws[threadIdx.x] = a;
    s_ls[threadIdx.x] = b;
    const uint32_t pl = threadIdx.x & (kAdamParams - 1), cl = threadIdx.x / kAdamParams;
    const uint32_t p = blockIdx.x * kAdamParams + pl;
    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
    if (p < (uint32_t)kParams) {
        const float* q = partial + p;
        uint32_t c = cl;
        for (; c + 3 * kAdamLanes < chunks; c += 4 * kAdamLanes) {
            g0 += q[(size_t)c * kPartial];
            g1 += q[(size_t)(c + kAdamLanes) * kPartial];
            g2 += q[(size_t)(c + 2 * kAdamLanes) * kPartial];
            g3 += q[(size_t)(c + 3 * kAdamLanes) * kPartial];
        }
        for (; c < chunks; c += kAdamLanes) g0 += q[(size_t)c * kPartial];
    }
    s_g[threadIdx.x] = (g0 + g1) + (g2 + g3);
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            s_ws[threadIdx.x] += s_ws[threadIdx.x + k];
            s_ls[threadIdx.x] += s_ls[threadIdx.x + k];
        }
        // the lane sums of each parameter: s_g[pl + kAdamParams * lane], a tree over lanes
        if ((int)threadIdx.x < k && k >= kAdamParams) s_g[threadIdx.x] += s_g[threadIdx.x + k];
        __syncthreads();
    }
    const float inv = s_ws[0] > 0.0f ? 1.0f / s_ws[0] : 0.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) *loss_out = s_ls[0] * inv;
    if (cl != 0 || p >= (uint32_t)kParams) return;
    const float g = s_g[pl] * inv;
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

// Fixed-order sum of the first n floats of x (n = min(*count, cap) when
// count is given): block b of kSumBlocks adds its contiguous slice, strided
// over its threads and then as a tree in LDS; k_sum_finish adds the block
// sums in order and writes mul * sum (mul * sum / n with `mean`, 0 when n = 0).
constexpr int kSumBlocks = 256;
__global__ __launch_bounds__(256) void k_sum_slices(const float* __restrict__ x, uint32_t n_host,
                                                     const uint32_t* __restrict__ count, uint32_t cap,
                                                     float* __restrict__ part) {
    __shared__ float sh[256];
    const uint32_t n = count ? min(*count, cap) : n_host;
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * per, hi = min(n, lo + per);
    float a0 = 0.0f, a1 = 0.0f;
    uint32_t r = lo + threadIdx.x;
    for (; r + 256 < hi; r += 512) {
        a0 += x[r];
        a1 += x[r + 256];
    }
    if (r < hi) a0 += x[r];
    sh[threadIdx.x] = a0 + a1;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_sum_finish(const float* __restrict__ part, uint32_t n_host,
                                                     const uint32_t* __restrict__ count, uint32_t cap, float mul,
                                                     int mean, float* __restrict__ out) {
    __shared__ float sh[256];
    sh[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t n = count ? min(*count, cap) : n_host;
        *out = mean ? (n ? mul * sh[0] / (float)n : 0.0f) : mul * sh[0];
    }
}

// w = min(w, *lim) over the first min(*count, cap) records.
__global__ void k_clamp_weights(uint32_t cap, const uint32_t* __restrict__ count, const float* __restrict__ lim,
                                float* __restrict__ w) {
    const uint32_t n = min(*count, cap);
    const float l = *lim;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
        w[r] = fminf(w[r], l);
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
    float* partial = nullptr;  // training: one kPartial slot per workgroup (max_chunks)
    uint32_t max_chunks = 0;
    float rad_weight = 0.0f;   // weight of the radiance loss (0: the radiance output is not trained)
    // device scalars: [0] radiance-loss scale of the step, [1] weight-clamp
    // limit, then kSumBlocks partial sums
    float* scal = nullptr;
};

namespace {

void free_guide(lr_guide* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    void* ptrs[] = {g->w, g->m, g->v, g->frag, g->fragT, g->partial, g->scal};
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

// out = mul * (sum or mean) of the first n floats of x (n = min(*count, cap)
// when count is given), in a fixed order; no allocation.
void sum_floats(lr_guide* g, const float* x, uint32_t n, const uint32_t* count, uint32_t cap, float mul, bool mean,
                float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_slices, dim3(kSumBlocks), dim3(256), 0, s, x, n, count, cap, g->scal + 2);
    hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(kSumBlocks), 0, s, g->scal + 2, n, count, cap, mul, mean ? 1 : 0,
                       out);
}

// k_infer over rows < n (or *n_dev): row-major mixture rows when ld is 0,
// else columns of ld floats.
int infer_rows(lr_guide* g, const float* inputs, uint32_t n, const uint32_t* n_dev, float* mixture, size_t ld,
               hipStream_t s) {
    const uint32_t tiles = (n + kInferRows - 1) / kInferRows;
    const uint32_t grid = std::max(1u, std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves,
                                                          (uint32_t)g->ctx->n_cu * 3));
    if (ld)
        hipLaunchKernelGGL(k_infer<true>, dim3(grid), dim3(kInferBlock), kInferSmem, s, g->frag, g->w + kWtElems,
                           inputs, n, n_dev, mixture, ld);
    else
        hipLaunchKernelGGL(k_infer<false>, dim3(grid), dim3(kInferBlock), kInferSmem, s, g->frag, g->w + kWtElems,
                           inputs, n, n_dev, mixture, (size_t)0);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

}  // namespace

int lr_guide_infer_columns(lr_guide* g, const float* inputs, uint32_t n, const uint32_t* n_dev, float* mixture,
                           size_t ld, hipStream_t s) {
    if (!g || (n && (!inputs || !mixture)) || ld < n) return lr::fail(LR_E_INVALID, "bad lr_guide_infer_columns arguments");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(g->ctx)) return rc;
    return infer_rows(g, inputs, n, n_dev, mixture, ld, s);
}

extern "C" {

int lr_guide_param_count(void) { return kParams; }

int lr_guide_set_radiance_weight(lr_guide* g, float weight) {
    if (!g) return lr::fail(LR_E_INVALID, "NULL guide");
    if (!(weight >= 0.0f && weight < INFINITY)) return lr::fail(LR_E_INVALID, "radiance weight must be >= 0");
    g->rad_weight = weight;
    return LR_OK;
}

int lr_guide_create(lr_context* ctx, uint64_t seed, float lr, lr_guide** out) {
    if (!ctx || !out) return lr::fail(LR_E_INVALID, "NULL context or out");
    *out = nullptr;
    if (!(lr >= 0.0f)) return lr::fail(LR_E_INVALID, "learning rate must be >= 0");
    if (int rc = lr::bind(ctx)) return rc;
    LR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_infer<false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kInferSmem));
    LR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_infer<true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kInferSmem));
    {
        const void* fns[] = {reinterpret_cast<const void*>(k_train<false, 4>),
                             reinterpret_cast<const void*>(k_train<true, 4>),
                             reinterpret_cast<const void*>(k_train<false, 8>),
                             reinterpret_cast<const void*>(k_train<true, 8>)};
        const size_t smem[] = {TrainShape<4>::kSmem, TrainShape<4>::kSmem, TrainShape<8>::kSmem, TrainShape<8>::kSmem};
        for (int k = 0; k < 4; ++k)
            LR_HIP_CHECK(hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem[k]));
    }
    lr_guide* g = new (std::nothrow) lr_guide();
    if (!g) return lr::fail(LR_E_NOMEM, "out of host memory");
    g->ctx = ctx;
    g->lr = lr;
    // a training step runs at most one workgroup per CU: one partial each
    g->max_chunks = (uint32_t)std::max(1, std::min(ctx->n_cu, kMaxChunks));
    hipError_t e = hipMalloc(&g->w, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&g->m, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&g->v, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&g->frag, sizeof(uint16_t) * kWtElems);
    if (e == hipSuccess) e = hipMalloc(&g->fragT, sizeof(uint16_t) * kFragsT * 512);
    if (e == hipSuccess) e = hipMalloc(&g->partial, sizeof(float) * kPartial * (size_t)g->max_chunks);
    if (e == hipSuccess) e = hipMalloc(&g->scal, sizeof(float) * (2 + kSumBlocks));
    if (e == hipSuccess) e = hipMemset(g->scal, 0, sizeof(float) * (2 + kSumBlocks));
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
    return infer_rows(g, inputs, n, n_dev, mixture, 0, static_cast<hipStream_t>(stream));
}

}  // extern "C"

namespace {

// One Adam step on the rows R: k_train (one workgroup per row group, at most
// one per CU) and k_adam.  No allocation: the partials were allocated
// with the guide.
int train_rows(lr_guide* g, const TrainRows& R, bool gather, float* loss_out, float* grad_out, hipStream_t s) {
    const bool wide = R.n > kTrainWideRows;
    const uint32_t rows = wide ? TrainShape<8>::kGroupRows : TrainShape<4>::kGroupRows;
    const uint32_t groups = (R.n + rows - 1) / rows;
    const uint32_t grid = std::max(1u, std::min(groups, g->max_chunks));
    auto launch = [&](auto fn, int waves, size_t smem) {
        hipLaunchKernelGGL(fn, dim3(grid), dim3(64 * waves), smem, s, g->frag, g->fragT, g->w + kWtElems, R, g->partial);
    };
    if (wide)
        gather ? launch(k_train<true, 8>, 8, TrainShape<8>::kSmem) : launch(k_train<false, 8>, 8, TrainShape<8>::kSmem);
    else
        gather ? launch(k_train<true, 4>, 4, TrainShape<4>::kSmem) : launch(k_train<false, 4>, 4, TrainShape<4>::kSmem);
    g->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)g->step), c2 = 1.0f - std::pow(b2, (float)g->step);
    hipLaunchKernelGGL(k_adam, dim3((kParams + kAdamParams - 1) / kAdamParams), dim3(256), 0, s, g->partial, grid,
                       g->w, g->m, g->v,
                       g->frag, g->fragT, g->lr, b1, b2, 1e-8f, c1, c2, grad_out, loss_out);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

}  // namespace

int lr_guide_train_records(lr_guide* g, const float* r_in, const float* r_dir, const float* r_w,
                           const uint32_t* count, uint32_t cap, uint64_t seed, uint32_t chunk, uint32_t step,
                           uint32_t batch, hipStream_t s) {
    if (!g || !r_in || !r_dir || !r_w || !count) return lr::fail(LR_E_INVALID, "NULL argument");
    if (batch == 0) return LR_OK;
    if (int rc = lr::bind(g->ctx)) return rc;
    const TrainRows R{r_in, r_dir, r_w, batch, count, cap, (uint32_t)seed, (uint32_t)(seed >> 32), chunk, step,
                      g->rad_weight > 0.0f ? g->scal : nullptr};
    return train_rows(g, R, true, nullptr, nullptr, s);
}

int lr_guide_prepare_records(lr_guide* g, float* r_w, const uint32_t* count, uint32_t cap, float weight_clamp,
                             hipStream_t s) {
    if (!g || !r_w || !count) return lr::fail(LR_E_INVALID, "NULL argument");
    if (int rc = lr::bind(g->ctx)) return rc;
    if (weight_clamp > 0.0f) {
        sum_floats(g, r_w, 0u, count, cap, weight_clamp, true, g->scal + 1, s);
        hipLaunchKernelGGL(k_clamp_weights, dim3(1024), dim3(256), 0, s, cap, count, g->scal + 1, r_w);
    }
    if (g->rad_weight > 0.0f) sum_floats(g, r_w, 0u, count, cap, g->rad_weight, true, g->scal, s);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

extern "C" {

int lr_guide_train_step(lr_guide* g, const float* inputs, const float* dirs, const float* weights,
                                   uint32_t n, float* loss_out, float* grad_out, void* stream) {
    if (!g || (n && (!inputs || !dirs || !weights))) return lr::fail(LR_E_INVALID, "NULL argument");
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(g->ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (g->rad_weight > 0.0f) sum_floats(g, weights, n, nullptr, 0u, g->rad_weight, true, g->scal, s);
    const TrainRows R{inputs, dirs, weights, n, nullptr, 0u, 0u, 0u, 0u, 0u, g->rad_weight > 0.0f ? g->scal : nullptr};
    return train_rows(g, R, false, loss_out, grad_out, s);
}

}  // extern "C"
