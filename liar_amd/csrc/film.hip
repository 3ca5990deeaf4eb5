// film.hip -- film output: RGB spaces and Raster tone mapping on the GPU.
//
// LiAR turns the accumulated XYZ raster into display RGBA in
// Raster::tonemap (src/liar/output/raster.cpp:306-494): per pixel the
// weighted mean XYZ times the exposure gain, one of nine tone-mapping
// operators, the destination RgbSpace's XYZ -> linear RGB matrix
// (src/liar/kernel/rgb_space.cpp:131-137, 335-386, including the Bradford
// adaptation from the space's white point to CIE E) and its gamma curve.
// Automatic exposure (raster.cpp:99-150) uses the log-average luminance of
// the covered pixels.
//
// Here the whole film is one launch sequence on the caller's stream:
//   k_film_stats   per-block partial sums of log Y and counts, and max Y
//   k_exposure     one block: fixed-order sum of the partials, the gain
//   k_tonemap      one thread per pixel, float RGBA and/or 8-bit RGBA
// No host synchronisation unless the caller asks for the exposure values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "lr_internal.h"

namespace {

constexpr int kStatsBlock = 256;
constexpr uint32_t kMaxStatsBlocks = 2048;

struct StatsPartial {
    double sum_log;
    uint32_t count;
    float max_y;
};

struct ToneState {
    float gain;       // 2^(stops + correction)
    float stops;      // exposure stops used
    float avg_y;      // log-average luminance of the covered pixels
    float max_y;      // largest weighted Y
    float inv_lw2;    // 1 / (gain * max_y)^2 (Reinhard 2002)
    float pad[3];
};

struct ToneArgs {
    float m[9];          // XYZ -> linear RGB, row-major
    float inv_gamma;     // 1 / gamma (1: no curve)
    uint32_t op;         // LR_TONEMAP_*
    uint32_t auto_exposure;
    float exposure_stops, correction_stops, middle_grey;
};

// x^g as exp2(g log2 x) on the hardware v_log_f32 / v_exp_f32 (relative
// error ~1e-6 for the exponents and ranges of display curves).  HIP's powf
// and __powf are both the accurate library pow, whose range reduction made
// the per-pixel kernel compute bound.
__device__ inline float fast_pow(float x, float g) {
    return x > 0.0f ? __builtin_amdgcn_exp2f(g * __builtin_amdgcn_logf(x)) : (g > 0.0f ? 0.0f : 1.0f / x);
}

__device__ inline float filmic(float x) {  // raster.cpp:62-67
    x = fmaxf(0.0f, x - 0.004f);
    const float y = (x * (6.2f * x + 0.5f)) / (x * (6.2f * x + 1.7f) + 0.06f);
    return fast_pow(y, 2.2f);
}

__device__ inline float inv_filmic(float y) {  // raster.cpp:69-80
    y = fmaxf(0.0f, fminf(y, 0.99999f));
    y = powf(y, 1.0f / 2.2f);
    const float a = 6.2f * y - 6.2f;
    const float b = 1.7f * y - 0.5f;
    const float c = 0.06f * y;
    const float D = b * b - 4.0f * a * c;
    const float x = (-b - sqrtf(D)) / (2.0f * a);
    return x + 0.004f;
}

__device__ inline float tonecurve(float x, float g) {  // rgb_space.cpp:156-162
    return x < 0.0f ? -fast_pow(-x, g) : fast_pow(x, g);
}

__global__ __launch_bounds__(kStatsBlock) void k_film_stats(const float* __restrict__ xyz,
                                                           const float* __restrict__ w, uint32_t n,
                                                           StatsPartial* __restrict__ part) {
    __shared__ double s_sum[kStatsBlock];
    __shared__ uint32_t s_cnt[kStatsBlock];
    __shared__ float s_max[kStatsBlock];
    double sum = 0.0;
    uint32_t cnt = 0;
    float mx = 0.0f;
    const float* Y = xyz + n;
    for (uint32_t k = blockIdx.x * kStatsBlock + threadIdx.x; k < n; k += gridDim.x * kStatsBlock) {
        const float wk = w[k], yk = Y[k];
        if (wk > 0.0f && yk > 0.0f) {
            const float l = yk / wk;
            // natural log from the hardware log2 (1 ulp), summed in FP64
            sum += (double)__builtin_amdgcn_logf(l) * 0.6931471805599453;
            ++cnt;
            mx = fmaxf(mx, l);
        }
    }
    s_sum[threadIdx.x] = sum;
    s_cnt[threadIdx.x] = cnt;
    s_max[threadIdx.x] = mx;
    __syncthreads();
    for (int h = kStatsBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
            s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = StatsPartial{s_sum[0], s_cnt[0], s_max[0]};
}

// Raster::exposureStops / sceneGain (raster.cpp:99-150, 613-617).  One
// block: each thread sums the partials b = tid, tid + 256, ... in order, then
// a fixed-shape tree, so the result does not depend on scheduling.
__global__ __launch_bounds__(kStatsBlock) void k_exposure(const StatsPartial* __restrict__ part, uint32_t nb,
                                                         ToneArgs A, ToneState* out) {
    __shared__ double s_sum[kStatsBlock];
    __shared__ uint32_t s_cnt[kStatsBlock];
    __shared__ float s_max[kStatsBlock];
    {
        double ps = 0.0;
        uint32_t pc = 0;
        float pm = 0.0f;
        for (uint32_t b = threadIdx.x; b < nb; b += kStatsBlock) {
            const StatsPartial q = part[b];
            ps += q.sum_log;
            pc += q.count;
            pm = fmaxf(pm, q.max_y);
        }
        s_sum[threadIdx.x] = ps;
        s_cnt[threadIdx.x] = pc;
        s_max[threadIdx.x] = pm;
    }
    __syncthreads();
    for (int h = kStatsBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
            s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double sum = s_sum[0];
    const uint32_t cnt = s_cnt[0];
    const float mx = s_max[0];
    const float avg = cnt ? (float)exp(sum / (double)cnt) : 0.0f;
    float stops = A.exposure_stops;
    if (A.auto_exposure) {
        float auto_gain = 1.0f;
        const float g = A.middle_grey;
        if (avg > 0.0f) {
            switch (A.op) {
            case LR_TONEMAP_LINEAR:
                auto_gain = g / avg;
                break;
            case LR_TONEMAP_COMPRESS_Y:
            case LR_TONEMAP_COMPRESS_RGB:
            case LR_TONEMAP_REINHARD2002_Y:
            case LR_TONEMAP_REINHARD2002_RGB:
                auto_gain = g / (avg * (1.0f - g));
                break;
            case LR_TONEMAP_EXPONENTIAL_Y:
            case LR_TONEMAP_EXPONENTIAL_RGB:
                auto_gain = -log1pf(-g) / avg;
                break;
            default:  // Duiker
                auto_gain = inv_filmic(g) / avg;
                break;
            }
        }
        const float s = log2f(auto_gain);
        stops = roundf(3.0f * s) / 3.0f;  // one third of a stop
    }
    const float gain = powf(2.0f, stops + A.correction_stops);
    const float lw = gain * mx;
    out->gain = gain;
    out->stops = stops;
    out->avg_y = avg;
    out->max_y = mx;
    out->inv_lw2 = 1.0f / (lw * lw);
}

// Raster::tonemap per pixel (raster.cpp:320-488).
__global__ void k_tonemap(const float* __restrict__ xyz, const float* __restrict__ wgt,
                          const float* __restrict__ alpha, uint32_t n, ToneArgs A, const ToneState* __restrict__ st,
                          float4* __restrict__ rgba, uchar4* __restrict__ rgba8) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gain = st->gain;
    const float w = wgt[k];
    float X = 0.0f, Y = 0.0f, Z = 0.0f, a = 0.0f;
    if (w > 0.0f) {  // Raster::weighted (raster.cpp:645-651)
        const float s = gain / w;
        X = xyz[k] * s;
        Y = xyz[n + k] * s;
        Z = xyz[2 * (size_t)n + k] * s;
        a = alpha ? alpha[k] / w : 1.0f;
    }
    const bool y_op = A.op == LR_TONEMAP_LINEAR || A.op == LR_TONEMAP_COMPRESS_Y || A.op == LR_TONEMAP_REINHARD2002_Y ||
                      A.op == LR_TONEMAP_EXPONENTIAL_Y || A.op == LR_TONEMAP_DUIKER_Y;
    float r, g, b;
    bool nan;
    if (y_op) {
        float scale = 1.0f;
        switch (A.op) {
        case LR_TONEMAP_COMPRESS_Y:
            scale = 1.0f / (1.0f + Y);
            break;
        case LR_TONEMAP_REINHARD2002_Y:
            scale = (1.0f + Y * st->inv_lw2) / (1.0f + Y);
            break;
        case LR_TONEMAP_EXPONENTIAL_Y:
            scale = Y < 1e-8f ? 1.0f - Y / 2.0f : -expm1f(-Y) / Y;
            break;
        case LR_TONEMAP_DUIKER_Y:
            scale = filmic(Y) / fmaxf(Y, 0.004f);
            break;
        default:
            break;
        }
        X *= scale;
        Y *= scale;
        Z *= scale;
        nan = isnan(X) || isnan(Y) || isnan(Z) || isnan(a);
        r = A.m[0] * X + A.m[1] * Y + A.m[2] * Z;
        g = A.m[3] * X + A.m[4] * Y + A.m[5] * Z;
        b = A.m[6] * X + A.m[7] * Y + A.m[8] * Z;
    } else {
        r = A.m[0] * X + A.m[1] * Y + A.m[2] * Z;
        g = A.m[3] * X + A.m[4] * Y + A.m[5] * Z;
        b = A.m[6] * X + A.m[7] * Y + A.m[8] * Z;
        switch (A.op) {
        case LR_TONEMAP_COMPRESS_RGB:
            r = r / (1.0f + r);
            g = g / (1.0f + g);
            b = b / (1.0f + b);
            break;
        case LR_TONEMAP_REINHARD2002_RGB: {
            const float q = st->inv_lw2;
            r = r * (1.0f + r * q) / (1.0f + r);
            g = g * (1.0f + g * q) / (1.0f + g);
            b = b * (1.0f + b * q) / (1.0f + b);
            break;
        }
        case LR_TONEMAP_EXPONENTIAL_RGB:
            r = -expm1f(-r);
            g = -expm1f(-g);
            b = -expm1f(-b);
            break;
        default:  // Duiker RGB
            r = filmic(r);
            g = filmic(g);
            b = filmic(b);
            break;
        }
        nan = isnan(r) || isnan(g) || isnan(b) || isnan(a);
    }
    float4 out;
    if (nan) {
        out = make_float4(1.0f, 0.0f, 1.0f, 1.0f);  // pink
    } else if (A.inv_gamma != 1.0f) {
        out = make_float4(tonecurve(r, A.inv_gamma), tonecurve(g, A.inv_gamma), tonecurve(b, A.inv_gamma), a);
    } else {
        out = make_float4(r, g, b, a);
    }
    if (rgba) rgba[k] = out;
    if (rgba8) {  // the PNG writer's quantisation (codecs/lodepng.cpp:255-258)
        auto q = [](float v) -> unsigned char {
            const float c = fminf(fmaxf(255.0f * v, 0.0f), 255.0f);
            return (unsigned char)c;
        };
        rgba8[k] = make_uchar4(q(out.x), q(out.y), q(out.z), q(out.w));
    }
}

// 3x3 helpers in double (host).
void mat_mul(const double* a, const double* b, double* c) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    std::memcpy(c, t, sizeof(t));
}

bool mat_inv(const double* m, double* o) {
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    o[0] = c00 * id;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = c01 * id;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = c02 * id;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return true;
}

}  // namespace

extern "C" {

int lr_rgb_space_matrices(const lr_rgb_space* sp, float* rgb_to_xyz, float* xyz_to_rgb) {
    if (!sp) return lr::fail(LR_E_INVALID, "NULL rgb space");
    const float* prim[4] = {sp->red, sp->green, sp->blue, sp->white};
    const char* names[4] = {"red", "green", "blue", "white"};
    for (int k = 0; k < 4; ++k) {  // RgbSpace::enforceChromaticity (rgb_space.cpp:389-396)
        const double x = prim[k][0], y = prim[k][1];
        if (!(x >= 0 && y >= 0 && x + y <= 1.001))
            return lr::fail(LR_E_INVALID, std::string("invalid chromaticity for ") + names[k]);
    }
    if (sp->white[1] == 0.0f) return lr::fail(LR_E_INVALID, "y component of white point should not be zero");
    if (!(sp->gamma > 0.0f)) return lr::fail(LR_E_INVALID, "gamma must be strictly positive");
    // RGB -> XYZ under the space's white (rgb_space.cpp:335-367): columns are
    // the primaries' xyz chromaticities scaled so that RGB (1,1,1) -> W
    const double xr = sp->red[0], yr = sp->red[1], xg = sp->green[0], yg = sp->green[1];
    const double xb = sp->blue[0], yb = sp->blue[1];
    const double Mxyz[9] = {xr, xg, xb, yr, yg, yb, 1 - xr - yr, 1 - xg - yg, 1 - xb - yb};
    const double W[3] = {sp->white[0] / (double)sp->white[1], 1.0,
                         (1.0 - sp->white[0] - sp->white[1]) / (double)sp->white[1]};
    double Mi[9];
    if (!mat_inv(Mxyz, Mi)) return lr::fail(LR_E_INVALID, "rgb primaries are collinear");
    double S[3];
    for (int i = 0; i < 3; ++i) S[i] = Mi[3 * i] * W[0] + Mi[3 * i + 1] * W[1] + Mi[3 * i + 2] * W[2];
    double MXYZ[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) MXYZ[3 * i + j] = Mxyz[3 * i + j] * S[j];
    // Bradford adaptation W -> E (rgb_space.cpp:424-440)
    const double B[9] = {0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296};
    double Bi[9];
    mat_inv(B, Bi);
    double Cs[3], Cd[3];
    for (int i = 0; i < 3; ++i) {
        Cs[i] = B[3 * i] * W[0] + B[3 * i + 1] * W[1] + B[3 * i + 2] * W[2];
        Cd[i] = B[3 * i] + B[3 * i + 1] + B[3 * i + 2];
    }
    double DB[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) DB[3 * i + j] = (Cd[i] / Cs[i]) * B[3 * i + j];
    double CAT[9], M[9], MI[9];
    mat_mul(Bi, DB, CAT);
    mat_mul(CAT, MXYZ, M);
    if (!mat_inv(M, MI)) return lr::fail(LR_E_INVALID, "singular rgb space");
    for (int k = 0; k < 9; ++k) {
        if (rgb_to_xyz) rgb_to_xyz[k] = (float)M[k];
        if (xyz_to_rgb) xyz_to_rgb[k] = (float)MI[k];
    }
    return LR_OK;
}

int lr_tonemap(lr_context* ctx, const float* film_xyz, const float* film_weight, const float* film_alpha,
               uint32_t width, uint32_t height, const lr_tonemap_params* p, float* rgba, uint8_t* rgba8,
               lr_tonemap_info* info, void* stream) {
    if (!ctx || !p) return lr::fail(LR_E_INVALID, "NULL context or params");
    if (!film_xyz || !film_weight) return lr::fail(LR_E_INVALID, "NULL film buffer");
    if (!rgba && !rgba8) return lr::fail(LR_E_INVALID, "no output buffer");
    if (reinterpret_cast<uintptr_t>(rgba) % 16 || reinterpret_cast<uintptr_t>(rgba8) % 4)
        return lr::fail(LR_E_INVALID, "rgba must be 16-byte and rgba8 4-byte aligned");
    if (p->tone_mapping > LR_TONEMAP_DUIKER_RGB) return lr::fail(LR_E_INVALID, "unknown tone mapping");
    if (!(p->middle_grey > 0.0f && p->middle_grey < 1.0f))
        return lr::fail(LR_E_INVALID, "middle_grey must be in (0, 1)");
    const uint64_t n64 = (uint64_t)width * height;
    if (n64 == 0 || n64 >= (1ull << 31)) return lr::fail(LR_E_INVALID, "film size out of range");
    const uint32_t n = (uint32_t)n64;
    if (int rc = lr::bind(ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ToneArgs A;
    float m_rgb[9];
    if (int rc = lr_rgb_space_matrices(&p->space, nullptr, m_rgb)) return rc;
    std::memcpy(A.m, m_rgb, sizeof(A.m));
    A.inv_gamma = (float)(1.0 / (double)p->space.gamma);
    A.op = p->tone_mapping;
    A.auto_exposure = p->auto_exposure ? 1u : 0u;
    A.exposure_stops = p->exposure_stops;
    A.correction_stops = p->exposure_correction_stops;
    A.middle_grey = p->middle_grey;
    const uint32_t nb = std::min<uint32_t>(kMaxStatsBlocks, (n + kStatsBlock - 1) / kStatsBlock);
    void* scratch = nullptr;
    const size_t part_bytes = (sizeof(StatsPartial) * kMaxStatsBlocks + 255) & ~(size_t)255;
    if (int rc = ctx->scratch(s, part_bytes + sizeof(ToneState), &scratch)) return rc;
    StatsPartial* part = static_cast<StatsPartial*>(scratch);
    ToneState* st = reinterpret_cast<ToneState*>(static_cast<char*>(scratch) + part_bytes);
    hipLaunchKernelGGL(k_film_stats, dim3(nb), dim3(kStatsBlock), 0, s, film_xyz, film_weight, n, part);
    hipLaunchKernelGGL(k_exposure, dim3(1), dim3(kStatsBlock), 0, s, part, nb, A, st);
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, film_xyz, film_weight, film_alpha, n, A,
                       st, reinterpret_cast<float4*>(rgba), reinterpret_cast<uchar4*>(rgba8));
    LR_HIP_CHECK(hipGetLastError());
    if (info) {
        ToneState h;
        LR_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
        LR_HIP_CHECK(hipStreamSynchronize(s));
        info->exposure_stops = h.stops;
        info->gain = h.gain;
        info->average_luminance = h.avg_y;
        info->max_luminance = h.max_y;
    }
    return LR_OK;
}

}  // extern "C"
