// shading.hip -- batched queries of the path tracer's shading functions on
// device arrays: Bsdf::evaluate / Bsdf::sample in the local BSDF frame
// (src/liar/kernel/bsdf.h:163-164, bsdf.cpp:52-95) and spectrum evaluation
// at given wavelengths (src/liar/spectra/sampled.cpp:102-131,
// src/liar/kernel/spectral.cpp:185-210, black_body.cpp:96-118).  They run the
// very device functions k_shade and k_dl_shade inline (lr_render.h), so a
// caller -- an external sampler, or a test against the CPU oracle -- sees
// what the renderer computes.
#include <hip/hip_runtime.h>

#include <vector>

#include "lr_internal.h"
#include "lr_render.h"

using lr::BsdfParams;
using lr::DevScene;
using lr::DevSpectrum;
using lr::V3;

namespace lr {
int build_spectra(const lr_spectrum* spectra, uint32_t n_spectra, std::vector<DevSpectrum>& dspec,
                  std::vector<float>& table);
}

namespace {

__global__ void k_bsdf_eval(BsdfParams B, const float* __restrict__ wi, const float* __restrict__ wo, uint32_t n,
                            float* __restrict__ value, float* __restrict__ pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = 0.0f, p = 0.0f;
    // shading normal == geometric normal: the hemisphere test is wo.z's sign
    const bool ok = lr::bsdf_evaluate(B, lr::ld3(wi + 3 * (size_t)i), lr::ld3(wo + 3 * (size_t)i),
                                      lr::v3(0.0f, 0.0f, 1.0f), v, p);
    value[i] = ok ? v : 0.0f;
    pdf[i] = ok ? p : 0.0f;
}

__global__ void k_bsdf_sample(BsdfParams B, const float* __restrict__ wi, const float* __restrict__ u, uint32_t n,
                              uint32_t allow, float* __restrict__ wo, float* __restrict__ value,
                              float* __restrict__ pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 w = lr::v3(0.0f, 0.0f, 0.0f);
    float v = 0.0f, p = 0.0f;
    const float* ui = u + 3 * (size_t)i;
    const bool ok = lr::bsdf_sample(B, lr::ld3(wi + 3 * (size_t)i), ui[0], ui[1], ui[2], w, v, p, allow);
    wo[3 * (size_t)i] = w.x;
    wo[3 * (size_t)i + 1] = w.y;
    wo[3 * (size_t)i + 2] = w.z;
    value[i] = ok ? v : 0.0f;
    pdf[i] = ok ? p : 0.0f;
}

__global__ void k_spectrum_eval(DevScene S, const float* __restrict__ w, uint32_t n, bool reflectant,
                                float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = lr::spectrum_eval(S, 0u, w[i], reflectant);
}

int to_params(const lr_bsdf* b, BsdfParams& B) {
    if (!b) return lr::fail(LR_E_INVALID, "NULL bsdf");
    if (b->type == LR_MAT_NONE || b->type > LR_MAT_COOK_TORRANCE_BECKMANN)
        return lr::fail(LR_E_INVALID, "unknown BSDF type " + std::to_string(b->type));
    B.type = b->type;
    B.refl = b->reflectance;
    B.trans = b->transmittance;
    B.eta = fmaxf(b->eta, 1e-9f);
    B.kappa = fmaxf(b->kappa, 1e-9f);
    B.ior = b->ior;
    const float ru = fmaxf(b->roughness_u, 1e-3f), rv = fmaxf(b->roughness_v, 1e-3f);
    B.alpha_u = ru * ru;
    B.alpha_v = rv * rv;
    B.fuzz = b->fuzz;
    if (!(B.fuzz >= 0.0f)) return lr::fail(LR_E_INVALID, "fuzz must be >= 0");
    return LR_OK;
}

}  // namespace

extern "C" {

int lr_bsdf_eval(lr_context* ctx, const lr_bsdf* bsdf, const float* wi, const float* wo, uint32_t n, float* value,
                 float* pdf, void* stream) {
    if (!ctx) return lr::fail(LR_E_INVALID, "NULL context");
    BsdfParams B;
    if (int rc = to_params(bsdf, B)) return rc;
    if (n == 0) return LR_OK;
    if (!wi || !wo || !value || !pdf) return lr::fail(LR_E_INVALID, "NULL array");
    if (int rc = lr::bind(ctx)) return rc;
    hipLaunchKernelGGL(k_bsdf_eval, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), B, wi, wo,
                       n, value, pdf);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

int lr_bsdf_sample(lr_context* ctx, const lr_bsdf* bsdf, const float* wi, const float* u, uint32_t n, uint32_t allow,
                   float* wo, float* value, float* pdf, void* stream) {
    if (!ctx) return lr::fail(LR_E_INVALID, "NULL context");
    BsdfParams B;
    if (int rc = to_params(bsdf, B)) return rc;
    if (n == 0) return LR_OK;
    if (!wi || !u || !wo || !value || !pdf) return lr::fail(LR_E_INVALID, "NULL array");
    if (allow == 0 || allow > 3) return lr::fail(LR_E_INVALID, "allow must be 1 (reflection), 2 or 3");
    if (int rc = lr::bind(ctx)) return rc;
    hipLaunchKernelGGL(k_bsdf_sample, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), B, wi, u,
                       n, allow, wo, value, pdf);
    LR_HIP_CHECK(hipGetLastError());
    return LR_OK;
}

int lr_spectrum_eval(lr_context* ctx, const lr_spectrum* spectrum, const float* wavelengths, uint32_t n,
                     uint32_t reflectant, float* out, void* stream) {
    if (!ctx || !spectrum) return lr::fail(LR_E_INVALID, "NULL context or spectrum");
    if (n && (!wavelengths || !out)) return lr::fail(LR_E_INVALID, "NULL array");
    std::vector<DevSpectrum> dspec;
    std::vector<float> table;
    if (int rc = lr::build_spectra(spectrum, 1u, dspec, table)) return rc;
    if (n == 0) return LR_OK;
    if (int rc = lr::bind(ctx)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevSpectrum* d_spec = nullptr;
    float* d_table = nullptr;
    LR_HIP_CHECK(hipMalloc(&d_spec, sizeof(DevSpectrum)));
    hipError_t e = hipMemcpy(d_spec, dspec.data(), sizeof(DevSpectrum), hipMemcpyHostToDevice);
    if (e == hipSuccess && !table.empty()) e = hipMalloc(&d_table, sizeof(float) * table.size());
    if (e == hipSuccess && d_table)
        e = hipMemcpy(d_table, table.data(), sizeof(float) * table.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        DevScene S{};
        S.spectra = d_spec;
        S.table = d_table;
        hipLaunchKernelGGL(k_spectrum_eval, dim3((n + 255) / 256), dim3(256), 0, s, S, wavelengths, n,
                           reflectant != 0u, out);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    (void)hipFree(d_table);
    (void)hipFree(d_spec);
    if (e != hipSuccess) return lr::fail(LR_E_DEVICE, std::string("lr_spectrum_eval: ") + hipGetErrorString(e));
    return LR_OK;
}

}  // extern "C"
