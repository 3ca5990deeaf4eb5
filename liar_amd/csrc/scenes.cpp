// scenes.cpp -- deterministic synthetic inputs (host, Philox4x32-10 keyed by
// (seed, stream, index)), shared by tests, the bench and the render paths.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/liar_amd.h"
#include "lr_geom.h"
#include "lr_util.h"

namespace {

template <class F>
void parallel_for(uint32_t n, F&& f) {
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t nt = std::min<uint32_t>(hw, std::max<uint32_t>(1u, n / 65536u));
    if (nt <= 1) {
        f(0u, n);
        return;
    }
    std::vector<std::thread> ts;
    const uint32_t chunk = (n + nt - 1) / nt;
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        ts.emplace_back([&f, b, e]() { f(b, e); });
    }
    for (auto& t : ts) t.join();
}

inline lr::u4 draw(uint64_t seed, uint32_t stream, uint32_t index, uint32_t block) {
    return lr::philox4x32((uint32_t)seed, (uint32_t)(seed >> 32), lr::u4{index, block, stream, 0u});
}

}  // namespace

extern "C" {

void lr_philox4x32(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t* out) {
    const lr::u4 r = lr::philox4x32((uint32_t)key, (uint32_t)(key >> 32), lr::u4{c0, c1, c2, c3});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

int lr_gen_random_triangles(uint64_t seed, uint32_t nt, float scale, float* verts, uint32_t* tris) {
    if ((nt && (!verts || !tris)) || !(scale >= 0.0f)) return lr::fail(LR_E_INVALID, "lr_gen_random_triangles: bad arguments");
    parallel_for(nt, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const lr::u4 r0 = draw(seed, 1u, i, 0u);
            const lr::u4 r1 = draw(seed, 1u, i, 1u);
            const lr::u4 r2 = draw(seed, 1u, i, 2u);
            const uint32_t xs[12] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
            const float c[3] = {lr::u01(xs[0]), lr::u01(xs[1]), lr::u01(xs[2])};
            for (int k = 0; k < 3; ++k) {
                for (int a = 0; a < 3; ++a) {
                    const float xi = 2.0f * lr::u01(xs[3 + 3 * k + a]) - 1.0f;
                    verts[9 * (size_t)i + 3 * k + a] = c[a] + scale * xi;
                }
                tris[3 * (size_t)i + k] = 3 * i + k;
            }
        }
    });
    return LR_OK;
}

int lr_gen_random_rays(uint64_t seed, uint32_t n, float tfar_lo, float tfar_hi, float* ox, float* oy, float* oz,
                       float* dx, float* dy, float* dz, float* tnear, float* tfar) {
    if (n && (!ox || !oy || !oz || !dx || !dy || !dz || !tnear || !tfar))
        return lr::fail(LR_E_INVALID, "lr_gen_random_rays: null output");
    const float two_pi = 6.283185307179586f;
    parallel_for(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const lr::u4 r0 = draw(seed, 2u, i, 0u);
            const lr::u4 r1 = draw(seed, 2u, i, 1u);
            ox[i] = lr::u01(r0.x);
            oy[i] = lr::u01(r0.y);
            oz[i] = lr::u01(r0.z);
            const float z = 1.0f - 2.0f * lr::u01(r0.w);
            const float phi = two_pi * lr::u01(r1.x);
            const float s = sqrtf(std::max(0.0f, 1.0f - z * z));
            dx[i] = s * cosf(phi);
            dy[i] = s * sinf(phi);
            dz[i] = z;
            tnear[i] = 0.0f;
            tfar[i] = tfar_hi > 0.0f ? tfar_lo + (tfar_hi - tfar_lo) * lr::u01(r1.y) : INFINITY;
        }
    });
    return LR_OK;
}

// Gradient noise on the integer lattice (Perlin's 2002 quintic fade), unit
// gradients at angles drawn from Philox keyed by (seed, octave, lattice).
static float grad_noise(uint64_t seed, uint32_t octave, float x, float y) {
    const float fx = floorf(x), fy = floorf(y);
    const int32_t ix = (int32_t)fx, iy = (int32_t)fy;
    const float tx = x - fx, ty = y - fy;
    float dots[4];
    for (int c = 0; c < 4; ++c) {
        const int32_t cx = ix + (c & 1), cy = iy + (c >> 1);
        const lr::u4 r = draw(seed, 100u + octave, (uint32_t)cx, (uint32_t)cy);
        const float a = 6.283185307179586f * lr::u01(r.x);
        dots[c] = cosf(a) * (tx - (float)(c & 1)) + sinf(a) * (ty - (float)(c >> 1));
    }
    auto fade = [](float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); };
    const float u = fade(tx), v = fade(ty);
    const float a = dots[0] + u * (dots[1] - dots[0]);
    const float b = dots[2] + u * (dots[3] - dots[2]);
    return a + v * (b - a);
}

int lr_gen_terrain(uint64_t seed, uint32_t n_side, uint32_t octaves, float gain, float base_frequency,
                   float height, float* verts, uint32_t* tris) {
    if (n_side < 2 || !verts || !tris || octaves == 0 || octaves > 30)
        return lr::fail(LR_E_INVALID, "lr_gen_terrain: need n_side >= 2, 1..30 octaves and outputs");
    const uint32_t n = n_side;
    const float inv = 1.0f / (float)(n - 1);
    parallel_for(n * n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t gx = i % n, gy = i / n;
            const float x = gx * inv, y = gy * inv;
            float h = 0.0f, amp = 1.0f, freq = base_frequency;
            for (uint32_t o = 0; o < octaves; ++o) {
                h += amp * grad_noise(seed, o, x * freq, y * freq);
                amp *= gain;
                freq *= 2.0f;
            }
            verts[3 * (size_t)i + 0] = x;
            verts[3 * (size_t)i + 1] = y;
            verts[3 * (size_t)i + 2] = height * h;
        }
    });
    const uint32_t cells = n - 1;
    parallel_for(cells * cells, [&](uint32_t b, uint32_t e) {
        for (uint32_t c = b; c < e; ++c) {
            const uint32_t cx = c % cells, cy = c / cells;
            const uint32_t v00 = cy * n + cx, v10 = v00 + 1, v01 = v00 + n, v11 = v01 + 1;
            uint32_t* t = tris + 6 * (size_t)c;
            t[0] = v00; t[1] = v10; t[2] = v11;
            t[3] = v00; t[4] = v11; t[5] = v01;
        }
    });
    return LR_OK;
}

}  // extern "C"
