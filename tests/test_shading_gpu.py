"""GPU shading functions against the pinned CPU oracle.

1. lr_spectrum_eval: the reference's golden Sampled vector
   (test/test_spectrum.cpp:256-278), BlackBody / Binned / Reflectant values
   against the oracle's lookup.
2. lr_bsdf_eval / lr_bsdf_sample: the device BSDFs the renderer inlines,
   for every BSDF type (anisotropic CookTorrance and fuzzy Mirror included),
   against the oracle's FP64 BSDFs on the same inputs: FP32 error bound
   stated in each assertion.
3. Renders through the C-ABI: a glossy floor under a constant sky equals the
   quadrature albedo (GGX, Beckmann), a Lambert floor lit by the sun disc
   equals rho L sin^2(a) cos(theta_s), and scenes with shading normals,
   fuzzy mirrors and anisotropic microfacets match the oracle render with
   the same seed.
"""
import math
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle as O
from liar_amd import scenery, shaders, shading, spectra, textures

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_oracle_bsdf import albedo, pixel_incidence, tilted_floor  # noqa: E402
from test_render_gpu import compare, gpu_render, say  # noqa: E402
from test_spectrum_golden import VS, WS_M, golden  # noqa: E402

pytestmark = pytest.mark.gpu


def test_spectrum_lookup_matches_golden_and_oracle(gpu_ctx):
    say("[gpu] lr_spectrum_eval: golden Sampled vector, BlackBody, Binned, Reflectant")
    np.testing.assert_allclose(golden().evaluate(gpu_ctx, np.array(WS_M) * 1e9), VS, atol=1e-6)
    w = np.linspace(360.0, 830.0, 941)
    for sp, refl in ((spectra.BlackBody(5800, 3.0), False), (spectra.BlackBody(3000, 1.0), False),
                     (spectra.Binned(380, 780, np.linspace(0.1, 2.0, 10)), True),
                     (spectra.Sampled([380, 450, 600, 780], [0.2, 1.5, 0.7, 3.0]), True),
                     (spectra.Sampled([380, 450, 600, 780], [0.2, 1.5, 0.7, 3.0]), False),
                     (spectra.Constant(0.37), False)):
        got = sp.evaluate(gpu_ctx, w, reflectant=refl)
        want = O.spectrum_eval(sp, w.astype(np.float32).astype(np.float64), reflectant=refl)
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7, err_msg=type(sp).__name__)


def random_hemisphere(rng, n, zmin=0.02):
    v = rng.normal(size=(n, 3))
    v[:, 2] = np.abs(v[:, 2])
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v[v[:, 2] > zmin]


BSDFS = [
    ("lambert", dict(kind=1, refl=0.7)),
    ("mirror", dict(kind=2, refl=0.9)),
    ("mirror_fuzz", dict(kind=2, refl=0.9, fuzz=0.4)),
    ("dielectric_in", dict(kind=3, refl=1.0, trans=1.0, ior=1 / 1.5)),
    ("dielectric_out", dict(kind=3, refl=0.9, trans=0.8, ior=1.5)),
    ("conductor", dict(kind=4, refl=1.0, eta=0.43, kappa=2.45)),
    ("ggx", dict(kind=5, refl=0.8, eta=0.3, kappa=3.0, roughness_u=0.3)),
    ("ggx_aniso", dict(kind=5, refl=0.8, eta=0.3, kappa=3.0, roughness_u=0.15, roughness_v=0.5)),
    ("beckmann", dict(kind=6, refl=0.8, eta=1.2, kappa=2.0, roughness_u=0.25)),
    ("beckmann_aniso", dict(kind=6, refl=0.8, eta=1.2, kappa=2.0, roughness_u=0.5, roughness_v=0.2)),
]


def _pair(ctx, p):
    q = dict(p)
    kind = q.pop("kind")
    rv = q.pop("roughness_v", None)
    args = dict(refl=q.get("refl", 1.0), trans=q.get("trans", 1.0), eta=q.get("eta", 1.0), kappa=q.get("kappa", 0.0),
                ior=q.get("ior", 1.0), roughness_u=q.get("roughness_u", 0.1), roughness_v=rv, fuzz=q.get("fuzz", 0.0))
    gpu = shading.Bsdf(ctx, kind, reflectance=args["refl"], transmittance=args["trans"], eta=args["eta"],
                       kappa=args["kappa"], ior=args["ior"], roughness_u=args["roughness_u"],
                       roughness_v=args["roughness_v"], fuzz=args["fuzz"])
    # the oracle gets the FP32-rounded inputs the GPU sees
    f32 = {k: float(np.float32(v)) if isinstance(v, float) else v for k, v in args.items()}
    prm = O.bsdf_params(kind, f32["refl"], f32["trans"], f32["eta"], f32["kappa"], f32["ior"], f32["roughness_u"],
                        f32["roughness_v"], f32["fuzz"])
    return gpu, prm


@pytest.mark.parametrize("name, p", BSDFS, ids=[b[0] for b in BSDFS])
def test_bsdf_eval_matches_oracle(gpu_ctx, name, p):
    import torch

    rng = np.random.default_rng(11)
    gpu, prm = _pair(gpu_ctx, p)
    wi = random_hemisphere(rng, 20000).astype(np.float32)
    wo = rng.normal(size=(len(wi), 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wo = wo.astype(np.float32)
    val, pdf = gpu.evaluate(torch.from_numpy(wi).cuda(), torch.from_numpy(wo).cuda())
    val, pdf = val.cpu().numpy().astype(np.float64), pdf.cpu().numpy().astype(np.float64)
    ok, ov, op = O.bsdf_eval(prm, wi.astype(np.float64), wo.astype(np.float64))
    # the renderer masks wo below the geometric hemisphere (bsdf_evaluate)
    ok &= wo[:, 2] > 0
    ov, op = np.where(ok, ov, 0), np.where(ok, op, 0)
    # FP32 vs FP64: relative 1e-4 away from grazing configurations
    safe = (np.abs(wo[:, 2]) > 0.02)
    err_v = np.abs(val - ov) / np.maximum(np.abs(ov), 1e-3)
    err_p = np.abs(pdf - op) / np.maximum(np.abs(op), 1e-3)
    say(f"[gpu] bsdf_eval {name}: {int(ok.sum())} valued; max rel err value {err_v[safe].max():.2e} "
        f"pdf {err_p[safe].max():.2e}")
    assert (err_v[safe] > 1e-4).mean() < 1e-3 and (err_p[safe] > 1e-4).mean() < 1e-3


@pytest.mark.parametrize("name, p", BSDFS, ids=[b[0] for b in BSDFS])
def test_bsdf_sample_matches_oracle(gpu_ctx, name, p):
    import torch

    rng = np.random.default_rng(12)
    gpu, prm = _pair(gpu_ctx, p)
    wi = random_hemisphere(rng, 20000, zmin=0.05).astype(np.float32)
    u = rng.random((len(wi), 3)).astype(np.float32)
    wo, val, pdf = gpu.sample(torch.from_numpy(wi).cuda(), torch.from_numpy(u).cuda())
    wo, val, pdf = (x.cpu().numpy().astype(np.float64) for x in (wo, val, pdf))
    ok, owo, ov, op = O.bsdf_sample(prm, wi.astype(np.float64), u.astype(np.float64))
    gok = pdf > 0
    agree = (gok == ok).mean()
    both = gok & ok
    dw = np.abs(wo - owo).max(axis=1)[both]
    ev = (np.abs(val - ov) / np.maximum(np.abs(ov), 1e-3))[both]
    ep = (np.abs(pdf - op) / np.maximum(np.abs(op), 1e-3))[both]
    say(f"[gpu] bsdf_sample {name}: agree {agree:.5f}; max |dwo| {dw.max():.2e}, value {ev.max():.2e}, "
        f"pdf {ep.max():.2e}")
    assert agree > 0.999
    assert (dw > 1e-4).mean() < 2e-3 and (ev > 1e-3).mean() < 2e-3 and (ep > 1e-3).mean() < 2e-3


@pytest.mark.parametrize("kind, mdf", [(5, shaders.MicrofacetTrowbridgeReitz), (6, shaders.MicrofacetBeckmann)])
def test_glossy_floor_furnace_equals_quadrature_albedo(gpu_ctx, kind, mdf):
    theta, rough, eta, kappa = 1.0, 0.4, 0.43, 2.45
    ct = shaders.CookTorrance(textures.Constant(eta), textures.Constant(kappa))
    ct.roughnessU = ct.roughnessV = textures.Constant(rough)
    ct.mdf = mdf()
    floor, cam = tilted_floor(ct, theta)
    cs = la.compile_scene(scenery.List([floor, scenery.LightSky(textures.Constant(1.0))]))
    size = 32
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, size, size, 16384, max_depth=4, seed=5)
    Y = (gx / gw)[1]
    prm = O.bsdf_params(kind, refl=1.0, eta=eta, kappa=kappa, roughness_u=rough)
    cos_i = pixel_incidence(cam, size)
    grid = np.linspace(cos_i.min() - 1e-3, cos_i.max() + 1e-3, 9)
    table = np.array([albedo(prm, math.acos(c), n_t=512, n_p=256) for c in grid])
    expect = float(np.interp(cos_i, grid, table).mean())
    sigma = Y.std() / math.sqrt(Y.size)
    say(f"[gpu] glossy floor kind {kind}: Y {Y.mean():.5f} albedo {expect:.5f} (3 sigma {3 * sigma:.1e})")
    assert abs(Y.mean() - expect) < 1e-3 + 3 * sigma


def test_sun_lit_lambert_floor_closed_form(gpu_ctx):
    rho, a, ts = 0.7, math.radians(0.5) / 2, 0.6
    L = 0.5 / (rho * math.sin(a) ** 2 * math.cos(ts))
    floor, cam = tilted_floor(shaders.Lambert(textures.Constant(rho)), 0.3)
    sun = scenery.LightSun((math.sin(ts), 0.0, math.cos(ts)), 2 * a, textures.Constant(L))
    cs = la.compile_scene(scenery.List([floor, sun]))
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 32, 32, 16384, max_depth=2, seed=6)
    Y = (gx / gw)[1]
    expect = rho * L * math.sin(a) ** 2 * math.cos(ts)
    sigma = Y.std() / math.sqrt(Y.size)
    say(f"[gpu] sun floor: Y {Y.mean():.5f} expect {expect:.5f} (3 sigma {3 * sigma:.1e})")
    assert abs(Y.mean() - expect) < 1e-3 * expect + 3 * sigma


def smooth_sphere_mesh(center, radius, n=24):
    """A UV-sphere TriangleMesh whose corners carry the true sphere normals."""
    th = np.linspace(0, math.pi, n + 1)
    ph = np.linspace(0, 2 * math.pi, 2 * n + 1)[:-1]
    T, P = np.meshgrid(th, ph, indexing="ij")
    nrm = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    verts = np.asarray(center) + radius * nrm
    tris = []
    m = 2 * n
    for i in range(n):
        for j in range(m):
            a, b = i * m + j, i * m + (j + 1) % m
            c, d = (i + 1) * m + j, (i + 1) * m + (j + 1) % m
            if i > 0:
                tris.append(((a, a), (c, c), (b, b)))
            if i < n - 1:
                tris.append(((b, b), (c, c), (d, d)))
    return scenery.TriangleMesh(verts.tolist(), nrm.tolist(), [], tris)


def shading_scene():
    floor = scenery.Parallelogram((-4, -4, 0), (8, 0, 0), (0, 8, 0))
    floor.shader = shaders.Lambert(textures.Constant(0.6))
    objs = [floor]
    m1 = smooth_sphere_mesh((-1.6, 0.0, 0.6), 0.55)
    m1.shader = shaders.Lambert(textures.Constant(0.8))
    fuzzy = shaders.Mirror(textures.Constant(0.9))
    fuzzy.fuzz = textures.Constant(0.25)
    s2 = scenery.Sphere((-0.4, 0.0, 0.55), 0.5)
    s2.shader = fuzzy
    aniso = shaders.CookTorrance(textures.Constant(0.3), textures.Constant(3.0))
    aniso.mdf = shaders.MicrofacetTrowbridgeReitz()
    aniso.roughnessU, aniso.roughnessV = textures.Constant(0.1), textures.Constant(0.45)
    s3 = scenery.Sphere((0.8, 0.0, 0.55), 0.5)
    s3.shader = aniso
    m4 = smooth_sphere_mesh((2.0, 0.0, 0.6), 0.55)
    beck = shaders.CookTorrance(textures.Constant(1.2), textures.Constant(2.0))
    beck.roughnessU, beck.roughnessV = textures.Constant(0.4), textures.Constant(0.15)
    m4.shader = beck
    objs += [m1, s2, s3, m4]
    light = scenery.LightArea(scenery.Parallelogram((-1, -1, 4), (2, 0, 0), (0, -2, 0)))
    light.radiance = textures.Constant(spectra.BlackBody(4500, 8.0))
    scene = scenery.List(objs + [light, scenery.LightSky(textures.Constant(0.3))])
    cam = la.cameras.PerspectiveCamera()
    cam.position = (0.0, -6.0, 2.2)
    cam.lookAt((0.2, 0.0, 0.5))
    cam.fovAngle = math.radians(45)
    cam.aspectRatio = 1.5
    return scene, cam


def test_shading_normals_fuzz_anisotropy_same_seed_matches_oracle(gpu_ctx):
    say("[gpu] shading normals / fuzz / anisotropic CookTorrance vs oracle")
    sc, cam = shading_scene()
    cs = la.compile_scene(sc)
    assert cs.tri_normals is not None and (np.abs(cs.tri_normals).sum(1) > 0).sum() > 2000
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 96, 64, 16, nl=2, max_depth=10, rr_depth=4, seed=9)
    ox, ow = O.render(cs, cam, 96, 64, 16, wavelengths=2, max_depth=10, rr_depth=4, seed=9)
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "shading")
    assert rel_total < 5e-3
    assert frac_bad < 0.05
