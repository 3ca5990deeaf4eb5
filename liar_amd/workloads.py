"""Synthetic workloads of the benchmark configurations, built with the
LiAR-style scene API."""
from __future__ import annotations

import math

import numpy as np

from . import cameras, scenery, shaders, spectra, textures

# 10 bins of 40 nm over [380, 780) nm; a bin takes the value of its centre
# (400, 440, ..., 760 nm).
_CENTRES = 400.0 + 40.0 * np.arange(10)
WHITE = spectra.Binned(380.0, 780.0, np.full(10, 0.75))
RED = spectra.Binned(380.0, 780.0, np.where(_CENTRES > 600.0, 0.6, 0.05))
GREEN = spectra.Binned(380.0, 780.0, np.where((_CENTRES >= 500.0) & (_CENTRES <= 570.0), 0.6, 0.05))


def _quad(o, a, b, shader):
    p = scenery.Parallelogram(o, a, b)
    p.shader = shader
    return p


def _block(corners, heights, shader):
    """Rotated box: 4 floor corners (x, z) extruded between two heights."""
    ps = [(x, y, z) for y in heights for (x, z) in corners]
    faces = [(0, 2, 1), (0, 3, 2), (4, 6, 7), (4, 5, 6), (0, 5, 4), (0, 1, 5),
             (1, 6, 5), (1, 2, 6), (2, 7, 6), (2, 3, 7), (3, 4, 7), (3, 0, 4)]
    m = scenery.TriangleMesh(ps, [], [], faces)
    m.shader = shader
    return m


def cornell_box(light_scale: float = 50.0):
    """Cornell-box-style scene: 5 axis-aligned walls (10 triangles), two rotated
    boxes (24 triangles), one quad light (2 triangles) emitting a 3000 K
    blackbody (peak radiance ``light_scale``).  All surfaces Lambert with the
    10-bin reflectances above.  Returns (scene, camera)."""
    s = 0.01
    W, H, D = 5.56, 5.488, 5.592
    white = shaders.Lambert(textures.Constant(WHITE))
    red = shaders.Lambert(textures.Constant(RED))
    green = shaders.Lambert(textures.Constant(GREEN))
    walls = [
        _quad((0, 0, 0), (W, 0, 0), (0, 0, D), white),      # floor
        _quad((0, H, 0), (0, 0, D), (W, 0, 0), white),      # ceiling
        _quad((0, 0, D), (W, 0, 0), (0, H, 0), white),      # back wall
        _quad((W, 0, 0), (0, 0, D), (0, H, 0), red),        # x = W wall
        _quad((0, 0, 0), (0, H, 0), (0, 0, D), green),      # x = 0 wall
    ]
    short = _block([(1.30, 0.65), (0.82, 2.25), (2.40, 2.72), (2.90, 1.14)], [0.0, 1.65], white)
    tall = _block([(4.23, 2.47), (2.65, 2.96), (3.14, 4.56), (4.72, 4.06)], [0.0, 3.30], white)
    # light quad just below the ceiling, facing down (normal = dx x dy = -y)
    surf = scenery.Parallelogram((2.13, H - 0.001, 2.27), (1.30, 0, 0), (0, 0, 1.05))
    light = scenery.LightArea(surf)
    bb = spectra.BlackBody(3000.0, light_scale)
    light.radiance = textures.Constant(bb)
    light.shader = shaders.Sum([shaders.Unshaded(textures.Constant(bb)), shaders.Lambert(textures.Constant(WHITE))])
    scene = scenery.List([light] + walls + [short, tall])
    cam = cameras.PerspectiveCamera()
    cam.position = (2.78, 2.73, -8.0)
    cam.direction = (0.0, 0.0, 1.0)
    cam.sky = (0.0, 1.0, 0.0)
    cam.fovAngle = math.radians(40.0)
    cam.aspectRatio = 1.0
    del s
    return scene, cam


def furnace_sphere(shader, sky=1.0):
    """White-furnace set-up (test/white_furnace/ashikhmin_shirley.py:20-79):
    unit sphere under a constant sky, camera at (0, -10, 5) looking at the
    origin with an 11 degree field of view."""
    sph = scenery.Sphere((0.0, 0.0, 0.0), 1.0)
    sph.shader = shader
    scene = scenery.List([sph, scenery.LightSky(textures.Constant(sky))])
    cam = cameras.PerspectiveCamera()
    cam.position = (0.0, -10.0, 5.0)
    cam.lookAt((0.0, 0.0, 0.0))
    cam.sky = (0.0, 0.0, 1.0)
    cam.fovAngle = math.radians(11.0)
    cam.aspectRatio = 1.0
    return scene, cam


def lit_floor(shader, sky=None, area=None, cam_height=1.0, fov_deg=4.0, half=50.0):
    """A square floor (z = 0, side 2 half) seen straight down from
    (0, 0, cam_height), lit by a constant sky (``sky`` = radiance) and/or a
    parallelogram light ``area`` = dict(origin, dx, dy, radiance, samples),
    whose front face (cross(dx, dy)) should face the floor.  The camera's
    footprint is the square |x|, |y| <= cam_height tan(fov / 2)."""
    parts = [_quad((-half, -half, 0.0), (2 * half, 0.0, 0.0), (0.0, 2 * half, 0.0), shader)]
    if sky is not None:
        parts.append(scenery.LightSky(textures.Constant(sky)))
    if area is not None:
        surf = scenery.Parallelogram(area["origin"], area["dx"], area["dy"])
        surf.shader = shaders.Unshaded(textures.Constant(area["radiance"]))
        light = scenery.LightArea(surf)
        light.radiance = textures.Constant(area["radiance"])
        if "samples" in area:
            light.numberOfEmissionSamples = int(area["samples"])
        parts.append(light)
    cam = cameras.PerspectiveCamera()
    cam.position = (0.0, 0.0, float(cam_height))
    cam.lookAt((0.0, 0.0, 0.0))
    cam.sky = (0.0, 1.0, 0.0)
    cam.fovAngle = math.radians(fov_deg)
    cam.aspectRatio = 1.0
    return scenery.List(parts), cam


# Gold (n, k): Hagemann, Gudat & Kunz 1974 (liar_amd/data/Au.spd).
def gold_nk():
    rows = [l.split() for l in (__import__("pathlib").Path(__file__).resolve().parent / "data" / "Au.spd")
            .read_text().splitlines() if l.strip() and not l.startswith("#")]
    a = np.array([[float(x) for x in r] for r in rows])
    return spectra.Sampled(a[:, 0], a[:, 1]), spectra.Sampled(a[:, 0], a[:, 2])


def terrain(n_side: int = 1581, n_spheres: int = 10000, seed: int = 7, tile: int = 16,
            sun_scale: float = 5e4, sky: float = 0.2):
    """Heightfield terrain over [0,1]^2 (fBm, 8 octaves, gain 0.5) with
    `n_spheres` spheres of radius U[0.002, 0.02] resting on it.  Shaders are
    split 50/25/25 over Lambert / Cook-Torrance GGX gold (roughness
    U[0.05, 0.5] in 256 levels, 0.0018 apart) / Dielectric 1.5, per 16x16-cell terrain tile and
    per sphere.  Lights: a 0.5-degree sun disc (5800 K) and a constant sky.
    Returns (CompiledScene, camera) -- built directly in the engine's
    flattened form."""
    from . import _native as N
    from .engine import CompiledScene

    n = int(n_side)
    verts = np.empty((n * n, 3), np.float32)
    tris = np.empty((2 * (n - 1) ** 2, 3), np.uint32)
    N.check(N.lib().lr_gen_terrain(seed, n, 8, 0.5, 4.0, 0.1, N.ptr(verts), N.ptr(tris)), "lr_gen_terrain")
    rng = np.random.default_rng(seed)
    # materials: 0 Lambert, 1..256 Cook-Torrance GGX gold (roughness levels), 257 Dielectric
    n_au, k_au = gold_nk()
    table = [spectra.Constant(0.45), n_au, k_au, spectra.Constant(1.5), spectra.Constant(1.0),
             spectra.BlackBody(5800.0, sun_scale), spectra.Constant(sky)]
    NONE = N.LR_NONE
    mats = [{"type": N.MAT_LAMBERT, "reflectance": 0, "transmittance": NONE, "eta": NONE, "kappa": NONE,
             "emission": NONE, "light": NONE, "roughness": 0.0}]
    n_lv = 256
    levels = 0.05 + 0.45 * (np.arange(n_lv) + 0.5) / n_lv  # midpoints of 256 equal bins of [0.05, 0.5]
    for r in levels:
        mats.append({"type": N.MAT_COOK_TORRANCE, "reflectance": 4, "transmittance": NONE, "eta": 1, "kappa": 2,
                     "emission": NONE, "light": NONE, "roughness": float(r)})
    mats.append({"type": N.MAT_DIELECTRIC, "reflectance": 4, "transmittance": 4, "eta": 3, "kappa": 4,
                 "emission": NONE, "light": NONE, "roughness": 0.0})

    def pick(k):
        u = rng.random(k)
        m = np.where(u < 0.5, 0, np.where(u < 0.75, 1 + rng.integers(0, n_lv, k), n_lv + 1))
        return m.astype(np.uint32)

    cells = n - 1
    tiles = (cells + tile - 1) // tile
    tile_mat = pick(tiles * tiles).reshape(tiles, tiles)
    c = np.arange(cells * cells)
    cx, cy = c % cells, c // cells
    cell_mat = tile_mat[cy // tile, cx // tile]
    tri_mat = np.repeat(cell_mat, 2)
    # spheres resting on the surface (height at the centre's (x, y) on its triangle)
    r = rng.uniform(0.002, 0.02, n_spheres)
    xy = rng.uniform(0.03, 0.97, (n_spheres, 2))
    gx, gy = xy[:, 0] * cells, xy[:, 1] * cells
    ix, iy = np.minimum(gx.astype(int), cells - 1), np.minimum(gy.astype(int), cells - 1)
    fx, fy = gx - ix, gy - iy
    z = verts[:, 2].reshape(n, n)
    z00, z10, z01, z11 = z[iy, ix], z[iy, ix + 1], z[iy + 1, ix], z[iy + 1, ix + 1]
    # cell split along v00-v11: lower triangle (v00, v10, v11) when fx >= fy
    h = np.where(fx >= fy, z00 + fx * (z10 - z00) + fy * (z11 - z10), z00 + fy * (z01 - z00) + fx * (z11 - z01))
    spheres = np.stack([xy[:, 0], xy[:, 1], h + r, r], 1).astype(np.float32)
    sph_mat = pick(n_spheres)
    lights = [{"type": N.LIGHT_SUN, "radiance": 5, "direction": (0.35, 0.45, 0.82), "angle": float(np.radians(0.5))},
              {"type": N.LIGHT_SKY, "radiance": 6}]
    cs = CompiledScene(verts, tris, spheres, np.concatenate([tri_mat, sph_mat]).astype(np.uint32), mats, table,
                       lights)
    cam = cameras.PerspectiveCamera()
    cam.position = (0.5, -0.25, 0.32)
    cam.lookAt((0.5, 0.55, 0.0))
    cam.sky = (0.0, 0.0, 1.0)
    cam.fovAngle = math.radians(55.0)
    cam.aspectRatio = 16.0 / 9.0
    return cs, cam


def slit_room(n_side: int = 1581, n_spheres: int = 10000, seed: int = 7, slit_fraction: float = 0.02,
              sun_scale: float = 5e4, sky: float = 0.2):
    """The terrain workload enclosed in a closed box room (Lambert 0.6 walls,
    floor below the terrain, ceiling at z = 0.45) whose ceiling has one slit
    covering `slit_fraction` of its area; the sun (0.5 degree, 5800 K) and the
    sky outside light the room only through the slit.  Camera inside the room.
    Returns (CompiledScene, camera)."""
    from . import _native as N
    from .engine import CompiledScene

    cs, _ = terrain(n_side, n_spheres, seed, sun_scale=sun_scale, sky=sky)
    x0, x1, y0, y1, z0, z1 = -0.02, 1.02, -0.02, 1.02, -0.15, 0.45
    wall_mat = len(cs.materials)
    cs.spectra.append(spectra.Constant(0.6))
    NONE = N.LR_NONE
    cs.materials.append({"type": N.MAT_LAMBERT, "reflectance": len(cs.spectra) - 1, "transmittance": NONE,
                         "eta": NONE, "kappa": NONE, "emission": NONE, "light": NONE, "roughness": 0.0})
    quads = []

    def quad(o, a, b):
        o, a, b = (np.asarray(x, np.float32) for x in (o, a, b))
        quads.append(np.stack([o, o + a, o + a + b, o + b]))

    W, D, H = x1 - x0, y1 - y0, z1 - z0
    quad((x0, y0, z0), (W, 0, 0), (0, D, 0))          # floor
    quad((x0, y0, z0), (0, 0, H), (W, 0, 0))          # y = y0 wall
    quad((x0, y1, z0), (W, 0, 0), (0, 0, H))          # y = y1 wall
    quad((x0, y0, z0), (0, D, 0), (0, 0, H))          # x = x0 wall
    quad((x1, y0, z0), (0, 0, H), (0, D, 0))          # x = x1 wall
    # ceiling with a slit along y at x in [xs, xs + w]
    w = slit_fraction * W
    xs = 0.5 * (x0 + x1) - 0.5 * w + 0.15
    quad((x0, y0, z1), (xs - x0, 0, 0), (0, D, 0))
    quad((xs + w, y0, z1), (x1 - xs - w, 0, 0), (0, D, 0))
    v = np.concatenate(quads).astype(np.float32)
    base = len(cs.verts)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    tris = np.concatenate([t + base + 4 * k for k in range(len(quads))])
    nt_old = len(cs.tris)
    cs.verts = np.concatenate([cs.verts, v])
    cs.tris = np.concatenate([cs.tris, tris])
    pm = cs.prim_material
    cs.prim_material = np.concatenate([pm[:nt_old], np.full(len(tris), wall_mat, np.uint32), pm[nt_old:]])
    # the sun shines down through the slit onto a stripe of terrain near
    # x = 0.76; the camera looks at the x < 0.45 half of the room, so every
    # visible point is lit only indirectly (stripe, walls, ceiling)
    cs.lights[0]["direction"] = (-0.25, 0.1, 0.963)
    cam = cameras.PerspectiveCamera()
    cam.position = (0.42, 0.02, 0.3)
    cam.lookAt((0.15, 0.75, -0.02))
    cam.sky = (0.0, 0.0, 1.0)
    cam.fovAngle = math.radians(50.0)
    cam.aspectRatio = 16.0 / 9.0
    return cs, cam


def random_cube(n_tris: int, seed: int = 2, scale: float = 0.01, sun_scale: float = 5e4, sky: float = 0.2):
    """The intersection microbenchmark's geometry as a rendered scene: n_tris
    triangles in the unit cube (each a random offset of `scale` from a uniform
    centre), split 70/30 over Lambert 0.6 and Cook-Torrance GGX gold
    (roughness 0.3), under the terrain workload's sun and sky.  Camera outside
    the cube.  Returns (CompiledScene, camera)."""
    from . import _native as N
    from .engine import CompiledScene
    from .qbvh import random_triangles

    v, t = random_triangles(int(n_tris), seed=seed, scale=scale)
    rng = np.random.default_rng(seed)
    n_au, k_au = gold_nk()
    table = [spectra.Constant(0.6), n_au, k_au, spectra.Constant(1.0), spectra.BlackBody(5800.0, sun_scale),
             spectra.Constant(sky)]
    NONE = N.LR_NONE
    mats = [{"type": N.MAT_LAMBERT, "reflectance": 0, "transmittance": NONE, "eta": NONE, "kappa": NONE,
             "emission": NONE, "light": NONE, "roughness": 0.0},
            {"type": N.MAT_COOK_TORRANCE, "reflectance": 3, "transmittance": NONE, "eta": 1, "kappa": 2,
             "emission": NONE, "light": NONE, "roughness": 0.3}]
    pm = np.where(rng.random(len(t)) < 0.7, 0, 1).astype(np.uint32)
    lights = [{"type": N.LIGHT_SUN, "radiance": 4, "direction": (0.35, 0.45, 0.82), "angle": float(np.radians(0.5))},
              {"type": N.LIGHT_SKY, "radiance": 5}]
    cs = CompiledScene(v.reshape(-1, 3), t.reshape(-1, 3), np.zeros((0, 4), np.float32), pm, mats, table, lights)
    cam = cameras.PerspectiveCamera()
    cam.position = (0.5, -1.1, 1.1)
    cam.lookAt((0.5, 0.5, 0.45))
    cam.sky = (0.0, 0.0, 1.0)
    cam.fovAngle = math.radians(45.0)
    cam.aspectRatio = 1.0
    return cs, cam


def batched_scene_specs(n_scenes: int = 64, seed: int = 11, min_tris: int = 10_000, max_tris: int = 2_000_000,
                        sizes=(128, 256, 512, 1024), spps=(4, 16, 64), depth=(4, 16)):
    """The draws of the batched uneven workload, without building a scene:
    per scene (index, generator, triangle count asked for, image side, spp,
    max depth, render seed).  Scene k uses the random-cube generator for even
    k and the terrain generator for odd k, with generator seed seed * 1000 + k."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        nt = int(round(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris)))))
        side = int(rng.choice(sizes))
        out.append({"index": k, "generator": "random_cube" if k % 2 == 0 else "terrain", "tris_asked": nt,
                    "gen_seed": seed * 1000 + k, "width": side, "height": side, "spp": int(rng.choice(spps)),
                    "max_depth": int(rng.integers(depth[0], depth[1] + 1)), "seed": 100 + k})
    return out


def batched_scenes(n_scenes: int = 64, seed: int = 11, min_tris: int = 10_000, max_tris: int = 2_000_000,
                   sizes=(128, 256, 512, 1024), spps=(4, 16, 64), depth=(4, 16), only=None):
    """The batched uneven workload: `n_scenes` independent scenes with triangle
    counts log-uniform in [min_tris, max_tris] (random-cube or terrain
    generator, alternately), square images with sides drawn from `sizes`, spp
    from `spps` and max depth uniform in `depth`.  Returns a list of dicts
    (scene, camera, width, height, spp, max_depth, seed, n_tris, index);
    `only`: the scene indices to build (the others are skipped; the draws of
    every scene stay those of the full workload)."""
    out = []
    for sp in batched_scene_specs(n_scenes, seed, min_tris, max_tris, sizes, spps, depth):
        k, nt = sp["index"], sp["tris_asked"]
        if only is not None and k not in only:
            continue
        if sp["generator"] == "random_cube":
            cs, cam = random_cube(nt, seed=sp["gen_seed"])
        else:
            n_side = max(2, int(round(np.sqrt(nt / 2.0))) + 1)
            cs, cam = terrain(n_side, n_spheres=min(10000, nt // 200), seed=sp["gen_seed"])
        cam.aspectRatio = 1.0
        out.append({"scene": cs, "camera": cam, "width": sp["width"], "height": sp["height"], "spp": sp["spp"],
                    "max_depth": sp["max_depth"], "seed": sp["seed"], "n_tris": int(len(cs.tris)), "index": k})
    return out
