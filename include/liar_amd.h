/*
 * liar_amd.h -- C-ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Every entry point returns LR_OK (0) on success or a negative LR_E* code;
 * lr_last_error() then returns a thread-local, human-readable message.
 *
 * Pointer conventions
 *   - "host" pointers are borrowed for the duration of the call only.
 *   - "device" pointers address HBM of the context's device (hipMalloc'd or
 *     torch CUDA tensors); they are borrowed for the duration of the call and
 *     the work is ordered on `stream` (a hipStream_t, NULL = default stream).
 *   - Calls on distinct streams are thread-safe; one lr_qbvh may serve
 *     concurrent read-only traces.
 *
 * Geometry semantics follow LiAR (reference paths are relative to the
 * reference repository root):
 *   - ray directions are unit vectors, so t is a world distance
 *     (src/liar/kernel/kernel_common.h:116);
 *   - a hit t is "in range" iff t > tnear*(1+eps) && t < tfar*(1-eps),
 *     eps = LR_TOLERANCE (src/liar/kernel/bounded_ray.h:69-77,
 *     src/liar/kernel/kernel_common.cpp:29-33);
 *   - closest-hit on the triangle set takes the nearest t > tnear over ALL
 *     triangles and then reports a miss if that t is not in range
 *     (src/liar/scenery/triangle_mesh.cpp:291-317); each sphere takes its
 *     nearest root t > tnear and keeps it iff in range
 *     (src/liar/scenery/sphere.cpp:104-118); the overall result is the
 *     minimum over the triangle set and the spheres
 *     (src/liar/scenery/list.cpp:106-133);
 *   - any-hit: a triangle with tnear < t < tfar, or a sphere whose nearest
 *     root t > tnear is in range (src/liar/scenery/triangle_mesh.cpp:321-330,
 *     src/liar/scenery/sphere.cpp:122-127, src/liar/scenery/list.cpp:137-150);
 *   - equal t is broken by the smaller primitive id (LiAR keeps the
 *     first-found, which depends on traversal order; this library fixes the
 *     rule so that every traversal order gives the same answer).
 *   - primitive ids: triangle k -> k; sphere j -> n_triangles + j
 *     (src/liar/scenery/triangle_mesh.cpp:310-311).
 */
#ifndef LIAR_AMD_H
#define LIAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LR_OK 0
#define LR_E_INVALID (-1)   /* bad argument (null pointer, size, handle)     */
#define LR_E_DEVICE (-2)    /* HIP runtime error                            */
#define LR_E_NOMEM (-3)     /* allocation failed                            */
#define LR_E_UNSUPPORTED (-4)

#define LR_MISS 0xFFFFFFFFu /* prim id of a miss                           */
#define LR_TOLERANCE 1e-6f  /* LiAR's global tolerance (kernel_common.cpp:31) */

/* Flattened 4-wide QBVH node, 128 B (one HBM/L2 line).  SoA over the four
 * children so one lane fetches a whole node with 8 dwordx4 loads.
 * child[i] encoding:
 *   LR_CHILD_EMPTY           no child
 *   bit31 == 0               inner node index
 *   bit31 == 1               leaf: first prim record = (c >> 2) & 0x1FFFFFFF,
 *                            prim count = (c & 3) + 1 (1..4)
 * Bounds are padded outward at build time so FP32 slab tests are
 * conservative. */
#define LR_CHILD_EMPTY 0xFFFFFFFFu
typedef struct lr_node {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t reserved[4];
} lr_node;

/* The device traversal's node, 64 B: lr_node with each child plane stored
 * as 8 bits on a per-node, per-axis grid.  A plane decodes as
 *   fmaf((float)q, scale[axis], origin[axis])
 * (scale a power of two, so q * scale is exact and the decode has one
 * rounding, the same on the host and the GPU).  lo planes are rounded down
 * and hi planes up, so every decoded box contains the lr_node box: box tests
 * stay conservative and the hits of a traversal are unchanged, while a node
 * fetch is half an L2 line.  No counterpart in LiAR, whose TriangleMesh
 * keeps a Lass QBVH with FP32 boxes (src/liar/scenery/triangle_mesh.cpp:69-85);
 * the library derives it from lr_node. */
typedef struct lr_node_q8 {
    float origin[3];  /* per axis: the smallest lo of the node's children */
    float scale[3];   /* per axis: grid step, a power of two */
    uint8_t qlo_x[4], qhi_x[4];
    uint8_t qlo_y[4], qhi_y[4];
    uint8_t qlo_z[4], qhi_z[4];
    uint32_t child[4]; /* as lr_node.child */
} lr_node_q8;

/* Primitive record, 48 B, stored in leaf order.
 *   triangle: a = (v0, id), b = (v1 - v0, kind=0), c = (v2 - v0, 0)
 *   sphere:   a = (centre, id), b = (radius, radius^2, 0, kind=1), c = 0
 * id / kind are stored as uint32 bit patterns in the .w lanes. */
typedef struct lr_prim {
    float a[4];
    float b[4];
    float c[4];
} lr_prim;

#define LR_PRIM_TRIANGLE 0u
#define LR_PRIM_SPHERE 1u

/* Device SoA ray batch, 32 B per ray (all device pointers, length n). */
typedef struct lr_rays_soa {
    const float *ox, *oy, *oz;
    const float *dx, *dy, *dz;
    const float *tnear, *tfar;
} lr_rays_soa;

/* Device SoA hit batch, 16 B per ray. Miss: prim = LR_MISS, t = +inf,
 * u = v = 0.  Triangles: (u, v) barycentrics, point = v0 + u e1 + v e2
 * (src/liar/scenery/triangle.cpp:85-98).  Spheres: u = v = 0. */
typedef struct lr_hits_soa {
    uint32_t *prim;
    float *t, *u, *v;
} lr_hits_soa;

typedef struct lr_qbvh_info {
    uint32_t n_nodes;
    uint32_t n_prims;
    uint32_t n_triangles;
    uint32_t n_spheres;
    uint32_t max_depth;   /* levels below the root                        */
    uint32_t max_stack;   /* traversal stack bound = 3*(max_depth+1) + 1  */
    double sah_cost;      /* SAH cost, node cost 1, prim cost 1           */
    double build_seconds;
    float scene_lo[3], scene_hi[3];
} lr_qbvh_info;

typedef struct lr_context lr_context;
typedef struct lr_qbvh lr_qbvh;

/* Thread-local message of the last failing call on this thread. */
const char* lr_last_error(void);

/* Library version string. */
const char* lr_version(void);

/* Replaces the process-global RenderEngine set-up of LiAR
 * (src/liar/kernel/render_engine.cpp:170-247): binds one HIP device. */
int lr_context_create(int device, lr_context** out);
int lr_context_destroy(lr_context* ctx);

/* Replaces the TriangleMesh ctor + Lass TriangleMesh3D QBVH build and the
 * object-level QbvhTree::doPreProcess over Sphere children
 * (src/liar/scenery/triangle_mesh.cpp:69-85, src/liar/scenery/object_trees.h:100-108).
 *   verts   host, 3*nv floats (x, y, z per vertex)
 *   tris    host, 3*nt uint32 vertex indices
 *   spheres host, 4*ns floats (cx, cy, cz, radius); may be NULL if ns == 0
 * SAH 4-wide build on the host, then the node and prim arrays are uploaded
 * to the context's device.  ctx == NULL builds a host-only handle (no device
 * memory; lr_qbvh_get_info / lr_qbvh_export work, tracing it is an error). */
int lr_qbvh_create(lr_context* ctx, const float* verts, uint32_t nv,
                   const uint32_t* tris, uint32_t nt, const float* spheres,
                   uint32_t ns, lr_qbvh** out);
int lr_qbvh_destroy(lr_qbvh* bvh);
int lr_qbvh_get_info(const lr_qbvh* bvh, lr_qbvh_info* info);
/* Copies the flattened arrays back to host memory (nodes: n_nodes records,
 * prims: n_prims records); either pointer may be NULL. */
int lr_qbvh_export(const lr_qbvh* bvh, lr_node* nodes, lr_prim* prims);
/* Copies the quantised nodes the GPU traverses (n_nodes records, the same
 * indexing and child encoding as lr_qbvh_export's nodes).  No counterpart in
 * LiAR: it exposes the device format so that it can be checked. */
int lr_qbvh_export_q8(const lr_qbvh* bvh, lr_node_q8* nodes);

/* Closest hit for n rays.  Replaces SceneObject::intersect over the scene
 * (src/liar/kernel/scene_object.h:110-113) and its TriangleMesh / Sphere /
 * QbvhTree / List implementations (triangle_mesh.cpp:291-317,
 * sphere.cpp:104-118, object_trees.h:110-135, list.cpp:106-133). */
int lr_trace_closest(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n,
                     lr_hits_soa* hits, void* stream);

/* lr_trace_closest without allocation: no LiAR counterpart.  For callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * lr_trace_closest (hits are bit-identical); the traversal's work counter
 * and stack spill live in a fixed 32 MiB device array per GPU, held until the
 * process ends, instead of the context's grow-only per-stream scratch.  The
 * persistent grid shrinks so that its spill fits; a tree whose stack needs
 * more than the array holds for one workgroup returns LR_E_INVALID.  Every
 * call on one GPU shares that array: calls that overlap on a GPU (other
 * streams, other contexts, other threads) race and must be serialised by
 * the caller. */
int lr_trace_closest_noalloc(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n,
                             lr_hits_soa* hits, void* stream);

/* Any hit (occlusion) for n rays; occluded[i] = 1 or 0.  Replaces
 * SceneObject::isIntersecting (scene_object.h:110-113, triangle_mesh.cpp:321-330,
 * sphere.cpp:122-127, object_trees.h:137-143, list.cpp:137-150). */
int lr_trace_any(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n,
                 uint8_t* occluded, void* stream);

/* ---- instancing: two-level trees ------------------------------------------
 * A top-level tree over instances, each an lr_qbvh (the instanced geometry,
 * built once) placed by a local-to-world affine transform.  Replaces
 * Transformation::doIntersect / doIsIntersecting
 * (src/liar/scenery/transformation.cpp:117-131): the ray goes to the
 * instance's local space, origin W o and direction normalise(W d) with W the
 * inverse transform, and its bounds scale by s = |W d|; the instance is then
 * intersected with the single-level rules above in local space, and its hit
 * distance is t_local / s in world units.  Instances combine as a List
 * (src/liar/scenery/list.cpp:106-150): nearest world t, equal t broken by
 * the smaller instance index, then the smaller primitive id.  The local ray
 * is computed in FP32 as ((W00 ox + W01 oy) + W02 oz) + W03 per row, s =
 * sqrt((x x + y y) + z z), d_local = W d / s. */
typedef struct lr_instance {
    uint32_t geometry;        /* index into the geometry (lr_qbvh) array     */
    float local_to_world[12]; /* row-major 3x4 affine transform              */
} lr_instance;

typedef struct lr_tlas lr_tlas;

typedef struct lr_tlas_info {
    uint32_t n_instances;     /* instances with geometry (empty ones dropped) */
    uint32_t n_nodes;         /* top-level nodes                              */
    uint32_t max_depth;       /* top-level depth                              */
    uint32_t max_stack;       /* two-level traversal stack bound              */
    uint64_t n_geometry_nodes, n_geometry_prims;
    double build_seconds;
    float scene_lo[3], scene_hi[3];
} lr_tlas_info;

/* geometry: n_geometry lr_qbvh of ctx (each may be shared by many
 * instances; they must outlive the tlas only until this call returns: their
 * trees are copied).  The local-to-world transforms must be invertible. */
int lr_tlas_create(lr_context* ctx, lr_qbvh* const* geometry, uint32_t n_geometry,
                   const lr_instance* instances, uint32_t n_instances, lr_tlas** out);
int lr_tlas_destroy(lr_tlas* tlas);
int lr_tlas_get_info(const lr_tlas* tlas, lr_tlas_info* info);

/* Closest hit over the instances: hits->prim is the primitive id within the
 * instance's geometry, hits->t the world distance, u / v as for
 * lr_trace_closest; instance (device, n) receives the instance index
 * (LR_MISS on a miss). */
int lr_trace_closest_instanced(lr_tlas* tlas, const lr_rays_soa* rays, uint32_t n, lr_hits_soa* hits,
                               uint32_t* instance, void* stream);
/* Any hit over the instances (occluded[i] = 1 or 0). */
int lr_trace_any_instanced(lr_tlas* tlas, const lr_rays_soa* rays, uint32_t n, uint8_t* occluded,
                           void* stream);

/* Traversal launch parameters.  0 keeps the current value.
 *   block          threads per workgroup (64, 128 or 256)
 *   blocks_per_cu  cap on resident workgroups per CU (the grid is the
 *                  occupancy limit x #CUs, persistent)
 *   lds_stack      traversal stack entries kept in LDS (8, 16 or 24); deeper
 *                  entries spill to global memory
 *   refill_min     a wave fetches new rays once at least this many of its 64
 *                  lanes are idle (1..64) */
/* schedule: LR_TRACE_SCHEDULE_STEP -- each lane visits one node or one leaf
 * per iteration (the two code paths share the iteration); PHASED -- the
 * wave alternates node and leaf iterations, a lane postponing one leaf and
 * visiting nodes on until every lane has one (fuller waves for short rays,
 * more registers).  In lr_set_trace_config a 0 field keeps its value. */
#define LR_TRACE_SCHEDULE_STEP 1u
#define LR_TRACE_SCHEDULE_PHASED 2u
typedef struct lr_trace_config {
    uint32_t block;
    uint32_t blocks_per_cu;
    uint32_t lds_stack;
    uint32_t refill_min;
    uint32_t schedule;
} lr_trace_config;
int lr_set_trace_config(lr_context* ctx, const lr_trace_config* cfg);
int lr_get_trace_config(const lr_context* ctx, lr_trace_config* cfg);

/* Traversal autotuner (no counterpart in LiAR, whose CPU tracer has no launch
 * parameters).  Times the caller's own batch under every instantiated
 * (block, lds_stack) kernel x blocks_per_cu in {2, 4, 8, 16} x refill_min in
 * {8, 16, 32, 48} x both schedules (288 configurations): one warm-up launch, then `repeats` (0: 3) launches between
 * HIP events on `stream`.  Tunes closest hit when hits is non-NULL, else any
 * hit into occluded.  The fastest configuration becomes the context's
 * (lr_set_trace_config) and is returned in *best; entries (optional, up to
 * max_entries) receive every configuration with its mean milliseconds per
 * launch.  Synchronises the stream; other traces on the same context must not
 * run meanwhile.  Results are those of the last timed launch. */
typedef struct lr_tune_entry {
    lr_trace_config cfg;
    float ms;
} lr_tune_entry;
int lr_autotune_trace(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n, lr_hits_soa* hits,
                      uint8_t* occluded, uint32_t repeats, lr_tune_entry* entries,
                      uint32_t max_entries, uint32_t* n_entries, lr_trace_config* best,
                      void* stream);

/* ---- spectral path tracing (wavefront) ------------------------------------ */

#define LR_NONE 0xFFFFFFFFu

/* Spectra are evaluated at one sampled wavelength per path, as LiAR's
 * SPECTRAL_MODE=Single (src/liar/kernel/spectral.h:64-65).  Wavelengths are
 * in nanometres. */
#define LR_SPEC_CONSTANT 0u  /* value                                         */
#define LR_SPEC_SAMPLED 1u   /* piecewise linear through (w[i], v[i]), 0 outside
                                (src/liar/kernel/spectral.cpp:185-210); as a
                                reflectance divided by max(v) when above 1, or
                                by `value` when value > 0 (a recovered XYZ
                                spectrum's blended maximum,
                                recovery_meng_simon.cpp:169-178)              */
#define LR_SPEC_BINNED 2u    /* `count` equal bins over [lo, hi), constant per
                                bin, 0 outside                                 */
#define LR_SPEC_BLACKBODY 3u /* Planck at `temperature` K, peak = `scale`
                                (src/liar/spectra/black_body.cpp:47-59,96-118) */
typedef struct lr_spectrum {
    uint32_t kind;
    uint32_t count;
    const float* w; /* host, SAMPLED: count strictly increasing wavelengths   */
    const float* v; /* host, SAMPLED / BINNED: count values                   */
    float value;
    float lo, hi;
    float temperature;
    float scale;
} lr_spectrum;

/* Surface shaders (src/liar/shaders/). Spectrum slots index the
 * lr_spectrum array given to lr_scene_create; LR_NONE = unused.  Reflectance
 * slots use LiAR's Reflectant convention (a SAMPLED/BINNED spectrum whose
 * maximum exceeds 1 is divided by that maximum). */
#define LR_MAT_NONE 0u          /* no BSDF (absorbs)                         */
#define LR_MAT_LAMBERT 1u       /* lambert.cpp:76-124                        */
#define LR_MAT_MIRROR 2u        /* mirror.cpp:99-161, with fuzz               */
#define LR_MAT_DIELECTRIC 3u    /* dielectric.cpp:148-247                    */
#define LR_MAT_CONDUCTOR 4u     /* conductor.cpp:125-198                     */
#define LR_MAT_COOK_TORRANCE 5u /* cook_torrance.cpp:189-294 with the
                                   Trowbridge-Reitz (GGX) distribution,
                                   microfacet_trowbridge_reitz.cpp:46-84     */
#define LR_MAT_COOK_TORRANCE_BECKMANN 6u /* same with the Beckmann
                                   distribution (LiAR's default mdf),
                                   microfacet_beckmann.cpp:50-106            */
typedef struct lr_material {
    uint32_t type;
    uint32_t reflectance;   /* Lambert diffuse; Mirror/Dielectric/Conductor/
                               CookTorrance reflectance                      */
    uint32_t transmittance; /* Dielectric                                    */
    uint32_t eta;           /* Conductor/CookTorrance n; Dielectric inner IOR */
    uint32_t kappa;         /* Conductor/CookTorrance k; Dielectric outer IOR */
    uint32_t emission;      /* Unshaded emission seen at generation 0        */
    uint32_t light;         /* index of the area light this surface is       */
    float roughness;        /* CookTorrance roughnessU: alpha_u =
                               max(roughness, 1e-3)^2 along the BSDF frame's
                               u tangent (cook_torrance.cpp:189-200)         */
    float roughness_v;      /* CookTorrance roughnessV (alpha_v likewise)    */
    float fuzz;             /* Mirror fuzz (mirror.cpp:142-161), >= 0; 0 is a
                               perfect mirror                               */
} lr_material;

/* ---- batched shading queries ---------------------------------------------
 * One BSDF, its inputs already looked up at one wavelength, in its local
 * frame (z = shading normal towards the viewer, wi.z > 0): what
 * Shader::bsdf(sample, context) returns for a hit (shader.h:60-87). */
typedef struct lr_bsdf {
    uint32_t type;           /* LR_MAT_LAMBERT .. LR_MAT_COOK_TORRANCE_BECKMANN */
    float reflectance;       /* Lambert diffuse / reflectance (Reflectant)   */
    float transmittance;     /* Dielectric                                   */
    float eta, kappa;        /* Conductor / CookTorrance n, k (>= 1e-9)      */
    float ior;               /* Dielectric n_incident / n_transmitted        */
    float roughness_u, roughness_v; /* CookTorrance                          */
    float fuzz;              /* Mirror                                       */
} lr_bsdf;

/* Bsdf::evaluate(wi, wo, all caps) for n direction pairs (device arrays,
 * 3 floats each) -> value, pdf (0, 0 where the BSDF has no value), with the
 * geometric normal equal to the shading normal (bsdf.h:163, bsdf.cpp:52-76). */
int lr_bsdf_eval(lr_context* ctx, const lr_bsdf* bsdf, const float* wi, const float* wo, uint32_t n,
                 float* value, float* pdf, void* stream);
/* Bsdf::sample(wi, (u0, u1), u2, caps) for n inputs (device; u: 3 floats
 * each) -> wo (3 floats), value, pdf (pdf 0: no sample).  allow: bit 0
 * reflection, bit 1 transmission caps (bsdf.h:164, bsdf.cpp:80-95). */
int lr_bsdf_sample(lr_context* ctx, const lr_bsdf* bsdf, const float* wi, const float* u, uint32_t n,
                   uint32_t allow, float* wo, float* value, float* pdf, void* stream);
/* A spectrum's value at n wavelengths [nm] (device arrays), as the renderer
 * evaluates it for one sampled wavelength: Sampled::doEvaluate /
 * Spectral::fromSampled in Single mode (sampled.cpp:102-131,
 * spectral.cpp:185-210: 0 outside [w0, wn) and at the last wavelength,
 * reflectant != 0 divides by the table maximum when it exceeds 1), BlackBody
 * (black_body.cpp:96-118).  Synchronous. */
int lr_spectrum_eval(lr_context* ctx, const lr_spectrum* spectrum, const float* wavelengths, uint32_t n,
                     uint32_t reflectant, float* out, void* stream);

/* Lights (src/liar/scenery/light_area.cpp, light_sky.cpp). */
#define LR_LIGHT_PARALLELOGRAM 0u /* LightArea over a Parallelogram         */
#define LR_LIGHT_SKY 1u           /* LightSky with a constant radiance      */
#define LR_LIGHT_SUN 2u           /* distant disc of `angle` diameter      */
#define LR_LIGHT_MESH 3u          /* LightArea over a TriangleMesh: the
                                     scene triangles first_triangle ..
                                     first_triangle + n_triangles - 1
                                     (triangle_mesh.cpp:414-444)          */
typedef struct lr_light {
    uint32_t type;
    uint32_t radiance;      /* spectrum index                               */
    uint32_t double_sided;  /* parallelogram: emit on both faces            */
    float origin[3], dx[3], dy[3]; /* parallelogram: origin + u dx + v dy,
                                      front face normal = cross(dx, dy)     */
    float direction[3];     /* sun: unit vector towards the sun             */
    float angle;            /* sun: apparent diameter [rad]                 */
    uint32_t samples;       /* DirectLighting emission samples per camera
                               sample (LightArea / LightSky
                               numberOfEmissionSamples); 0: LiAR's default,
                               9 for a parallelogram or mesh, 1 for sky and
                               sun                                          */
    uint32_t first_triangle, n_triangles; /* mesh: its triangles' primitive
                               ids (at most 4096 triangles)                 */
} lr_light;

/* PerspectiveCamera (src/liar/cameras/perspective_camera.cpp:406-434),
 * pinhole, default film width 0.036. */
typedef struct lr_camera {
    float position[3];
    float direction[3];
    float sky[3];
    float fov;        /* horizontal field of view [rad]                     */
    float aspect;     /* width / height                                     */
    float near_limit; /* rays start at max(1e-6, near_limit)                 */
} lr_camera;

typedef struct lr_render_params {
    uint32_t width, height;
    uint32_t spp;          /* camera samples per pixel                       */
    uint32_t wavelengths;  /* stratified wavelengths (paths) per camera sample */
    uint32_t max_depth;    /* ray generations per path (0: 64)              */
    uint32_t rr_depth;     /* Russian roulette from this generation on       */
    float rr_survival;     /* survival probability (LiAR: 0.5)               */
    float lambda_min, lambda_max; /* wavelength sampling domain [nm]       */
    uint64_t seed;
    uint32_t chunk_paths;  /* paths per wavefront chunk, whole pixels (0: 32M
                              for lr_render and lr_render_guided without
                              training, 4M for lr_render_guided with
                              training and for lr_render_direct)             */
    float max_sample_luminance; /* Raster::maxSampleLuminance: a sample whose
                                   Y exceeds it is scaled down to it
                                   (raster.cpp:567-570); 0 = no clamp       */
    uint32_t graph;        /* nonzero: the call's launches (every chunk and
                              bounce, guide inference and training) are
                              captured into a HIP graph, kept per scene and
                              updated in place on later calls, and launched
                              as one; the caller's stream is ordered before
                              and after it by events                         */
    uint32_t sampler;      /* LR_SAMPLER_STRATIFIER (0) or LR_SAMPLER_HALTON */
    uint32_t sampler_generations; /* Halton: ray generations that draw from
                              the sequence (the tracer's maxRayGeneration;
                              later ones from the counter-based stream)     */
    uint64_t sample_offset; /* Halton: index of the call's first sample, so
                              that successive calls continue the sequence  */
} lr_render_params;

/* Samplers.  LR_SAMPLER_STRATIFIER: spp stratified samples in every pixel,
 * one per path and wavelength stratum (counter-based numbers).
 * LR_SAMPLER_HALTON: LiAR's Halton sampler (src/liar/samplers/halton.cpp:
 * 180-230): width * height * spp * wavelengths samples of one global
 * sequence over the whole film, sample j at index j + 1 + sample_offset:
 * screen = radical inverses in bases 2 and 3, wavelength base 13, and for
 * generation g < sampler_generations the AdjointPhotonTracer's sub-
 * sequences (per generation 1D strategy, component, light choice, scatter
 * and 2D BSDF, light, phase) on the primes from 17 on, 1D ones first, each
 * digit-scrambled by the permutation LiAR's initialize() draws
 * (std::mt19937 default-seeded, std::shuffle, in prime order).  A sample's
 * pixel is where its screen point falls, so the film weights vary per
 * pixel and are added with float atomics (reproducible up to FP32
 * summation order).  Path tracing (lr_render) only. */
#define LR_SAMPLER_STRATIFIER 0u
#define LR_SAMPLER_HALTON 1u

/* The Halton sampler's digit permutation of sub-sequence prime `index` (0:
 * prime 17, 1: 19, ...; at most 284): *prime and, when out has room for
 * prime entries, the permutation. */
int lr_halton_permutation(uint32_t index, uint32_t* prime, uint32_t* out, uint32_t capacity);

typedef struct lr_render_stats {
    uint64_t paths;
    uint64_t rays;         /* extension rays traced (all generations)        */
    uint32_t chunks;
    double seconds;        /* host wall time of lr_render (synchronised)     */
} lr_render_stats;

typedef struct lr_scene lr_scene;

/* Standard observer (LiAR: liar.Observer.setStandard, src/liar/__init__.py:51-69,
 * src/liar/kernel/observer.cpp:57-129): n wavelengths [nm] and x, y, z
 * colour matching values (xyz: 3*n, interleaved).  Process-global. */
int lr_set_standard_observer(const float* w, const float* xyz, uint32_t n);

/* A renderable scene: geometry (an lr_qbvh of the same context), one
 * material index per primitive (triangles then spheres), materials,
 * spectra and lights.  All arrays are host, copied. */
int lr_scene_create(lr_context* ctx, lr_qbvh* geometry, const uint32_t* prim_material,
                    const lr_material* materials, uint32_t n_materials,
                    const lr_spectrum* spectra, uint32_t n_spectra,
                    const lr_light* lights, uint32_t n_lights, lr_scene** out);
int lr_scene_destroy(lr_scene* scene);

/* Shading normals of the scene's triangles: TriangleMesh vertex normals,
 * interpolated at the hit's barycentric (u, v) (triangle_mesh.cpp:334-358,
 * flipped towards the viewer and used for the BSDF frame,
 * intersection_context.cpp:197-223,278-291).  corner_normals holds 9 floats
 * per triangle, (n0, n1, n2) of its corners in primitive-id order, for
 * n_triangles = the scene's triangle count; a triangle whose 9 values are all
 * zero keeps its geometric normal.  NULL removes them.  Call before
 * rendering and before lr_batch_create takes the scene. */
int lr_scene_set_normals(lr_scene* scene, const float* corner_normals, uint32_t n_triangles);

/* Renders with LiAR's AdjointPhotonTracer integrand
 * (src/liar/tracers/adjoint_photon_tracer.cpp:77-292) as a wavefront path
 * tracer and accumulates into a Raster-style film
 * (src/liar/output/raster.cpp:547-606): film_xyz (device, 3*W*H floats,
 * planes X, Y, Z) += sum of sample XYZ, film_weight (device, W*H) += sample
 * count.  Deterministic for a given seed.  stats may be NULL. */
int lr_render(lr_scene* scene, const lr_camera* camera, const lr_render_params* params,
              float* film_xyz, float* film_weight, lr_render_stats* stats, void* stream);

/* ---- DirectLighting (src/liar/tracers/direct_lighting.cpp:42-377) --------
 * Every surface hit adds its emission (camera rays, and objects that are not
 * lights), a light estimate with shadow rays (RayTracer::
 * estimateLightContribution, src/liar/kernel/ray_tracer.cpp:188-268) and
 * the radiance of its specular (and, with trace_glossy, glossy) reflection
 * and transmission directions, traced recursively.  High quality (camera
 * rays and specular continuations of high quality): every light with its
 * own number of samples (lr_light.samples), each a light sample and a BSDF
 * sample combined by the power heuristic; otherwise secondary_light_samples
 * lights picked by power.  Rays that escape add nothing.  The random
 * numbers are fixed by the seed, and a path's radiance terms (from its ray
 * tree's emission and unoccluded shadow rays) are summed in 64-bit fixed
 * point (2^-32 per unit, up to 2^31), so films are bit-reproducible; the
 * film is accumulated as by lr_render. */
typedef struct lr_direct_params {
    uint32_t max_generation;           /* RayTracer::maxRayGeneration: rays of
                                          generation 0 (camera) .. this are cast
                                          (LiAR default 8, at most 23)        */
    uint32_t secondary_light_samples;  /* numSecondaryLightSamples (0: 1, <= 8) */
    uint32_t trace_glossy;             /* traceGlossy (LiAR default true)    */
} lr_direct_params;

/* params->max_depth, rr_depth, rr_survival and graph are ignored;
 * chunk_paths 0 means 4M camera paths per chunk (each chunk keeps two branch
 * queues of 2x its paths and a shadow queue of as many rays as its hits can
 * make; a queue that overflows is doubled and the chunk repeated, and one
 * that would pass 2^30 entries returns LR_E_INVALID and asks for a smaller
 * chunk).  dparams NULL: LiAR's defaults (8, 1, true). */
int lr_render_direct(lr_scene* scene, const lr_camera* camera, const lr_render_params* params,
                     const lr_direct_params* dparams, float* film_xyz, float* film_weight,
                     lr_render_stats* stats, void* stream);

/* ---- film output -----------------------------------------------------------
 * RgbSpace (src/liar/kernel/rgb_space.cpp:73-79, 335-386; rgb_space.h:124-140):
 * primaries and white point as CIE xy chromaticities and a gamma.  Linear RGB
 * maps to XYZ relative to CIE illuminant E (the renderer's white is a flat
 * spectrum) through the primaries' matrix and a Bradford adaptation from the
 * white point to E.  LiAR's sRGB: red (0.64, 0.33), green (0.30, 0.60), blue
 * (0.15, 0.06), white (0.3127, 0.3290), gamma 2.2. */
typedef struct lr_rgb_space {
    float red[2], green[2], blue[2], white[2];
    float gamma;
} lr_rgb_space;

/* Row-major 3x3 matrices of a space (either pointer may be NULL). */
int lr_rgb_space_matrices(const lr_rgb_space* space, float* rgb_to_xyz, float* xyz_to_rgb);

/* Raster::ToneMapping (src/liar/output/raster.cpp:30-40), same order. */
#define LR_TONEMAP_LINEAR 0u
#define LR_TONEMAP_COMPRESS_Y 1u
#define LR_TONEMAP_COMPRESS_RGB 2u
#define LR_TONEMAP_REINHARD2002_Y 3u
#define LR_TONEMAP_REINHARD2002_RGB 4u
#define LR_TONEMAP_EXPONENTIAL_Y 5u
#define LR_TONEMAP_EXPONENTIAL_RGB 6u
#define LR_TONEMAP_DUIKER_Y 7u
#define LR_TONEMAP_DUIKER_RGB 8u

typedef struct lr_tonemap_params {
    lr_rgb_space space;              /* destination space (the image writer's)  */
    uint32_t tone_mapping;           /* LR_TONEMAP_*                           */
    uint32_t auto_exposure;          /* Raster.autoExposure (LiAR default on)  */
    float exposure_stops;            /* used when auto_exposure == 0           */
    float exposure_correction_stops; /* added to either                        */
    float middle_grey;               /* Raster.middleGrey (LiAR: 0.184)        */
} lr_tonemap_params;

typedef struct lr_tonemap_info {
    float exposure_stops;     /* stops used (automatic: rounded to 1/3)        */
    float gain;               /* 2^(stops + correction)                        */
    float average_luminance;  /* log-average weighted Y of the covered pixels  */
    float max_luminance;      /* largest weighted Y of the final film          */
} lr_tonemap_info;

/* Raster::tonemap (src/liar/output/raster.cpp:306-494) of a film written by
 * lr_render / lr_render_direct / lr_render_guided: film_xyz (device, planes
 * X, Y, Z of width*height floats), film_weight (device), film_alpha (device
 * sum of sample alphas, or NULL: alpha 1 wherever the weight is positive).
 * Writes display RGBA, rows top to bottom: rgba (device, 4 floats per pixel,
 * 16-byte aligned) and/or rgba8 (device, 4 bytes per pixel, quantised as
 * LiAR's PNG writer does: clamp(255 x, 0, 255) truncated).  NaN pixels become
 * pink.  Automatic exposure is computed on the device; info (host, may be
 * NULL) returns it and then synchronises the stream.  LiAR's
 * maxSceneLuminance is a running maximum over every splat of the render;
 * here it is the maximum over the final film. */
int lr_tonemap(lr_context* ctx, const float* film_xyz, const float* film_weight, const float* film_alpha,
               uint32_t width, uint32_t height, const lr_tonemap_params* params, float* rgba,
               uint8_t* rgba8, lr_tonemap_info* info, void* stream);

/* Image codecs (src/liar/kernel/image_codec.cpp:38-45, src/liar/codecs/):
 * writes host RGBA floats (rows top to bottom, values as lr_tonemap made
 * them for this format's space) to `path`, chosen by its extension:
 *   .png        8-bit RGBA, zlib-deflated (lodepng.cpp:248-260 quantisation)
 *   .exr        half RGBA, uncompressed scanlines, chromaticities attribute
 *               (openexr.cpp:95-112; linear space)
 *   .hdr .rgbe .pic  Radiance RGBE, flat scanlines (linear space)
 *   .pfm        32-bit float RGB, bottom row first (linear space)
 *   .tga .targa 8-bit BGRA, uncompressed, top-left origin
 * space (may be NULL: sRGB) fills the EXR chromaticities. */
int lr_image_write(const char* path, const float* rgba, uint32_t width, uint32_t height,
                   const lr_rgb_space* space);

/* ---- batches: many scenes in one wavefront (no counterpart in LiAR) -------
 * A batch holds scenes built with lr_scene_create (non-empty geometry each);
 * their trees are concatenated into one node array (indices rebased, each
 * ray starting at its scene's root), so that one sequence of launches renders
 * them all: chunks of up to 4M paths are packed from whole pixels of the
 * scenes, shortest scene first, and every bounce's queue mixes the paths of
 * all scenes in its chunk.  Every path draws the random numbers it would
 * draw in lr_render of its scene alone, so each film equals that render's.
 * The batch keeps its own copy of the trees: the scenes must outlive it. */
typedef struct lr_batch lr_batch;
typedef struct lr_batch_item {
    lr_camera camera;
    lr_render_params params;  /* graph is ignored; the batch's chunk size is
                                 the smallest nonzero chunk_paths (0: 4M)     */
    float* film_xyz;          /* device, 3 * width * height (accumulated)    */
    float* film_weight;       /* device, width * height                      */
} lr_batch_item;
#define LR_BATCH_GRAPH 1u     /* capture the call's launches into a HIP graph */
int lr_batch_create(lr_context* ctx, lr_scene* const* scenes, uint32_t n, lr_batch** out);
int lr_batch_destroy(lr_batch* batch);
/* items: one per scene, in the batch's scene order.  scene_ms (optional, n
 * floats): milliseconds from the start of the call's first launch to the end
 * of the last chunk holding the scene's pixels (device wall clock). */
int lr_render_batch(lr_batch* batch, const lr_batch_item* items, uint32_t flags, float* scene_ms,
                    lr_render_stats* stats, void* stream);

/* ---- path-guiding network (no counterpart in LiAR) -------------------------
 * A learned proposal for the integrand's direction sampling: an MLP
 * (6 inputs -> 64-wide frequency encoding -> 5 hidden layers of 64, ReLU ->
 * 48 outputs) in BF16 on MFMA with FP32 accumulation.  Outputs 0..39 are an
 * 8-lobe von Mises-Fisher mixture over incident directions, 40 and 41 the
 * logits of the per-vertex strategy probabilities (light, guide, BSDF) =
 * 0.025 + 0.925 (r_l, (1 - r_l) r_g, (1 - r_l)(1 - r_g)), r = sigmoid, and
 * 42 the log of an estimate of the radiance the vertex reflects towards the
 * previous one (a linear readout of the last hidden layer, trained when the
 * radiance weight is > 0: lr_guide_set_radiance_weight).
 * Trained online with Adam on radiance samples the tracer produces.
 * Inputs: device float [n][8], per row (px, py, pz, theta_o/pi, phi_o/(2 pi),
 * (lambda - lambda_min)/(lambda_max - lambda_min), p_light, p_bsdf), the
 * first six in [0, 1]; the last two are read by training only (below).
 * Mixture: device float [n][48], per lobe (weight = softmax of the logits,
 * unit mean, kappa = 0.1 + e^min(s, 7.6)), then
 * (c_light, c_guide, c_bsdf, radiance e^clamp(out42, -20, 20), 0, 0, 0, 0).
 * Parameters (lr_guide_param_count() floats): per layer l (0..5) Wt_l
 * [n_out][64] row-major (n_out = 64, output layer 48 with rows 43..47 unused),
 * then the biases (5 x 64 + 48).
 * A guide's training scratch (workgroup partials, weight sums) is its own and
 * shared by its calls: calls that train one guide on different streams at the
 * same time race; inference alone may overlap. */
typedef struct lr_guide lr_guide;
int lr_guide_param_count(void);
int lr_guide_create(lr_context* ctx, uint64_t seed, float learning_rate, lr_guide** out);
int lr_guide_destroy(lr_guide* guide);
int lr_guide_get_params(const lr_guide* guide, float* host_params);
int lr_guide_set_params(lr_guide* guide, const float* host_params);
/* n_dev: optional device count (<= n) read by the kernel. */
int lr_guide_infer(lr_guide* guide, const float* inputs, uint32_t n, const uint32_t* n_dev,
                   float* mixture, void* stream);
/* One Adam step on sum_i w_i (-log q(dir_i | input_i) - [p_l + p_b > 0]
 * log(c_l p_l + c_g q + c_b p_b)) / sum_i w_i, where p_l, p_b are inputs 6
 * and 7 of row i (the light and BSDF sampling pdfs of dir_i) and q is held
 * fixed inside the second term, which trains the strategy probabilities;
 * with a radiance weight lambda > 0, plus lambda / n sum_i (1 - min(w_i / L_i,
 * 50))^2 with L_i the radiance output held fixed in the denominator, whose
 * gradient trains output row 42 only.
 * dirs: device float [n][3] unit vectors; weights: device float [n] >= 0.
 * loss_out (device float, optional): the summed normalised loss.
 * grad_out (device float [param_count], optional): the gradient.  A guide
 * created with learning_rate 0 computes gradients without updating. */
int lr_guide_train_step(lr_guide* guide, const float* inputs, const float* dirs,
                        const float* weights, uint32_t n, float* loss_out, float* grad_out,
                        void* stream);

/* Guided rendering (no counterpart in LiAR): at diffuse / glossy vertices the
 * integrand's one-sample MIS gains a third strategy, sampling the guide's
 * vMF mixture: light sampling takes `light_share` (LiAR: 1/2), and the rest
 * is split between the guide (`selection`) and BSDF sampling
 * (1 - selection), or the network sets all three per vertex
 * (learned_selection).  The power (or balance) heuristic over all three keeps
 * the estimator unbiased for any network, since the probabilities depend on
 * the vertex only and each is at least 0.025 when learned.  With `train`, the vertices of each
 * chunk are recorded (up to max_records, 0: 2^22) and `train_steps` Adam
 * steps, each on `train_batch` records (0: 65536) drawn uniformly with
 * replacement from all recorded ones, run after the chunk on the
 * weights f|cos|/p_mix * L_incident. */
typedef struct lr_guide_params {
    float selection;
    uint32_t train;
    uint32_t train_steps;
    uint32_t train_batch;
    uint32_t max_records;
    float weight_clamp; /* training weights clamped at this multiple of their
                           mean (0: no clamp); outliers such as sun caustics
                           otherwise dominate the loss */
    float light_share;  /* probability of light sampling at diffuse / glossy
                           vertices (LiAR: 0.5); the guide and BSDF share the
                           rest as selection : 1 - selection */
    uint32_t learned_selection; /* nonzero: the network's per-vertex strategy
                           probabilities (outputs 40, 41; each >= 0.025) replace
                           selection / light_share, and training fits them to
                           the recorded radiance (KL of the strategy mixture) */
    uint32_t balance_heuristic; /* nonzero: balance heuristic (the optimal
                           one-sample MIS weight) instead of LiAR's power heuristic */
    uint32_t image_weighting; /* nonzero: each training record's weight is also
                           multiplied by the path throughput before its vertex
                           (its share of the pixel), so vertices that matter to
                           the image dominate the fit rather than the brightest */
    uint32_t record_depth; /* record only the vertices of the first record_depth
                           bounces (0: all) */
    uint32_t guide_depth;  /* guided MIS at the vertices of the first
                           guide_depth bounces only; later ones sample as the
                           plain tracer (LiAR's 0.5 BSDF / 0.5 light) and run
                           no inference (0: every bounce) */
    uint32_t learned_rr;   /* nonzero: radiance-driven Russian roulette at the
                           guided diffuse / glossy vertices after a path's
                           first one (adjoint-driven RR, Vorba & Krivanek
                           2016): with L the network's estimate of the
                           radiance a vertex reflects towards the previous one
                           (mixture float 43) and I = throughput x L at the
                           path's first diffuse / glossy vertex, the path
                           survives with q = clamp(throughput x L / (rr_scale
                           x I), B x rr_min_survival, 1) and its throughput is
                           divided by q, B being the product of the path's
                           earlier 1 / q (so that it stays <= 1 /
                           rr_min_survival).
                           Unbiased for any network; LiAR's roulette
                           (rr_depth) keeps precedence */
    float rr_scale;        /* see learned_rr (0: 1) */
    float rr_min_survival; /* learned_rr: a path's roulette weight (the
                              product of its 1 / q) is at most the inverse
                              of this (0: 0.05) */
    uint32_t rr_start;     /* learned_rr only at vertices of generation >=
                              rr_start (0: from the path's second diffuse /
                              glossy vertex on) */
} lr_guide_params;
/* Weight of the radiance loss in the guide's training (default 0: the
 * radiance output, mixture float 43, is not trained).  Each training row adds
 * weight x mean(w) x (1 - min(w / L, 50))^2, w the row's weight (a one-sample
 * estimate of the radiance its vertex reflects) and L the output: the
 * gradient pulls L to the mean of w.  No LiAR counterpart. */
int lr_guide_set_radiance_weight(lr_guide* guide, float weight);
int lr_render_guided(lr_scene* scene, const lr_camera* camera, const lr_render_params* params,
                     lr_guide* guide, const lr_guide_params* guide_params, float* film_xyz,
                     float* film_weight, lr_render_stats* stats, void* stream);

/* ---- scene import --------------------------------------------------------
 * PLY meshes for TriangleMesh::loadPly (src/liar/scenery/triangle_mesh.cpp:183-281):
 * ascii and binary (either endianness), any numeric property types; reads
 * vertex x y z [nx ny nz] [u v] and face vertex_indices (or vertex_index).
 * Triangles as LiAR makes them: a polygon keeps only its last fan triangle
 * (i0, i_n-2, i_n-1) (LiAR stores the fan triangle after its loop), or every
 * fan triangle with LR_PLY_FAN_ALL.  Arrays are allocated by the library
 * (NULL when absent) and released with lr_mesh_data_free. */
typedef struct lr_mesh_data {
    uint32_t n_verts, n_tris;
    float* verts;     /* 3 * n_verts                   */
    float* normals;   /* 3 * n_verts, or NULL           */
    float* uvs;       /* 2 * n_verts, or NULL           */
    uint32_t* tris;   /* 3 * n_tris vertex indices       */
} lr_mesh_data;
#define LR_PLY_FAN_ALL 1u
int lr_load_ply(const char* path, uint32_t flags, lr_mesh_data* out);
void lr_mesh_data_free(lr_mesh_data* mesh);

/* ---- deterministic synthetic inputs (host, counter-based Philox4x32-10) --- */

/* Random triangle soup: centre c ~ U[0,1)^3, v_k = c + scale * xi_k,
 * xi_k ~ U[-1,1)^3.  verts: 9*nt floats, tris: 3*nt uint32 (identity). */
int lr_gen_random_triangles(uint64_t seed, uint32_t nt, float scale,
                            float* verts, uint32_t* tris);
/* Rays: origin ~ U[0,1)^3, direction uniform on S^2 (z = 1 - 2u,
 * phi = 2 pi v), tnear = 0, tfar = +inf or ~ U[tfar_lo, tfar_hi) when
 * tfar_hi > 0.  All outputs host arrays of n floats (SoA). */
int lr_gen_random_rays(uint64_t seed, uint32_t n, float tfar_lo, float tfar_hi,
                       float* ox, float* oy, float* oz, float* dx, float* dy,
                       float* dz, float* tnear, float* tfar);
/* Heightfield terrain over [0,1]^2: n_side^2 vertices, 2 (n_side-1)^2
 * triangles (two per cell, counter-clockwise seen from +z), z = height *
 * fBm(x, y): `octaves` octaves of gradient noise, frequency base_frequency *
 * 2^o, amplitude gain^o.  verts: 3 n_side^2 floats, tris: 6 (n_side-1)^2. */
int lr_gen_terrain(uint64_t seed, uint32_t n_side, uint32_t octaves, float gain,
                   float base_frequency, float height, float* verts, uint32_t* tris);
/* Philox4x32-10 block (Salmon et al., SC'11) for tests: out[4]. */
void lr_philox4x32(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2,
                   uint32_t c3, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
