/*
 * milton_pt.h -- C ABI of the MI355X path-tracing core for Milton.
 *
 * The library replaces Milton's unidirectional path-tracing path: the
 * renderer selected as "pathTracer" (mlt/src/milton/loaders/
 * MiltonJSONSceneLoader.cpp:718-719), i.e. PathTracer::_evaluate
 * (renderers/renderers/PathTracer.cpp:28-116) and everything below it
 * (kd-tree traversal, primitive intersection, BSDF and area-light sampling,
 * Russian roulette, film accumulation).  Paths in citations are relative to
 * mlt/src/milton/ of the reference.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative MPT_E* code on
 *    failure; mpt_last_error() returns the message of the last failure on the
 *    context (the reference has no error codes: it turns load failures into
 *    scene-parse errors, MiltonJSONSceneLoader.cpp:1782-1810).
 *  - Host arrays are owned by the caller and copied in or out.  Device memory
 *    is owned by the library and released by the matching destroy call.
 *  - A context is used by one host thread at a time.
 *  - All structs are plain data with explicit widths.
 *
 * Geometry conventions (shared with the CPU oracle):
 *  - triangles are two-sided Moller-Trumbore triangles of one mesh
 *    (shapes/MeshTriangle.cpp:92-169); per-vertex normals are the normalised
 *    sum of unit face normals (shapes/Mesh.cpp:335-377);
 *  - a sphere is (cx, cy, cz, r) (shapes/Sphere.cpp:16-45);
 *  - a quad is the reference "plane" shape (shapes/Plane.cpp:25-106): the
 *    unit square y=0, |x|,|z|<=0.5 mapped to center + x*ex + z*ez; its normal
 *    is normalize(ez x ex) and its area |ex x ez|;
 *  - quads, spheres and (mesh) triangles with an emissive material are the
 *    area lights sampled by next-event estimation, in that order (quads,
 *    spheres, then emitting shapes by first triangle); standalone one-sided
 *    triangles may not emit;
 *  - global primitive ids: triangles [0,nt), spheres [nt, nt+ns), quads
 *    [nt+ns, nt+ns+nq); MPT_MISS marks no hit.
 */
#ifndef MILTON_PT_H_
#define MILTON_PT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPT_ABI_VERSION 8u
#define MPT_MISS 0xFFFFFFFFu

enum {
  MPT_OK = 0,
  MPT_E_INVALID = -1,     /* bad argument or scene description */
  MPT_E_DEVICE = -2,      /* HIP runtime failure */
  MPT_E_NOMEM = -3,       /* device or host allocation failed */
  MPT_E_BUDGET = -4,      /* request exceeds the admission budget */
  MPT_E_NOTFOUND = -5,    /* unknown job handle */
  MPT_E_CANCELLED = -6,   /* job was cancelled */
  MPT_E_IO = -7           /* image file could not be written, or its format is unsupported */
};

enum { MPT_BSDF_DIFFUSE = 0, MPT_BSDF_PHONG = 1, MPT_BSDF_DIELECTRIC = 2 };

/* Replaces the Material property map keys read on the path:
 * "bsdf" type (materials/Material.cpp:67-92), "kd"/"ks" (DiffuseBSDF.cpp:45,
 * ModifiedPhongBSDF.cpp:51-62), "n" (ModifiedPhongBSDF.cpp:73), "ior" and
 * "transparency" (DielectricBSDF.cpp:48-53), and the oriented emitter's
 * radiance Le = power / (pi * area) (Emitter.cpp:34-48,
 * emitters/OrientedEmitter.cpp:17-23).  Le == 0 means not an emitter. */
typedef struct mpt_material {
  uint32_t type;
  float kd[3];
  float ks[3];
  float exponent;
  float ior;          /* index of refraction; the red channel when ior_g / ior_b are set */
  float transparency;
  float Le[3];
  /* Per-channel IOR of a dispersive dielectric (DielectricBSDF's "ior"
   * spectrum, materials/bsdfs/DielectricBSDF.cpp:49, read at the path's
   * channel, :65-87); 0 = ior. */
  float ior_g, ior_b;
  uint32_t pad;
} mpt_material;

/* Reconstruction kernels of KernelFilter::create (filters/KernelFilter.cpp:
 * 18-45): box, triangle, gaussian (sigma), mitchell (B, C), lanczosSinc (tau). */
enum {
  MPT_FILTER_BOX = 0,
  MPT_FILTER_TRIANGLE = 1,
  MPT_FILTER_GAUSSIAN = 2,
  MPT_FILTER_MITCHELL = 3,
  MPT_FILTER_LANCZOS = 4
};

/* A lat-long environment image for the background: EnvironmentMap
 * (materials/emitters/EnvironmentMap.cpp:24-95) looked up through its
 * material's kernel filter (KernelFilter::apply, filters/KernelFilter.cpp:
 * 124-160).  A ray that leaves the scene in direction d sees the image at
 * the UV of -d (Scene::getBackgroundRadiance, core/Scene.cpp:116-118). */
typedef struct mpt_env_map {
  uint32_t width, height; /* texels; both at least 1 */
  const float *rgb;       /* width * height * 3 floats, row-major from the top row (v = 0) */
  float repeat_u, repeat_v; /* "repeatU" / "repeatV" (EnvironmentMap.cpp:49-50); 0 = default 1 */
  uint32_t filter;        /* MPT_FILTER_*; the material's "filter", default triangle (Material.cpp:53) */
  uint32_t support;       /* "support", default 2 when 0; the kernel is (ceil(support) | 1)^2 taps, at most 15^2 */
  float params[2];        /* gaussian: sigma (0 = 1); mitchell: B, C (both 0 = 1/3, 1/3);
                             lanczosSinc: tau (0 = 3) */
} mpt_env_map;

/* Replaces the scene graph the loader flattens (loaders/SceneGraph.cpp:23-135)
 * and Scene::init (core/Scene.cpp:31-71). */
typedef struct mpt_scene_desc {
  uint32_t version; /* MPT_ABI_VERSION */
  uint32_t num_vertices;
  const float *vertices; /* 3 floats per vertex */
  uint32_t num_triangles;
  const uint32_t *indices;      /* 3 per triangle */
  const uint32_t *tri_material; /* 1 per triangle */
  uint32_t num_spheres;
  const float *spheres; /* cx, cy, cz, r */
  const uint32_t *sphere_material;
  uint32_t num_quads;
  const float *quads; /* center[3], ex[3], ez[3] */
  const uint32_t *quad_material;
  uint32_t num_materials;
  const mpt_material *materials;
  float env_radiance[3]; /* constant background radiance (Scene.cpp:116-118) */
  /* Optional (NULL = all two-sided): per-triangle flags.  MPT_TRI_ONE_SIDED
   * marks the reference's standalone Triangle, which culls back faces:
   * a ray with e1 . (d x e2) < EPSILON misses it (shapes/Triangle.cpp:245-270);
   * mesh triangles are two-sided (shapes/MeshTriangle.cpp:142-167). */
  const uint32_t *tri_flags;
  /* Optional (NULL = none): per-vertex normals (3 floats) for the vertices
   * whose vertex_normal_mask entry is nonzero (a mesh file's normals,
   * shapes/Mesh.cpp:335-377 computes the others); both or neither. */
  const float *vertex_normals;
  const uint8_t *vertex_normal_mask;
  /* Optional (NULL = the constant env_radiance): an environment image in
   * place of the constant background. */
  const mpt_env_map *env_map;
  /* Optional point lights: the reference's "point" shape with an "omni"
   * emitter (shapes/PointShape.cpp, emitters/OmniEmitter.cpp).  6 floats per
   * light: position x, y, z and power r, g, b; radiance P / (4 pi) in every
   * direction, one light sample per vertex (DirectIllumination.cpp:96-98). */
  uint32_t num_points;
  const float *points;
  /* Optional (NULL = triangles that share an emitting material form one
   * shape): the reference shape (Mesh) each triangle belongs to.  The
   * triangles of one emitting shape are one light, sampled as
   * Mesh::getPosition does (shapes/Mesh.cpp:442-450: a triangle picked by
   * the sample's u, a point at barycentrics (r1, v, r2) / (r1 + v + r2) with
   * two more uniforms) and weighted by the shape's world-space area
   * (Mesh::_getSurfaceArea, :452-466). */
  const uint32_t *tri_shape;
  /* Optional (NULL = 2r times the identity): per sphere, the linear part of
   * its object-to-world transform (9 floats, column-major: the images of the
   * object axes).  An emitting sphere places its light samples through it,
   * as Sphere::getPosition does (shapes/Sphere.cpp:117-124: a point at
   * object-space distance 1 from the centre near the +z axis), with area pi
   * (Sphere::_getSurfaceArea ignores the scale, :138-141). */
  const float *sphere_frames;
} mpt_scene_desc;

enum { MPT_TRI_ONE_SIDED = 1u };

/* Replaces kdTreeAccel::BuildParams (accel/kdTreeAccel.h:68-116). */
typedef struct mpt_kd_params {
  uint32_t leaf_size;     /* a node with <= leaf_size references becomes a leaf (kdMinPrimitives);
                             0 = default 4 */
  uint32_t max_depth;     /* kdMaxDepth (at most 48) */
  float cost_traversal;   /* kdCostTraversal */
  float cost_intersect;   /* parsed but unused by the reference (kdCostIntersect) */
  float empty_bias;       /* kdEmptyBias: SAH cost factor of a split with an empty side; 0 = default
                             1.0 (no bonus: on this device a node visit costs more than the
                             references a bonus split saves) */
  uint32_t num_threads;   /* kdNoThreads; 0 = all host cores */
  uint32_t bins;          /* SAH bins per axis (0 = default) */
  uint32_t no_clip;       /* 1: split triangles by their bounds only (0: clip them to the cells) */
} mpt_kd_params;

/* Replaces ThinLensCamera's keys with aperture 0 (cameras/PinholeCamera.cpp:22,
 * cameras/ThinLensCamera.cpp:62-86). */
typedef struct mpt_camera {
  float eye[3];
  float look[3];
  float up[3];
  float height_angle; /* degrees, "heightAngle" */
  float aspect;       /* "aspectRatio" */
  /* Thin lens (ThinLensCamera::init / getWorldRay, cameras/ThinLensCamera.cpp:
   * 62-129, 131-172); aperture 0 is the pinhole (PinholeCamera.cpp:21-25). */
  float aperture;       /* lens diameter in metres (the file's "aperture" is in mm) */
  float focal_length;   /* metres: fstop * aperture */
  float focal_distance; /* metres; with autofocus, the value used when the focus ray misses */
  uint32_t autofocus;   /* nonzero: focal distance = hit distance of the ray through focal_point */
  float focal_point[2]; /* [0,1]^2 film position of the focus ray ("focalPoint", default 0.5, 0.5) */
} mpt_camera;

/* One render job: the film ("width"/"height", RenderOutput.cpp:34), samples
 * per pixel ("noSuperSamples", SuperSampleGenerator.cpp:26-28), a maximum
 * path depth (a cap the reference does not have) and the RNG seed ("-seed="). */
typedef struct mpt_job {
  mpt_camera camera;
  uint32_t width, height, spp, max_depth;
  uint32_t seed, job_id;
  /* light samples per light per non-specular vertex (DirectIllumination's
   * noDirectSamples, renderers/DirectIllumination.cpp:33,98-133), jittered
   * over floor(sqrt(n))^2 strata of the light; 0 means 1 */
  uint32_t direct_samples;
  /* Film reconstruction (ReconstructionRenderOutput::addSample,
   * renderers/outputs/ReconstructionRenderOutput.cpp:30-83): 0 bins each
   * sample into its own pixel with weight 1 (RenderOutput.cpp:72-76); else
   * 1 + MPT_FILTER_* splats it onto the 3 x 3 pixels around its own with
   * weight filter(x - X, y - Y) > 0, for integer pixel corners (x, y) and
   * the continuous position (X, Y) in pixels; the film holds numerator and
   * denominator.  A filtered render adds into the film (mpt_render_film
   * clears it first, mpt_render_film_range does not: tiles then merge). */
  uint32_t filter;
  uint32_t filter_support; /* the filter's "support", 0 = 2 */
  float filter_params[2];  /* gaussian sigma (0 = 1); mitchell B, C (both 0 = 1/3); lanczosSinc tau (0 = 3) */
  /* 1: the renderer's efficientDirect = false (renderers/renderers/
   * PathTracer.cpp:23, 50-52, 70-73): emitted radiance is counted at every
   * vertex and next-event estimation is off.  0 (the reference default,
   * efficientDirect = true): emission only at depth 0 or after a reflective
   * specular bounce, next-event estimation at every non-specular vertex. */
  uint32_t direct_off;
  /* Next-event estimation's choice of lights.  0 (the reference's
   * DirectIllumination, renderers/DirectIllumination.cpp:90-135): every
   * light gets direct_samples shadow rays at each vertex.  1: each of the
   * direct_samples shadow rays goes to one light drawn with probability
   * proportional to its emitted power, as the scene's EmitterSampler draws
   * lights (stats/samplers/EmitterSampler.cpp:19-61: power averaged over the
   * channels, the normalized powers walked in order until their sum reaches
   * the uniform), its contribution divided by that probability; point lights
   * are drawn with the area lights.  Same expectation, direct_samples shadow
   * rays per vertex however many lights the scene has.  With one light and
   * no point light both choices give the same paths. */
  uint32_t light_select;
  /* MPT_RENDERER_PATH (0): PathTracer (renderers/renderers/PathTracer.cpp).
   * MPT_RENDERER_BIDIR (1): BidirectionalPathTracer (renderers/renderers/
   * BidirectionalPathTracer.cpp:34-174 over renderers/utils/Path.cpp): an
   * eye and a light subpath, every split of every prefix/suffix weighted by
   * the power heuristic, the sum clamped to [0, 1]; max_depth caps the
   * bounces of each subpath; direct_samples, direct_off and light_select do
   * not apply.  It needs a pinhole camera and at least one quad, mesh or
   * point light (emitting spheres: MPT_E_INVALID); the background is
   * not an emitter of this renderer.
   * MPT_RENDERER_MLT (2): MLTRenderer (renderers/mlt/MLTRenderer.cpp:
   * 70-235, MLTMarkovProcess.cpp:38-176): mlt_chains Markov chains (the
   * reference's noRenderThreads) of mlt_mutations mutations each, started
   * from the splits of mlt_seed_paths generated paths weighted by their
   * contribution; the film is RenderOutput's MLT histogram
   * (renderers/RenderOutput.cpp:129-216).  width and height give the film;
   * spp, max_depth, direct_samples, direct_off, light_select and filter do
   * not apply (filter must be 0).  Needs a pinhole camera and a quad, mesh
   * or point light (emitting spheres: MPT_E_INVALID). */
  uint32_t renderer;
  /* Metropolis light transport (MPT_RENDERER_MLT only). */
  uint32_t mlt_chains;         /* Markov chains, >= 1 */
  uint32_t mlt_mutations;      /* mutations per chain, at most 32767 */
  uint32_t mlt_seed_paths;     /* mltNoInitialPaths (reference default 10000; 0 means 1) */
  uint32_t mlt_max_depth;      /* mltMaxDepth: most vertices of a path, 2..16 (reference default 10) */
  uint32_t mlt_max_rejections; /* mltMaxConsequtiveRejections (reference default 500; 0: no limit) */
  uint32_t mlt_flags;          /* bit 0: mltFilterProposed = false (the proposal-count median filter off) */
  float mlt_probs[3];          /* mltBidirPathMutationProb, mltLensSubpathMutationProb,
                                  mltPerturbationPathMutationProb (reference default 1, 1, 1) */
} mpt_job;

enum { MPT_RENDERER_PATH = 0, MPT_RENDERER_BIDIR = 1, MPT_RENDERER_MLT = 2 };

typedef struct mpt_ray {
  float o[3];
  float d[3]; /* unit length */
} mpt_ray;

typedef struct mpt_hit {
  float t;       /* INFINITY on a miss */
  uint32_t prim; /* global primitive id or MPT_MISS */
} mpt_hit;

typedef struct mpt_render_stats {
  uint64_t paths;
  uint64_t camera_rays;
  uint64_t extension_rays; /* closest-hit rays, camera rays included */
  uint64_t shadow_rays;
  double render_ms;        /* device time of the whole render (HIP events) */
  double extend_ms;        /* sum of closest-hit launch durations (MPT_OPT_TIMING) */
  double shade_ms;
  double shadow_ms;
  uint64_t extend_launches;
  uint64_t nodes_visited;  /* kd nodes fetched by closest-hit rays (MPT_OPT_COUNTERS) */
  uint64_t prims_tested;   /* leaf references tested by closest-hit rays (MPT_OPT_COUNTERS) */
  double extend_wall_ms;   /* time during which at least one closest-hit launch ran: launches on
                              different lanes overlap, so this is below extend_ms */
  double shade_wall_ms;
  double shadow_wall_ms;
  /* Metropolis light transport: mutations accepted, samples added to the
   * film (RenderOutput's sample count) and b, the seed paths' mean weight */
  uint64_t mlt_accepted;
  uint64_t mlt_samples;
  double mlt_b;
} mpt_render_stats;

/* Context options (mpt_set_option).  No reference counterpart: these are the
 * knobs of the device path (queue depth) and its instrumentation. */
enum {
  MPT_OPT_BATCH_PATHS = 1, /* paths per wavefront batch (queue depth), default 2^25 */
  MPT_OPT_TIMING = 2,      /* 1: time every kernel launch with HIP events into *_ms */
  MPT_OPT_COUNTERS = 3,    /* 1: count kd nodes and references visited by closest-hit rays */
  MPT_OPT_GRAPHS = 4,      /* 1 (default): replay each batch's kernels from a captured hipGraph */
  MPT_OPT_STACK = 5,       /* traversal short-stack rows per ray: 8 (default), 12, 16 or 24; a ring of n rows holds n - 1 entries */
  MPT_OPT_PERSISTENT = 6,  /* closest-hit kernel form: 0 one ray per lane, grid-stride over 8
                              blocks per CU; 1 (default) persistent waves that refill finished
                              lanes from 128-ray chunks of the queue each wave claims; 2 one ray
                              per lane, one grid slot per queue slot (also the shadow-ray form) */
  MPT_OPT_LANES = 7,       /* streams (each with its own batch workspace) that consecutive batches
                              alternate over, 1..4, default 3: one batch's kernel tails overlap
                              another's bulk */
  MPT_OPT_TRACE_BLOCK = 10, /* threads per block of the closest-hit kernel in persistent waves
                               (form 1): 128 (default) or 256 (the same waves in half the blocks;
                               with MPT_OPT_TREE_TOP, 256) */
  MPT_OPT_SHADE_BLOCK = 11, /* threads per block of the shading kernel: 512 (default) or 256 (one
                               queue atomic per block, so 512 halves the atomics) */
  MPT_OPT_TREE_TOP = 9,    /* 1: the closest-hit kernel in persistent waves (form 1) stages the top
                              of the kd-tree (its first 512 nodes, laid out breadth-first) in LDS
                              and reads those nodes from there; 0 (default) reads every node
                              from memory */
  MPT_OPT_FAIL_BATCH = 8   /* failure injection for tests of the callers' error paths: n > 0 makes
                              the n-th batch started from now on fail with MPT_E_DEVICE after its
                              descriptors are staged and before its kernels are enqueued (the
                              setting then returns to 0); 0 (default) off */
};

typedef struct mpt_context mpt_context;
typedef struct mpt_scene mpt_scene;
typedef struct mpt_film mpt_film;

/* Opens a context on a HIP device.  (No reference counterpart: the
 * reference has no device.) */
int mpt_open(int device, mpt_context **ctx);
/* Closes the context and destroys the scenes and films still open on it;
 * their handles become invalid. */
int mpt_close(mpt_context *ctx);
/* The message of the context's last failure, from any thread (a server's
 * worker records its batches' failures here too); the string is the calling
 * thread's copy, valid until that thread calls mpt_last_error again. */
const char *mpt_last_error(const mpt_context *ctx);
int mpt_abi_version(void);

/* Uploads a scene and builds its kd-tree.  Replaces Scene::init
 * (core/Scene.cpp:31-71) with ShapeSet::init / Mesh::init /
 * kdTreeAccel::init (shapes/ShapeSet.cpp:36-80, shapes/Mesh.cpp:211-251,
 * accel/kdTreeAccel.cpp:403-472).  kd may be NULL for defaults. */
int mpt_scene_create(mpt_context *ctx, const mpt_scene_desc *desc, const mpt_kd_params *kd,
                     mpt_scene **scene);
int mpt_scene_destroy(mpt_scene *scene);
/* out: [nodes, leaves, references, max_depth, device_bytes, build_ms, max_leaf, triangles] */
int mpt_scene_stats(const mpt_scene *scene, uint64_t out[8]);

/* Host-only kd build and trace: the tree mpt_scene_create would upload, built
 * without a device, and closest hits through it computed on host threads by
 * the same traversal and intersection code the kernels run.  For inspecting
 * and tuning trees and for checking the traversal where no device is present.
 * counters (optional) = [kd nodes fetched, leaf references tested]. */
typedef struct mpt_kd_tree mpt_kd_tree;
int mpt_kd_build(const mpt_scene_desc *desc, const mpt_kd_params *kd, mpt_kd_tree **tree);
/* out: [nodes, leaves, references, max_depth, build_ms, max_leaf] */
int mpt_kd_stats(const mpt_kd_tree *tree, uint64_t out[6]);
int mpt_kd_trace(const mpt_kd_tree *tree, const mpt_ray *rays, uint64_t n, mpt_hit *hits, uint64_t counters[2],
                 uint32_t nthreads);
int mpt_kd_destroy(mpt_kd_tree *tree);

/* Closest hit for n host rays.  Replaces Scene::getIntersection
 * (core/Scene.cpp:104-108) -> kdTreeAccel::getIntersection
 * (accel/kdTreeAccel.cpp:893-1016).  Hits with t <= 1e-7 are ignored. */
int mpt_trace_closest(mpt_scene *scene, const mpt_ray *rays, uint64_t n, mpt_hit *hits);
/* The same closest hits traced by the render's own closest-hit kernel (the
 * context's kernel form and short-stack depth) with its work counters:
 * counters = [kd nodes fetched, leaf references tested] summed over the rays,
 * the per-ray counts behind the algorithmic bytes of the closest-hit kernel;
 * mpt_kd_trace counts the same walk on the host.  At most 2^31 rays. */
int mpt_trace_closest_stats(mpt_scene *scene, const mpt_ray *rays, uint64_t n, mpt_hit *hits,
                            uint64_t counters[2]);
/* Any hit with 1e-7 < t < tmax[i].  Replaces Scene::intersects
 * (core/Scene.cpp:110-114) -> kdTreeAccel::intersects (:1018-1094). */
int mpt_trace_any(mpt_scene *scene, const mpt_ray *rays, const float *tmax, uint64_t n,
                  uint8_t *occluded);

/* One emitter point per sample, as EmitterSampler::sample and ::getPd draw
 * them for a light subpath's first vertex (stats/samplers/EmitterSampler.cpp:
 * 51-96, called from renderers/utils/Path.cpp:233-236): a light drawn by
 * emitted power (the rule of mpt_job.light_select), then a point from that
 * shape's getRandomPoint:
 *   quad    Shape::getRandomPoint (shapes/Shape.cpp:45-50): Plane::getPosition
 *           at two uniforms;
 *   mesh    Shape::getRandomPoint with Mesh::getPosition (shapes/Mesh.cpp:
 *           442-450): a triangle by u, barycentrics (r1, v, r2) / sum;
 *   sphere  Sphere::getRandomPoint (shapes/Sphere.cpp:107-115): a uniform
 *           point s on the unit sphere, its (u, v) from Sphere::_getUV
 *           (u = 1 - atan2(s_z, s_x) / 2pi, v = 1/2 - asin(2 s_y) / pi), then
 *           Sphere::getPosition(u, v) through the sphere's transform.  For
 *           |s_y| > 1/2 (half of the draws) asin's argument leaves [-1, 1] and,
 *           as in the reference, position and normal are NaN;
 *   point   its position, normal 0 (PointShape has none).
 * pdf_area is getPd's value: 1 / (sum of the lights' surface areas), the
 * same for every sample (points have area 0; 1 if the sum is 0);
 * pdf_select the drawn light's probability.  Light indices number the
 * scene's area lights (emitting quads, spheres, then meshes) followed by its
 * point lights.  Sample i uses Philox keyed by (seed, i): runs are
 * reproducible.  MPT_E_INVALID for a scene without lights. */
typedef struct mpt_emitter_point {
  float p[3];
  uint32_t light;
  float n[3];
  float pdf_area;
  float pdf_select;
  uint32_t pad;
} mpt_emitter_point;
int mpt_emitter_sample(mpt_scene *scene, uint64_t n, uint32_t seed, mpt_emitter_point *out);

/* Renders a job and copies the film (numerator / denominator per pixel,
 * RGB float, row-major) to film_rgb (W*H*3 floats).  Replaces
 * PointSampleRenderer::render (renderers/PointSampleRenderer.cpp:24-81) over
 * RayTracer::sample -> PathTracer::_evaluate -> RenderOutput::addSample.
 * stats may be NULL. */
int mpt_render(mpt_scene *scene, const mpt_job *job, float *film_rgb, mpt_render_stats *stats);

/* Device-resident film for benchmarking and serving: render into it without
 * a host copy, then download when needed. */
int mpt_film_create(mpt_context *ctx, uint32_t width, uint32_t height, mpt_film **film);
int mpt_film_destroy(mpt_film *film);
int mpt_render_film(mpt_scene *scene, const mpt_job *job, mpt_film *film, mpt_render_stats *stats);
/* Pixels [pix_begin, pix_end) (row-major) of the job only, accumulated into
 * film: a tile of the job, the unit a job is split into across devices or
 * over time.  Paths are the same as in mpt_render_film. */
int mpt_render_film_range(mpt_scene *scene, const mpt_job *job, uint32_t pix_begin, uint32_t pix_end,
                          mpt_film *film, mpt_render_stats *stats);
int mpt_film_download(mpt_film *film, float *film_rgb);
/* The film's accumulators as stored: numerator RGB and denominator per pixel
 * (W*H*4 floats, row-major), the ProgressiveFilterValue pair
 * (filters/ProgressiveFilterValue.h:20-37).  Upload replaces them: films
 * rendered as tiles or on several devices are merged by summing these. */
int mpt_film_download_raw(mpt_film *film, float *rgbw);
int mpt_film_upload_raw(mpt_film *film, const float *rgbw);

/* Tone maps.  Replaces ToneMap::create (renderers/ToneMap.cpp:17-31):
 * "reinhard" and "default" select Reinhard, any other name the gamma-only
 * map. */
enum {
  MPT_TONEMAP_REINHARD = 0, /* ReinhardToneMap::map (renderers/tonemaps/ReinhardToneMap.cpp:25-98) */
  MPT_TONEMAP_GAMMA = 1     /* NullToneMap::map (renderers/tonemaps/NullToneMap.cpp:21-46) */
};
uint32_t mpt_tonemap_from_name(const char *name);

/* The film's finalized 8-bit image, computed on the device: per pixel value
 * = numerator / denominator (0 where the denominator is 0), then the tone
 * map over the whole film.  rgba: W*H*4 bytes, row-major from the top row,
 * bytes r, g, b, a = 255.  Replaces RenderOutput::getFinalizedOutput
 * (renderers/RenderOutput.cpp:129-231) for point-sample renderers. */
int mpt_film_tonemap(mpt_film *film, uint32_t tonemap, uint8_t *rgba);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops or harnesses that
 * forbid hipMalloc): the same bytes as mpt_film_tonemap, computed with a
 * fixed scratch array the library holds on each device for the life of the
 * process (16 MiB: the reduction partials and the 8-bit staging image)
 * instead of buffers allocated with the film.  Allocates nothing.  Films
 * above MPT_NOALLOC_MAX_PIXELS pixels fail with MPT_E_INVALID.  All calls on
 * one device share that array: calls that overlap on one device race. */
#define MPT_NOALLOC_MAX_PIXELS 4194304u
int mpt_film_tonemap_noalloc(mpt_film *film, uint32_t tonemap, uint8_t *rgba);

/* Writes the film to a file chosen by the path's extension (case-insensitive),
 * as FileRenderOutput::finalize -> _save (renderers/outputs/FileRenderOutput.cpp:
 * 45-76) -> Image::save (common/image/Image.cpp:103-129) does:
 *   .png  8-bit RGB PNG of the tone-mapped image
 *   .ppm  binary PPM (P6) of the tone-mapped image
 *   .hdr / .rgbe  Radiance RGBE, run-length encoded (HDRUtils::saveHDR,
 *         common/image/HDRUtils.cpp:114-148)
 *   .pfm  little-endian RGB PFM (HDRUtils::savePFM, :181-184, is
 *         unimplemented in the reference; here it is written)
 * HDR formats hold the tone-mapped bytes / 255, as the reference's
 * RgbaImage output does, unless MPT_SAVE_RADIANCE asks for the linear film
 * values.  Other extensions (jpg, gif, tiff, exr) fail with MPT_E_IO. */
enum { MPT_SAVE_RADIANCE = 1u };
int mpt_film_save(mpt_film *film, const char *path, uint32_t tonemap, uint32_t flags);

/* The writers behind mpt_film_save, on host images (no device needed):
 * rgba is W*H*4 bytes (alpha ignored), rgb W*H*3 floats, both row-major
 * from the top row.  An LDR image written to an HDR format becomes
 * bytes / 255; an HDR image can only go to .hdr, .rgbe or .pfm. */
int mpt_image_write_ldr(const char *path, uint32_t width, uint32_t height, const uint8_t *rgba);
int mpt_image_write_hdr(const char *path, uint32_t width, uint32_t height, const float *rgb);

/* Reads an image as linear RGB floats (row-major from the top row), as
 * Image::load (common/image/Image.cpp:67-101) does for the formats it reads
 * without Qt or OpenEXR: .hdr / .rgbe (Radiance RGBE, flat or run-length
 * encoded scanlines, "-Y H +X W"; HDRUtils::loadHDR, HDRUtils.cpp:36-70,
 * decoding byte * 2^(e-136)), .pfm (unimplemented in the reference,
 * HDRUtils.cpp:108-112), and binary .ppm (bytes / 255, as 8-bit images read
 * through RgbaImage).  Sets *width and *height; fills rgb when it holds at
 * least width * height * 3 floats (capacity, in floats), else returns
 * MPT_E_INVALID after setting the size, so a first call with rgb = NULL
 * asks for it.  MPT_E_IO if the file cannot be read or parsed. */
int mpt_image_read(const char *path, uint32_t *width, uint32_t *height, float *rgb, size_t capacity);

/* Per-path radiance of pixels [pix_begin, pix_end) of a job, all samples,
 * ordered (pixel, sample): L_out holds (pix_end-pix_begin)*spp*3 floats.
 * Test hook: the value PathTracer::_evaluate returns for each camera ray. */
/* Metropolis light transport's chains step by step, for checking: an
 * MPT_RENDERER_MLT job's seeds and chain starts, then its first `steps`
 * mutations (at most mlt_mutations) of every chain, recording per chain and
 * mutation six doubles: the mutation (0 bidirectional, 1 lens subpath,
 * 2 perturbation), its acceptance probability, accepted (0/1), the mutated
 * path's vertices (0 when rejected as invalid), the tentative sample's film
 * bin, the chain's path vertices after.  rec holds mlt_chains * steps * 6
 * doubles; b and entries (optional) get the seeds' mean weight and count.
 * No reference counterpart: the reference's chains run unobserved. */
int mpt_mlt_trace(mpt_scene *scene, const mpt_job *job, uint32_t steps, double *rec, double *b,
                  uint64_t *entries);
int mpt_render_paths(mpt_scene *scene, const mpt_job *job, uint32_t pix_begin, uint32_t pix_end,
                     float *L_out);

/* Scene files.  Replaces MiltonJSONSceneLoader::parse
 * (loaders/MiltonJSONSceneLoader.cpp:133-337) for the shapes, materials,
 * camera, output and renderer keys this path renders: the scene graph is
 * flattened into an mpt_scene_desc (owned by the mpt_scenefile) and the
 * camera, film size, samples per pixel, depth cap and seed into an mpt_job.
 * Unsupported elements fail with a message in err instead of being dropped.
 * Host only; no context needed.  Relative mesh paths resolve against the
 * scene file's directory (mpt_scenefile_load) or base_dir (parse). */
typedef struct mpt_scenefile mpt_scenefile;
int mpt_scenefile_load(const char *path, mpt_scenefile **scenefile, char *err, size_t err_len);
int mpt_scenefile_parse(const char *text, const char *base_dir, mpt_scenefile **scenefile, char *err,
                        size_t err_len);
const mpt_scene_desc *mpt_scenefile_desc(const mpt_scenefile *scenefile);
int mpt_scenefile_job(const mpt_scenefile *scenefile, mpt_job *job);
/* The output's tone map (MPT_TONEMAP_*): "tonemap" of the output object,
 * read as ToneMap::create does (renderers/RenderOutput.cpp:52-56). */
uint32_t mpt_scenefile_tonemap(const mpt_scenefile *scenefile);
int mpt_scenefile_destroy(mpt_scenefile *scenefile);

/* Autotuning.  No reference counterpart: the reference's only knobs on this
 * path are kdTreeAccel::BuildParams (accel/kdTreeAccel.h:68-116), fixed per
 * scene file.  These entry points time a sample of a job (its camera, depth
 * and seed at fewer samples per pixel over a centred pixel window, about
 * sample_paths paths) under candidate settings and keep the fastest.  None of
 * the settings changes a rendered result. */
typedef struct mpt_tune_result {
  uint64_t batch_paths;         /* MPT_OPT_BATCH_PATHS chosen */
  uint32_t stack, persistent;   /* MPT_OPT_STACK, MPT_OPT_PERSISTENT chosen */
  uint32_t leaf_size;           /* kd leaf size of the scene */
  uint32_t trials;              /* sample renders timed */
  double paths_per_s;           /* sampled rate at the chosen settings */
  double baseline_paths_per_s;  /* sampled rate at the settings on entry */
  uint32_t trace_block, shade_block;   /* MPT_OPT_TRACE_BLOCK, MPT_OPT_SHADE_BLOCK chosen */
} mpt_tune_result;

/* Engine settings (persistent, stack, block sizes, batch paths) for a
 * resident scene; the fastest are left set on the scene's context. */
int mpt_tune(mpt_scene *scene, const mpt_job *job, uint64_t sample_paths, mpt_tune_result *result);
/* Also the kd leaf size (1..16, hill climb from kd->leaf_size or 4): builds
 * the scene once per candidate and returns the fastest build in *scene. */
int mpt_tune_scene(mpt_context *ctx, const mpt_scene_desc *desc, const mpt_kd_params *kd, const mpt_job *job,
                   uint64_t sample_paths, mpt_scene **scene, mpt_tune_result *result);

/* ------------------------------------------------------------------------
 * Batching and serving.  No reference counterpart: the reference renders one
 * job per Renderer::render call (renderers/PointSampleRenderer.cpp:24-81).
 * A server keeps scenes resident, admits jobs while their films fit an HBM
 * budget, and packs the remaining pixels of all admitted jobs that share a
 * scene into common batches, so small and large jobs run together and new
 * jobs join between batches.  Every pixel's samples stay in one batch and use
 * the job's own random streams, so a job's film is bit-identical whether it
 * runs alone or mixed with others.
 * ------------------------------------------------------------------------ */
typedef struct mpt_server mpt_server;

typedef struct mpt_server_config {
  uint64_t hbm_budget;   /* device bytes for the ray queues of the server's batches and the films of
                            admitted jobs (0 = 64 GiB).  The queues take at most half: the server's
                            batch size is lowered until every lane's workspace fits there, and a job
                            is admitted while its film fits the rest. */
  uint32_t max_active;   /* jobs admitted at once (0 = 256) */
  uint32_t max_inflight; /* batches enqueued ahead of completion (0 = 2) */
} mpt_server_config;

enum { MPT_JOB_QUEUED = 0, MPT_JOB_RUNNING = 1, MPT_JOB_DONE = 2, MPT_JOB_CANCELLED = 3, MPT_JOB_FAILED = 4 };

typedef struct mpt_job_status {
  uint32_t state;
  uint32_t error;          /* MPT_E_* of a failed job */
  uint64_t pixels_done, pixels_total;
  double submit_ms, start_ms, done_ms; /* server clock: ms since mpt_server_create */
} mpt_job_status;

int mpt_server_create(mpt_context *ctx, const mpt_server_config *cfg, mpt_server **server);
int mpt_server_destroy(mpt_server *server);
/* Queues a job on a resident scene.  MPT_E_BUDGET if its film alone exceeds
 * the budget; otherwise it waits until admission.  Thread-safe. */
int mpt_submit(mpt_server *server, mpt_scene *scene, const mpt_job *job, uint64_t *handle);
int mpt_status(mpt_server *server, uint64_t handle, mpt_job_status *status);
int mpt_cancel(mpt_server *server, uint64_t handle);
/* Copies a finished job's film (RGB, W*H*3 floats) and releases the job. */
int mpt_result(mpt_server *server, uint64_t handle, float *film_rgb);
/* Blocks until the job leaves the queued/running states; returns 1 if
 * timeout_ms passed first, 0 when it finished. */
int mpt_wait(mpt_server *server, uint64_t handle, uint32_t timeout_ms);
/* Synchronous driving: one scheduling round, or rounds until idle. */
int mpt_server_step(mpt_server *server, uint32_t *batches_run);
int mpt_server_drain(mpt_server *server);
/* Asynchronous driving: a worker thread runs rounds until stopped. */
int mpt_server_start(mpt_server *server);
int mpt_server_stop(mpt_server *server);
/* out: [batches, paths, extension rays, shadow rays, jobs done, peak film bytes] */
int mpt_server_stats(mpt_server *server, uint64_t out[6]);
/* out: [budget, bytes reserved for the lanes' ray queues, film bytes now, peak film bytes,
 * paths per batch, ray-queue bytes the lanes hold now] */
int mpt_server_memory(mpt_server *server, uint64_t out[6]);

/* Paths per batch the engine plans for a job with direct_samples light
 * samples per area light (0 = 1) on a scene with num_lights area lights and
 * num_points point lights, given the context's MPT_OPT_BATCH_PATHS: each
 * path of a batch reserves one shadow ray (32 bytes) per light sample, and
 * the batch is lowered so that a batch's shadow rays stay within 2^28 (but
 * not below 1024 paths).  Host only; no reference counterpart. */
uint64_t mpt_batch_paths_for(uint64_t batch_paths, uint32_t num_lights, uint32_t num_points,
                             uint32_t direct_samples);

/* Synchronises the context's stream. */
int mpt_sync(mpt_context *ctx);
/* Sets a context option (MPT_OPT_*). */
int mpt_set_option(mpt_context *ctx, uint32_t option, uint64_t value);
int mpt_get_option(mpt_context *ctx, uint32_t option, uint64_t *value);

#ifdef __cplusplus
}
#endif
#endif
