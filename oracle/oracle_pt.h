/*
 * oracle_pt.h -- C interface of the CPU oracle (test infrastructure only).
 *
 * The oracle is a double-precision CPU restatement of Milton's unidirectional
 * path tracer (mlt/src/milton/renderers/renderers/PathTracer.cpp) and of every
 * function it calls on the hot path: the SAH kd-tree build and traversal
 * (accel/kdTreeAccel.cpp), the naive linear accel (accel/NaiveSpatialAccel.cpp),
 * mesh/standalone triangle, plane and sphere intersection, the Lambertian,
 * modified-Phong and dielectric BSDFs, oriented area emitters, next-event
 * estimation (renderers/DirectIllumination.cpp), the jittered super-sample
 * pattern and the pinhole camera.
 *
 * It exists to check the HIP product path. Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it; the product never does.
 *
 * Scene arrays use the same conventions as include/milton_pt.h:
 *   vertices  float[3*nv]            triangle indices uint32[3*nt]
 *   spheres   float[4*ns]  (cx,cy,cz,r)
 *   quads     float[9*nq]  (center, ex, ez) -- the reference "plane" shape:
 *             the unit plane y=0, |x|,|z|<=0.5 mapped by ex (object x) and
 *             ez (object z); its normal is normalize(ez x ex)
 *   materials float[16*nm] (type, kd[3], ks[3], exponent, ior, transparency,
 *             Le[3], ior_g, ior_b, pad);  type 0 diffuse, 1 modified Phong,
 *             2 dielectric; ior_g / ior_b 0 = ior
 * Global primitive ids: triangles [0,nt), spheres [nt,nt+ns), quads after.
 */
#ifndef ORACLE_PT_H_
#define ORACLE_PT_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MISS 0xFFFFFFFFu

/* Philox4x32-10 block function (counter, key) -> 4 words. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct orc_scene orc_scene;

/* Builds the two-level scene (top kd-tree over shapes, a kd-tree over the
 * mesh triangles) with the reference build parameters.  kd_min_prims /
 * kd_max_depth of 0 select the reference defaults (3, 24).  use_naive != 0
 * replaces both kd-trees by NaiveSpatialAccel.  Returns NULL on bad input. */
int orc_interface_version(void);   /* 13 */
orc_scene *orc_scene_create(uint32_t nv, const float *vertices,
                            uint32_t nt, const uint32_t *indices, const uint32_t *tri_material,
                            uint32_t ns, const float *spheres, const uint32_t *sphere_material,
                            uint32_t nq, const float *quads, const uint32_t *quad_material,
                            uint32_t nm, const float *materials,
                            const float env_radiance[3],
                            const uint32_t *tri_flags,        /* optional: bit 0 = one-sided triangle */
                            const float *vertex_normals,      /* optional, with the mask */
                            const uint8_t *vertex_normal_mask,
                            uint32_t kd_min_prims, uint32_t kd_max_depth,
                            int use_naive, int nthreads);
void orc_scene_destroy(orc_scene *s);

/* kd statistics of the mesh tree: [internal, leaves, max_depth, refs,
 * max_leaf_prims, invariant_violations (prims overlapping a leaf cell but not
 * referenced by it; checked only when check != 0)] */
void orc_kd_stats(const orc_scene *s, int check, uint64_t out[6]);

/* Closest hit for n rays (rays = double[6*n]: origin, direction (unit)).
 * t_out = INFINITY and id_out = ORC_MISS on a miss. */
void orc_trace_closest(const orc_scene *s, uint64_t n, const double *rays,
                       double *t_out, uint32_t *id_out, int nthreads);

/* t of one primitive (global id) for n rays, INFINITY where it misses or
 * t <= EPSILON: used to classify closest-hit disagreements as exact ties. */
void orc_prim_t(const orc_scene *s, uint64_t n, const double *rays, const uint32_t *prim, double *t_out);

/* Any hit in (eps, tmax) following Scene::intersects (tmax - EPSILON). */
void orc_trace_any(const orc_scene *s, uint64_t n, const double *rays,
                   const double *tmax, uint8_t *occluded, int nthreads);

/* Camera: eye[3], look[3], up[3], heightAngle (degrees), aspect, then the
 * thin lens: aperture (m, 0 = pinhole), focal length (m), focal distance (m),
 * focal point x, y ([0,1]^2), autofocus (nonzero: the focal distance is the
 * hit distance of the ray through the focal point, cameras/ThinLensCamera.cpp:
 * 97-118, when the scene is known; the given distance otherwise). */
typedef struct orc_job {
  double camera[17];
  uint32_t width, height, spp, max_depth;
  uint32_t seed, job_id, pix_begin, pix_end;
  uint32_t direct_samples; /* light samples per light per vertex (0 = 1) */
  uint32_t filter;         /* film: 0 pixel bin, else 1 + filter type (0 box ... 4 lanczosSinc) */
  uint32_t filter_support; /* 0 = 2 */
  double filter_params[2]; /* gaussian sigma; mitchell B, C; lanczosSinc tau (0 = defaults) */
  uint32_t direct_off;     /* 1: efficientDirect = false (PathTracer.cpp:50-52, 70-73) */
  uint32_t light_select;   /* 1: each light sample draws one light by emitted power
                              (EmitterSampler's draw) and divides by its probability */
  uint32_t renderer;       /* 0 path tracer (PathTracer.cpp), 1 bidirectional path tracer
                              (BidirectionalPathTracer.cpp:34-174); max_depth caps each subpath */
} orc_job;

/* Renders pixels [pix_begin, pix_end) of the job.  sum/sumsq are
 * double[3*(pix_end-pix_begin)] (sum over the spp samples of L and L^2 per
 * channel); counters (optional, uint64[4]) = camera rays, extension rays
 * (camera included), shadow rays, paths. */
void orc_render(const orc_scene *s, const orc_job *job, double *sum, double *sumsq,
                uint64_t *counters, int nthreads);

/* Per-path radiance for explicit (pixel, sample) pairs: L_out double[3*n]. */
void orc_render_paths(const orc_scene *s, const orc_job *job, uint64_t n,
                      const uint32_t *pixel, const uint32_t *sample, double *L_out,
                      int nthreads);

/* The job's film as ReconstructionRenderOutput builds it
 * (renderers/outputs/ReconstructionRenderOutput.cpp:30-83, RenderOutput.cpp:
 * 72-101): rgbw_out double[4*W*H] = numerator RGB and denominator per pixel,
 * every sample splatted (filter) or binned (no filter). */
void orc_render_film(const orc_scene *s, const orc_job *job, double *rgbw_out, int nthreads);

/* Camera ray of (pixel, sample): ray_out double[6].  No scene: a thin lens
 * focuses at the given focal distance. */
void orc_camera_ray(const orc_job *job, uint32_t pixel, uint32_t sample, double *ray_out);

/* BSDF sampling of one material at a hit, for distribution tests:
 * normal n[3] (unit), incident wi[3]; u[4] uniforms; ior_index in {0,1,2}.
 * Outputs wo[3] (zero if absorbed), fs[3] = f/pdf, pdf, specular flag. */
void orc_bsdf_sample(const float *material, const double *n, const double *wi,
                     const double *u, uint32_t ior_index, double *wo, double *fs,
                     double *pdf, int *specular);
/* BSDF evaluate f(wi, wo) and pdf of a reflected direction (projected measure). */
void orc_bsdf_eval(const float *material, const double *n, const double *wi,
                   const double *wo, double *f, double *pdf);
/* Fresnel::getFr(wi, N, n1, n2). */
double orc_fresnel(const double *wi, const double *n, double n1, double n2);

/* Replaces the scene's constant background by a lat-long environment image
 * (materials/emitters/EnvironmentMap.cpp:24-95) looked up through a kernel
 * filter (filters/KernelFilter.cpp:124-176): rgb holds w*h*3 floats from
 * the top row; filter 0 box, 1 triangle, 2 gaussian (params[0] = sigma),
 * 3 mitchell (B, C), 4 lanczosSinc (tau); zero params and repeats take the
 * reference's defaults, support 0 means 2.  Returns 0, or -1 on bad
 * arguments. */
void orc_scene_set_points(orc_scene *s, uint32_t n, const float *points);
/* Emitting meshes and spheres (Scene::getLights over every emitting shape):
 * tri_shape (optional) names each triangle's shape, else triangles that
 * share an emitting material form one; sphere_frames (optional, 9 floats per
 * sphere, the images of the object axes) carry Sphere::getPosition's
 * object-space samples to the world, else 2r times the identity. */
void orc_scene_set_emitters(orc_scene *s, const uint32_t *tri_shape, const float *sphere_frames);
int orc_scene_set_env_map(orc_scene *s, uint32_t w, uint32_t h, const float *rgb, float repeat_u, float repeat_v,
                          int filter, uint32_t support, const float *params);
/* Background radiance seen by a ray leaving the scene in direction d. */
void orc_background(const orc_scene *s, const double *d, double *rgb);

/* Finalized 8-bit image of a film: rgbw holds n (numerator RGB, denominator)
 * quadruples; mode 0 Reinhard (renderers/tonemaps/ReinhardToneMap.cpp:25-98),
 * otherwise gamma only (NullToneMap.cpp:21-46); rgba gets n * 4 bytes
 * (r, g, b, 255). */
/* EmitterSampler::sample + ::getPd (stats/samplers/EmitterSampler.cpp:
 * 51-96) for samples 0..n-1 of `seed`: out double[9*n] = position (3),
 * normal (3), light index, pdf_area (1 / the lights' summed area), the
 * light's selection probability.  Returns 0, or -1 without lights. */
int orc_emitter_sample(const orc_scene *s, uint64_t n, uint32_t seed, double *out);

void orc_tonemap(const float *rgbw, uint64_t n, int mode, uint8_t *rgba);
/* Metropolis light transport (renderers/mlt/): `chains` Markov chains of
 * `mutations` mutations each, seeded from `seed_paths` generated paths of at
 * most `max_depth` vertices (mltNoInitialPaths, mltMaxDepth); max_rejections
 * = mltMaxConsequtiveRejections (0 = no limit); filter_proposed =
 * mltFilterProposed; probs = mltBidirPathMutationProb,
 * mltLensSubpathMutationProb, mltPerturbationPathMutationProb.  The job's
 * seed and job_id key the random streams; width x height is the film. */
typedef struct orc_mlt {
  uint32_t chains, mutations, seed_paths, max_depth, max_rejections, filter_proposed;
  double probs[3];
} orc_mlt;
/* The seed set: b (sum / entries), the number of entries, and up to cap
 * normalised weights and path lengths.  0, -1 for an unsupported scene or
 * camera, -2 when no seed path carries light. */
int orc_mlt_seeds(const orc_scene *s, const orc_job *job, const orc_mlt *p, double *b_out, uint64_t *n_out,
                  double *w_out, uint32_t *len_out, uint64_t cap, int nthreads);
/* The whole job: film_out double[3*W*H] finalised (RenderOutput's MLT
 * branch); optional num_out double[3*W*H] (numerators), proposed_out
 * uint64[W*H]; stats uint64[6] = samples added, accepted mutations,
 * mutations, extension rays, visibility rays, seed entries. */
int orc_mlt_render(const orc_scene *s, const orc_job *job, const orc_mlt *p, double *film_out, double *num_out,
                   uint64_t *proposed_out, uint64_t *stats, double *b_out, int nthreads);
/* Chains [c0, c1), their first `steps` mutations: rec double[6 per step] =
 * mutation (0 bidir, 1 lens subpath, 2 perturbation), acceptance
 * probability, accepted, the mutated path's length (0 when rejected as
 * invalid), the tentative sample's film bin, the chain's path length after. */
int orc_mlt_trace(const orc_scene *s, const orc_job *job, const orc_mlt *p, uint32_t c0, uint32_t c1,
                  uint32_t steps, double *rec, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
