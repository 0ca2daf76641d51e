/*
 * milton_pt_scenes.h -- seeded synthetic workloads for the path-tracing core.
 *
 * No reference counterpart: the reference reads scenes from JSON scenefiles
 * and OBJ/PLY meshes (loaders/MiltonJSONSceneLoader.cpp, loaders/MeshLoader.cpp)
 * and ships no benchmark scenes.  These generators build the three benchmark
 * scenes directly in the mpt_scene_desc layout, with float-representable
 * vertices so the device path and the double-precision oracle see identical
 * geometry.  Callers: tests, the benchmark, the serving loop.
 */
#ifndef MILTON_PT_SCENES_H_
#define MILTON_PT_SCENES_H_

#include "milton_pt.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MPT_SCENE_ROOM = 1,    /* 1x1x1 room of 5 quads (10 tris) + two boxes (24 tris) + a 0.25^2 ceiling light */
  MPT_SCENE_SOUP = 2,    /* `count` random triangles in the unit cube, lognormal edge lengths */
  MPT_SCENE_TERRAIN = 3  /* `count`^2-vertex value-noise heightfield + 500 dielectric spheres + distant light */
};

typedef struct mpt_gen mpt_gen;

/* count: SOUP triangle count (default 1,000,000); TERRAIN grid side in
 * vertices (default 2048); ignored for ROOM.  0 selects the default. */
int mpt_gen_create(uint32_t kind, uint32_t count, uint32_t seed, mpt_gen **gen);
/* Fills desc with pointers into the generator's arrays (valid until destroy)
 * and camera with the scene's default camera (aspect from the benchmark
 * resolution). */
int mpt_gen_describe(const mpt_gen *gen, mpt_scene_desc *desc, mpt_camera *camera);
int mpt_gen_destroy(mpt_gen *gen);

/* Seeded ray sets for traversal parity over a scene's bounding cube [0,1]^3:
 * rays [0, n/2) are camera rays of the scene camera through a w x h film,
 * the rest have origins uniform in the cube and directions uniform on S^2. */
int mpt_gen_rays(const mpt_gen *gen, uint64_t n, uint32_t seed, uint32_t film_w, uint32_t film_h,
                 mpt_ray *rays);

#ifdef __cplusplus
}
#endif
#endif
