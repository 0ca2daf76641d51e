// Loads the renderer plugin the way Milton's scenefile loader does
// (loaders/MiltonJSONSceneLoader.cpp:1770-1813: dlopen with RTLD_LAZY,
// plugin_type / plugin_new, type check) and renders the room scene through
// the PathTracer facade.  With --load-only it stops after the type check and
// plugin_delete (no device needed).  With a device it also renders the same
// job through the C ABI directly and requires bit-identical films.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "milton_pt_plugin.hpp"
#include "milton_pt_scenes.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <plugin.so> [--load-only]\n", argv[0]);
    return 2;
  }
  const bool load_only = argc > 2 && std::strcmp(argv[2], "--load-only") == 0;
  void *lib = dlopen(argv[1], RTLD_LAZY);
  if (!lib) {
    std::fprintf(stderr, "plugin failed to load: %s\n", dlerror());
    return 1;
  }
  typedef int (*type_fn)();
  typedef void *(*new_fn)();
  typedef void (*delete_fn)(void *);
  type_fn ptype = (type_fn)dlsym(lib, "plugin_type");
  new_fn pnew = (new_fn)dlsym(lib, "plugin_new");
  delete_fn pdel = (delete_fn)dlsym(lib, "plugin_delete");
  if (!ptype || !pnew || !pdel) {
    std::fprintf(stderr, "invalid plugin: missing symbol\n");
    return 1;
  }
  if (ptype() != MPT_PLUGIN_TYPE_RENDERER) {
    std::fprintf(stderr, "dynamic plugin expected type 0 but found %d\n", ptype());
    return 1;
  }
  milton_amd::PathTracer *pt = static_cast<milton_amd::PathTracer *>(pnew());
  if (!pt) return 1;
  if (load_only) {
    pdel(pt);
    std::printf("plugin ok: type %d\n", ptype());
    return 0;
  }
  mpt_gen *gen = nullptr;
  if (mpt_gen_create(MPT_SCENE_ROOM, 0, 1, &gen)) return 1;
  mpt_scene_desc desc;
  mpt_camera cam;
  mpt_gen_describe(gen, &desc, &cam);
  pt->setScene(desc);
  pt->setCamera(cam);
  pt->setOutput(64, 48);
  pt->setValue("noSuperSamples", 16);
  pt->setValue("maxDepth", 5);
  pt->setValue("seed", 7);
  int rc = pt->init();
  if (!rc) rc = pt->render();
  if (rc) {
    std::fprintf(stderr, "plugin render failed (%d): %s\n", rc, pt->lastError().c_str());
    return 1;
  }
  const std::vector<float> film = pt->getOutput();
  // the same job through the C ABI
  mpt_context *ctx = nullptr;
  mpt_scene *scene = nullptr;
  if (mpt_open(0, &ctx) || mpt_scene_create(ctx, &desc, nullptr, &scene)) return 1;
  mpt_job job{};
  job.camera = cam;
  job.width = 64;
  job.height = 48;
  job.spp = 16;
  job.max_depth = 5;
  job.seed = 7;
  std::vector<float> ref(3 * 64 * 48);
  if (mpt_render(scene, &job, ref.data(), nullptr)) return 1;
  double mean = 0;
  for (float v : film) mean += v;
  mean /= film.size();
  const bool same = std::memcmp(film.data(), ref.data(), 4 * ref.size()) == 0;
  std::printf("plugin render: mean %.6f, paths %llu, identical to C ABI: %s\n", mean,
              (unsigned long long)pt->getStats().paths, same ? "yes" : "no");
  // Metropolis light transport through the plugin's "mlt" key, against the
  // same job through the C ABI
  bool mlt_same = false;
  {
    milton_amd::PathTracer *mp = static_cast<milton_amd::PathTracer *>(pnew());
    mp->setScene(desc);
    mp->setCamera(cam);
    mp->setOutput(32, 24);
    mp->setValue("mlt", 1);
    mp->setValue("noRenderThreads", 64);
    mp->setValue("mltMutations", 16);
    mp->setValue("mltNoInitialPaths", 500);
    mp->setValue("seed", 7);
    int mrc = mp->init();
    if (!mrc) mrc = mp->render();
    mpt_job mj{};
    mj.camera = cam;
    mj.width = 32;
    mj.height = 24;
    mj.spp = 4;
    mj.max_depth = 8;
    mj.seed = 7;
    mj.renderer = MPT_RENDERER_MLT;
    mj.mlt_chains = 64;
    mj.mlt_mutations = 16;
    mj.mlt_seed_paths = 500;
    mj.mlt_max_depth = 10;
    mj.mlt_max_rejections = 500;
    mj.mlt_probs[0] = mj.mlt_probs[1] = mj.mlt_probs[2] = 1.0f;
    std::vector<float> mref(3 * 32 * 24);
    if (!mrc && !mpt_render(scene, &mj, mref.data(), nullptr)) {
      const std::vector<float> mf = mp->getOutput();
      mlt_same = mf.size() == mref.size() && std::memcmp(mf.data(), mref.data(), 4 * mref.size()) == 0;
    } else {
      std::fprintf(stderr, "plugin MLT render failed (%d): %s\n", mrc, mp->lastError().c_str());
    }
    std::printf("MLT identical to C ABI: %s\n", mlt_same ? "yes" : "no");
    pdel(mp);
  }
  // a scene description with a missing array is rejected at init
  milton_amd::PathTracer *bad = static_cast<milton_amd::PathTracer *>(pnew());
  mpt_scene_desc broken = desc;
  broken.indices = nullptr;
  bad->setScene(broken);
  bad->setCamera(cam);
  const bool rejected = broken.num_triangles > 0 && bad->init() == MPT_E_INVALID;
  std::printf("missing array rejected: %s (%s)\n", rejected ? "yes" : "no", bad->lastError().c_str());
  pdel(bad);
  mpt_scene_destroy(scene);
  mpt_close(ctx);
  pdel(pt);
  mpt_gen_destroy(gen);
  return same && mlt_same && rejected && std::isfinite(mean) && mean > 0 ? 0 : 1;
}
