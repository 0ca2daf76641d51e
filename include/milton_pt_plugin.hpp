/*
 * milton_pt_plugin.hpp -- the Milton "dynamic" renderer plugin of the
 * MI355X path-tracing core.
 *
 * Milton loads a renderer plugin from a scenefile entry
 *   "renderer": { "type": "dynamic", "path": "libmilton_pt_plugin.so", ... }
 * through _parse_dynamic_plugin<Renderer, MILTON_PLUGIN_TYPE_RENDERER>
 * (loaders/MiltonJSONSceneLoader.cpp:1770-1813): dlopen, then the C symbols
 * plugin_type / plugin_new / plugin_delete (dynamic/DynamicPlugin.h:39-58).
 * The plugin library exports exactly those three symbols; plugin_new returns
 * a milton_amd::PathTracer, a facade with the renderer lifecycle the loader
 * and the application drive (setScene / setCamera / setOutput, the
 * PropertyMap keys of the path, init(), render()).  A Milton build wraps it
 * in its own milton::Renderer subclass (the reference headers never ship
 * with this library).  As with Milton's own plugins, the host reaches the
 * object only through its virtual functions, so it never links the plugin.
 */
#ifndef MILTON_PT_PLUGIN_HPP_
#define MILTON_PT_PLUGIN_HPP_

#include <map>
#include <string>
#include <vector>

#include "milton_pt.h"

/* MILTON_PLUGIN_TYPE_RENDERER (dynamic/DynamicPlugin.h:39-50). */
enum { MPT_PLUGIN_TYPE_RENDERER = 0 };

/* The plugin library's entry points.  A Milton-side adapter that defines its
 * own plugin_type / plugin_new / plugin_delete (with Milton's
 * MILTON_PLUGIN_TYPE return type) defines MILTON_PT_PLUGIN_NO_ENTRY_POINTS
 * before including this header and constructs milton_amd::PathTracer
 * directly. */
#ifndef MILTON_PT_PLUGIN_NO_ENTRY_POINTS
extern "C" {
int plugin_type(void);
void *plugin_new(void);
void plugin_delete(void *plugin);
}
#endif

namespace milton_amd {

class PathTracer {
 public:
  PathTracer();
  virtual ~PathTracer();

  /* PropertyMap keys read on the path (numbers only):
   *   renderer  "noSuperSamples" (spp, SuperSampleGenerator.cpp:26-28),
   *             "maxDepth" (the cap), "seed" (-seed=, MiltonApp.cpp:329),
   *             "efficientDirect" (PathTracer.cpp:23),
   *             "noDirectSamples" (1..64, DirectIllumination.cpp:33),
   *             "device" (HIP device), "batchPaths" (queue depth),
   *             "bidirectional" (1: BidirectionalPathTracer instead,
   *             MiltonJSONSceneLoader.cpp:720-721), "mlt" (1: MLTRenderer,
   *             :731-738, with "noRenderThreads" (chains),
   *             "mltNoInitialPaths", "mltMaxDepth",
   *             "mltMaxConsequtiveRejections", "mltBidirPathMutationProb",
   *             "mltLensSubpathMutationProb",
   *             "mltPerturbationPathMutationProb", "mltFilterProposed" and
   *             "mltMutations" per chain, by default noSuperSamples x
   *             pixels / chains: the reference runs until stopped)
   *   output    "width" / "height" (RenderOutput.cpp:34)
   *   kd-tree   "kdMinPrimitives" (leaf size), "kdMaxDepth", "kdCostTraversal",
   *             "kdEmptyBias", "kdNoThreads" (kdTreeAccel.h:97-104)       */
  virtual void setValue(const std::string &key, double value);
  virtual double getValue(const std::string &key, double fallback) const;

  /* Renderer::setScene / setCamera / setOutput (Renderer.h:26-62).  The
   * scene arrays are copied. */
  virtual void setScene(const mpt_scene_desc &scene);
  virtual void setCamera(const mpt_camera &camera);
  virtual void setOutput(unsigned width, unsigned height);

  /* Renderer::init (Renderer.cpp:28-54): opens the device, uploads the
   * scene and builds its kd-tree.  Returns an MPT_* status. */
  virtual int init();
  /* PointSampleRenderer::render (PointSampleRenderer.cpp:24-81): renders the
   * job and keeps the film (numerator / denominator per pixel, RGB). */
  virtual int render();
  const std::vector<float> &getOutput() const { return film_; }
  const mpt_render_stats &getStats() const { return stats_; }
  virtual std::string lastError() const;

 private:
  std::map<std::string, double> props_;
  std::vector<float> verts_, spheres_, quads_, film_;
  std::vector<uint32_t> idx_, tri_mat_, sph_mat_, quad_mat_;
  std::vector<mpt_material> mats_;
  std::vector<uint32_t> tri_flags_;
  std::vector<float> vnormals_;
  std::vector<uint8_t> vmask_;
  std::vector<float> env_rgb_;   // the environment image, when the scene has one
  std::vector<float> points_;    // point lights
  std::vector<uint32_t> tri_shape_;
  std::vector<float> sphere_frames_;
  mpt_env_map env_{};
  mpt_scene_desc desc_{};
  mpt_camera camera_{};
  unsigned width_ = 480, height_ = 480;
  mpt_context *ctx_ = nullptr;
  mpt_scene *scene_ = nullptr;
  mpt_render_stats stats_{};
  std::string err_;
};

}  // namespace milton_amd

#endif
