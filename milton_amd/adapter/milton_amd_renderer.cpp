// Milton-side adapter of the MI355X path-tracing plugin: a
// milton::PointSampleRenderer that Milton's "dynamic" renderer loader
// (loaders/MiltonJSONSceneLoader.cpp:1770-1813) instantiates through this
// library's plugin_type / plugin_new, and that renders through
// milton_amd::PathTracer (include/milton_pt_plugin.hpp).  It is compiled
// inside a Milton build (with Milton's headers) and linked against
// libmilton_pt_plugin.so; the loader's _finalizeData gives it its output,
// scene and camera (:315-330), MiltonApp calls init() and RenderThread
// render() (gui/MiltonApp.cpp:218-222, gui/RenderThread.cpp:26-29).
//
// Its two converters follow the mapping of Milton's flattened scene to the
// C ABI's scene description:
//   flatten()      Scene -> mpt_scene_desc: Mesh triangles (world space, the
//                  mesh's own normals when it has them; an emitting mesh is
//                  one light), standalone Triangles (one-sided), Spheres (with
//                  their transforms), Planes (quads), PointShapes with omni
//                  emitters (point lights), and the materials' BSDF and
//                  emitter keys;
//   to_camera()    Pinhole / ThinLensCamera -> mpt_camera.
#include <milton.h>

#include <cmath>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#define MILTON_PT_PLUGIN_NO_ENTRY_POINTS
#include "milton_pt_plugin.hpp"

namespace milton_amd_adapter {

using namespace milton;

// Milton's scene flattened into the arrays an mpt_scene_desc points at.
struct Flat {
  std::vector<float> vertices, vertex_normals, spheres, quads, points, sphere_frames;
  std::vector<uint8_t> normal_mask;
  std::vector<uint32_t> indices, tri_material, tri_flags, tri_shape, sphere_material, quad_material;
  uint32_t shapes = 0;
  std::vector<mpt_material> materials;
  std::map<Material *, uint32_t> slot;   // Milton material -> index
  bool any_normals = false;
  std::string error;

  mpt_scene_desc desc() const {
    mpt_scene_desc d{};
    d.version = MPT_ABI_VERSION;
    d.num_vertices = (uint32_t)(vertices.size() / 3);
    d.vertices = vertices.data();
    d.num_triangles = (uint32_t)(indices.size() / 3);
    d.indices = indices.data();
    d.tri_material = tri_material.data();
    d.tri_flags = tri_flags.data();
    d.tri_shape = tri_shape.data();
    d.sphere_frames = sphere_frames.data();
    d.num_spheres = (uint32_t)(spheres.size() / 4);
    d.spheres = spheres.data();
    d.sphere_material = sphere_material.data();
    d.num_quads = (uint32_t)(quads.size() / 9);
    d.quads = quads.data();
    d.quad_material = quad_material.data();
    d.num_materials = (uint32_t)materials.size();
    d.materials = materials.data();
    if (any_normals) {
      d.vertex_normals = vertex_normals.data();
      d.vertex_normal_mask = normal_mask.data();
    }
    d.num_points = (uint32_t)(points.size() / 6);
    d.points = points.data();
    return d;
  }
};

double rgb_of(const SpectralSampleSet &s, unsigned c) { return s[c < s.getN() ? c : s.getN() - 1].value; }

// Material::getBSDF (materials/Material.cpp:67-92) and the BSDFs' keys:
// kd (DiffuseBSDF.cpp:45), kd / ks / n (ModifiedPhongBSDF.cpp:51-81),
// ks / ior / transparency (DielectricBSDF.cpp:48-53); "specular" and
// "transmissive" are dielectrics with transparency 0 and 1.  An oriented
// emitter of power P on a shape of area A has Le = P / (pi A)
// (Emitter.cpp:34-48, emitters/OrientedEmitter.cpp:17-23).
uint32_t material_of(Flat &f, Material *m, double emitter_area) {
  const auto it = f.slot.find(m);
  if (it != f.slot.end() && emitter_area <= 0) return it->second;
  mpt_material out{};
  out.exponent = 1;
  out.ior = 1.5f;
  out.transparency = 1;
  const std::string bsdf = m ? m->getValue<std::string>("bsdf", "diffuse") : "diffuse";
  const SpectralSampleSet kd =
      m ? m->getValue<SpectralSampleSet>("kd", SpectralSampleSet::fill(0.5)) : SpectralSampleSet::fill(0.5);
  const SpectralSampleSet ks =
      m ? m->getValue<SpectralSampleSet>("ks", SpectralSampleSet::fill(0.5)) : SpectralSampleSet::fill(0.5);
  for (unsigned c = 0; c < 3; ++c) {
    out.kd[c] = (float)rgb_of(kd, c);
    out.ks[c] = (float)rgb_of(ks, c);
  }
  if (bsdf == "diffuse") {
    out.type = MPT_BSDF_DIFFUSE;
  } else if (bsdf == "modifiedPhong" || bsdf == "phong") {
    out.type = MPT_BSDF_PHONG;
    out.exponent = (float)m->getValue<real_t>("n", 1);
  } else if (bsdf == "dielectric" || bsdf == "specular" || bsdf == "transmissive") {
    out.type = MPT_BSDF_DIELECTRIC;
    const SpectralSampleSet ior = m->getValue<SpectralSampleSet>("ior", SpectralSampleSet::fill(1.5));
    out.ior = (float)rgb_of(ior, 0);
    if (rgb_of(ior, 1) != rgb_of(ior, 0) || rgb_of(ior, 2) != rgb_of(ior, 0)) {
      out.ior_g = (float)rgb_of(ior, 1);
      out.ior_b = (float)rgb_of(ior, 2);
    }
    out.transparency = bsdf == "specular" ? 0.0f
                       : bsdf == "transmissive" ? 1.0f
                                                : (float)m->getValue<real_t>("transparency", 1);
  } else {
    f.error = "unsupported BSDF '" + bsdf + "'";
  }
  if (emitter_area > 0) {
    const SpectralSampleSet P = m->getValue<SpectralSampleSet>("power", SpectralSampleSet::black());
    for (unsigned c = 0; c < 3; ++c) out.Le[c] = (float)(rgb_of(P, c) / (M_PI * emitter_area));
  }
  f.materials.push_back(out);
  const uint32_t k = (uint32_t)f.materials.size() - 1;
  if (emitter_area <= 0) f.slot[m] = k;
  return k;
}

std::string emitter_type(Material *m) { return m ? m->getValue<std::string>("emitter", "null") : "null"; }

void push3(std::vector<float> &v, const Point3 &p) {
  for (int k = 0; k < 3; ++k) v.push_back((float)p[k]);
}
void push3(std::vector<float> &v, const Vector3 &p) {
  for (int k = 0; k < 3; ++k) v.push_back((float)p[k]);
}
Vector3 cross(const Vector3 &a, const Vector3 &b) {
  return Vector3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
double norm(const Vector3 &a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

void add_shape(Flat &f, Shape *s);

void add_mesh(Flat &f, Mesh *mesh) {
  const std::string et = emitter_type(mesh->getMaterial());
  if (et != "null" && et != "oriented") {
    f.error = "mesh emitters must be oriented";
    return;
  }
  const Matrix4x4 &M = mesh->getTransToWorld();
  unsigned nv = 0, nt = 0, nn = 0;
  const Vertex *v = mesh->getVertices(nv);
  const MeshTriangle *t = mesh->getTriangles(nt);
  const Normal *n = mesh->getNormals(nn);
  const uint32_t base = (uint32_t)(f.vertices.size() / 3);
  // an emitting mesh's radiance spreads its power over its world-space
  // area (Mesh::_getSurfaceArea, shapes/Mesh.cpp:452-466)
  double area = 0;
  if (et == "oriented")
    for (unsigned i = 0; i < nt; ++i) {
      const Point3 a = M * Point3(v[t[i].A][0], v[t[i].A][1], v[t[i].A][2]);
      const Point3 b = M * Point3(v[t[i].B][0], v[t[i].B][1], v[t[i].B][2]);
      const Point3 c = M * Point3(v[t[i].C][0], v[t[i].C][1], v[t[i].C][2]);
      const Vector3 e1(b[0] - a[0], b[1] - a[1], b[2] - a[2]), e2(c[0] - a[0], c[1] - a[1], c[2] - a[2]);
      area += 0.5 * norm(cross(e1, e2));
    }
  const uint32_t mat = material_of(f, mesh->getMaterial(), area);
  const uint32_t shape = f.shapes++;
  for (unsigned i = 0; i < nv; ++i) {
    push3(f.vertices, M * Point3(v[i][0], v[i][1], v[i][2]));
    // a mesh file's normals (Mesh::computeNormals fills the others,
    // shapes/Mesh.cpp:335-377); normals transform by the inverse transpose,
    // which for the similarity transforms of scene files is M up to scale
    const bool has = nn == nv;
    if (has) {
      const Vector3 w = M * n[i];
      const double l = norm(w);
      for (int k = 0; k < 3; ++k) f.vertex_normals.push_back((float)(l > 0 ? w[k] / l : 0));
      f.any_normals = true;
    } else {
      for (int k = 0; k < 3; ++k) f.vertex_normals.push_back(0);
    }
    f.normal_mask.push_back(has ? 1 : 0);
  }
  for (unsigned i = 0; i < nt; ++i) {
    f.indices.push_back(base + t[i].A);
    f.indices.push_back(base + t[i].B);
    f.indices.push_back(base + t[i].C);
    f.tri_material.push_back(mat);
    f.tri_flags.push_back(0);
    f.tri_shape.push_back(shape);
  }
}

void add_shape(Flat &f, Shape *s) {
  if (ShapeSet *set = dynamic_cast<ShapeSet *>(s)) {
    for (unsigned i = 0; i < set->size(); ++i) add_shape(f, (*set)[i]);
  } else if (Mesh *mesh = dynamic_cast<Mesh *>(s)) {
    add_mesh(f, mesh);
  } else if (Triangle *tri = dynamic_cast<Triangle *>(s)) {
    // a standalone Triangle culls back faces (shapes/Triangle.cpp:245-270)
    if (emitter_type(tri->getMaterial()) != "null") {
      f.error = "emitting triangles are not rendered by the device path";
      return;
    }
    const Matrix4x4 &M = tri->getTransToWorld();
    const uint32_t base = (uint32_t)(f.vertices.size() / 3);
    for (int k = 0; k < 3; ++k) {
      push3(f.vertices, M * Point3(tri->vertices[k][0], tri->vertices[k][1], tri->vertices[k][2]));
      for (int c = 0; c < 3; ++c) f.vertex_normals.push_back(0);
      f.normal_mask.push_back(0);
      f.indices.push_back(base + k);
    }
    f.tri_material.push_back(material_of(f, tri->getMaterial(), 0));
    f.tri_flags.push_back(MPT_TRI_ONE_SIDED);
    f.tri_shape.push_back(f.shapes++);
  } else if (Sphere *sph = dynamic_cast<Sphere *>(s)) {
    // the unit sphere of radius 0.5 (shapes/Sphere.cpp:16-45) under its
    // transform, whose linear part also places an emitting sphere's light
    // samples (Sphere::getPosition, :117-124); its area is pi whatever the
    // scale (Sphere::_getSurfaceArea, :138-141)
    const std::string et = emitter_type(sph->getMaterial());
    if (et != "null" && et != "oriented") {
      f.error = "sphere emitters must be oriented";
      return;
    }
    const Matrix4x4 &M = sph->getTransToWorld();
    push3(f.spheres, M * Point3(0, 0, 0));
    f.spheres.push_back((float)(0.5 * norm(M * Vector3(1, 0, 0))));
    for (int k = 0; k < 3; ++k) push3(f.sphere_frames, M * Vector3(k == 0, k == 1, k == 2));
    f.sphere_material.push_back(material_of(f, sph->getMaterial(), et == "oriented" ? M_PI : 0));
  } else if (Plane *pl = dynamic_cast<Plane *>(s)) {
    // the unit square y = 0 (shapes/Plane.cpp:25-106): centre and the images
    // of its x and z axes; normal ez x ex
    const Matrix4x4 &M = pl->getTransToWorld();
    Vector3 ex = M * Vector3(1, 0, 0), ez = M * Vector3(0, 0, 1);
    // a mirroring transform keeps the inverse-transpose normal: swap the axes
    const Vector3 ey = M * Vector3(0, 1, 0);
    const Vector3 c = cross(ex, ey);
    if (c[0] * ez[0] + c[1] * ez[1] + c[2] * ez[2] < 0) std::swap(ex, ez);
    push3(f.quads, M * Point3(0, 0, 0));
    push3(f.quads, ex);
    push3(f.quads, ez);
    const std::string et = emitter_type(pl->getMaterial());
    if (et != "null" && et != "oriented") {
      f.error = "plane emitters must be oriented";
      return;
    }
    f.quad_material.push_back(material_of(f, pl->getMaterial(), et == "oriented" ? norm(cross(ex, ez)) : 0));
  } else if (PointShape *pt = dynamic_cast<PointShape *>(s)) {
    // an omni emitter of power P at the transformed origin
    // (emitters/OmniEmitter.cpp; DirectIllumination.cpp:96-121)
    Material *m = pt->getMaterial();
    const std::string et = emitter_type(m);
    if (et != "omni" && et != "point") return;   // a point that does not emit adds nothing
    push3(f.points, pt->getTransToWorld() * Point3(0, 0, 0));
    const SpectralSampleSet P = m->getValue<SpectralSampleSet>("power", SpectralSampleSet::black());
    for (unsigned c = 0; c < 3; ++c) f.points.push_back((float)rgb_of(P, c));
  } else {
    f.error = "a shape type the device path does not render";
  }
}

// Scene -> flattened arrays.  The background stays black here; a constant
// or image environment goes through mpt_scene_desc.env_radiance / env_map.
bool flatten(Scene *scene, Flat &f) {
  if (!scene) {
    f.error = "no scene";
    return false;
  }
  add_shape(f, scene->getShapes());
  return f.error.empty();
}

// ThinLensCamera::init's keys (cameras/ThinLensCamera.cpp:62-118): eye,
// look, up, heightAngle, aspectRatio; aperture in mm, fstop, focalDistance
// and focalPoint for a thin lens; a PinholeCamera has aperture 0
// (PinholeCamera.cpp:21-25).
mpt_camera to_camera(Camera *camera) {
  mpt_camera c{};
  c.look[2] = -1;
  c.up[1] = 1;
  c.height_angle = 45;
  c.aspect = 1;
  ThinLensCamera *tl = dynamic_cast<ThinLensCamera *>(camera);
  if (!tl) return c;
  const Point3 e = tl->getEye();
  const Vector3 l = tl->getLook(), u = tl->getUp();
  for (int k = 0; k < 3; ++k) {
    c.eye[k] = (float)e[k];
    c.look[k] = (float)l[k];
    c.up[k] = (float)u[k];
  }
  c.height_angle = (float)tl->getValue<real_t>("heightAngle", 45);
  c.aspect = (float)tl->getValue<real_t>("aspectRatio", 1);
  if (!dynamic_cast<PinholeCamera *>(camera)) {
    c.aperture = (float)(tl->getValue<real_t>("aperture", 35.0) / 1000.0);
    c.focal_length = (float)(tl->getValue<real_t>("fstop", 14) * c.aperture);
    c.focal_distance = (float)tl->getValue<real_t>("focalDistance", 1);
    c.autofocus = tl->contains("focalDistance") ? 0u : 1u;
    c.focal_point[0] = c.focal_point[1] = 0.5f;
  }
  return c;
}

class AmdPathTracer : public PointSampleRenderer {
 public:
  AmdPathTracer() : pt_(new milton_amd::PathTracer()) {}
  virtual ~AmdPathTracer() { delete pt_; }

  // Renderer::init (renderers/Renderer.cpp:28-54) without the CPU kd build:
  // the renderer's keys go to the facade, the scene and camera through the
  // converters, the film size from the output.
  virtual void init() {
    ok_ = false;
    static const char *kKeys[] = {"noSuperSamples", "maxDepth", "seed", "noDirectSamples", "kdMinPrimitives",
                                  "kdMaxDepth", "kdCostTraversal", "kdEmptyBias", "kdNoThreads", "device",
                                  "batchPaths", "bidirectional", "mlt", "noRenderThreads", "mltNoInitialPaths",
                                  "mltMaxDepth", "mltMaxConsequtiveRejections", "mltBidirPathMutationProb",
                                  "mltLensSubpathMutationProb", "mltPerturbationPathMutationProb", "mltMutations"};
    for (const char *k : kKeys)
      if (contains(k)) pt_->setValue(k, getValue<real_t>(k, 0));
    if (contains("efficientDirect")) pt_->setValue("efficientDirect", getValue<bool>("efficientDirect", true));
    if (contains("mltFilterProposed")) pt_->setValue("mltFilterProposed", getValue<bool>("mltFilterProposed", true));
    if (!flatten(m_scene, flat_)) {
      std::cerr << "AmdPathTracer: " << flat_.error << std::endl;
      return;
    }
    pt_->setScene(flat_.desc());
    pt_->setCamera(to_camera(m_camera));
    if (m_output) pt_->setOutput(m_output->getWidth(), m_output->getHeight());
    const int rc = pt_->init();
    if (rc) {
      std::cerr << "AmdPathTracer: " << pt_->lastError() << std::endl;
      return;
    }
    ok_ = true;
  }

  // PointSampleRenderer::render (renderers/PointSampleRenderer.cpp:24-81):
  // the device renders the whole job; each pixel's value (numerator /
  // denominator of the pixel-bin film) goes to the output as one sample at
  // the pixel's centre, then the output is finalized.
  virtual void render() {
    if (!ok_ || !m_output) return;
    if (pt_->render()) {
      std::cerr << "AmdPathTracer: " << pt_->lastError() << std::endl;
      return;
    }
    const std::vector<float> &rgb = pt_->getOutput();
    const unsigned W = m_output->getWidth(), H = m_output->getHeight();
    for (unsigned y = 0; y < H; ++y)
      for (unsigned x = 0; x < W; ++x) {
        PointSample s;
        s.position[0] = (x + 0.5) / W;
        s.position[1] = (y + 0.5) / H;
        s.value = SpectralSampleSet::black();
        for (unsigned c = 0; c < 3; ++c) s.value[c].value = rgb[3ull * (y * W + x) + c];
        m_output->addSample(s);
      }
    m_output->finalize();
  }

  milton_amd::PathTracer *facade() { return pt_; }
  const Flat &flat() const { return flat_; }
  bool ok() const { return ok_; }

 private:
  milton_amd::PathTracer *pt_;
  Flat flat_;
  bool ok_ = false;
};

}  // namespace milton_amd_adapter

// dynamic/DynamicPlugin.h:34-62: the three symbols the loader binds.
extern "C" MILTON_PLUGIN_TYPE plugin_type() { return MILTON_PLUGIN_TYPE_RENDERER; }
extern "C" void *plugin_new() {
  return static_cast<milton::Renderer *>(new milton_amd_adapter::AmdPathTracer());
}
extern "C" void plugin_delete(void *p) { delete static_cast<milton::Renderer *>(p); }
