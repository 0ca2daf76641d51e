// Test driver of the Milton-side adapter (milton_amd/adapter/
// milton_amd_renderer.cpp), compiled with the interface subset in
// tests/adapter/milton.h.  It builds the benchmark room as Milton objects
// (one Mesh per run of triangles sharing a material, the ceiling light as an
// oriented-emitter Plane, a PinholeCamera), hands them to the renderer that
// plugin_new returns, as Milton's loader and application do, and checks:
//   --cpu  the flattened scene equals the room's scene description (same
//          triangles, materials and light) and traces the same closest hits
//          on the host (mpt_kd_build / mpt_kd_trace); init() fails cleanly
//          when no device is present;
//   --gpu  init() and render() succeed and the output the adapter writes is
//          bit-identical to mpt_render of the room's scene description.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../milton_amd/adapter/milton_amd_renderer.cpp"
#include "milton_pt_scenes.h"

using namespace milton;

namespace {

int fail(const char *what) {
  std::fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}

struct MiltonRoom {
  std::vector<Material *> mats;
  std::vector<Shape *> shapes;
  Scene scene;
  PinholeCamera camera;
  ~MiltonRoom() {
    for (Shape *s : shapes) delete s;
    for (Material *m : mats) delete m;
  }
};

SpectralSampleSet rgb(const float *v) {
  SpectralSampleSet s;
  for (int c = 0; c < 3; ++c) s[c].value = v[c];
  return s;
}

// The scene description as Milton objects.
void build(const mpt_scene_desc &d, const mpt_camera &cam, MiltonRoom &r) {
  std::vector<Material *> by_id(d.num_materials, nullptr);
  for (uint32_t i = 0; i < d.num_materials; ++i) {
    const mpt_material &m = d.materials[i];
    if (m.Le[0] > 0 || m.Le[1] > 0 || m.Le[2] > 0) continue;   // emitters are made per shape
    Material *mm = new Material();
    (*mm)["bsdf"] = std::string(m.type == MPT_BSDF_DIFFUSE ? "diffuse" : m.type == MPT_BSDF_PHONG ? "modifiedPhong"
                                                                                                 : "dielectric");
    (*mm)["kd"] = rgb(m.kd);
    (*mm)["ks"] = rgb(m.ks);
    (*mm)["n"] = (real_t)m.exponent;
    (*mm)["transparency"] = (real_t)m.transparency;
    (*mm)["ior"] = SpectralSampleSet::fill(m.ior);
    r.mats.push_back(mm);
    by_id[i] = mm;
  }
  for (uint32_t t0 = 0; t0 < d.num_triangles;) {
    uint32_t t1 = t0;
    while (t1 < d.num_triangles && d.tri_material[t1] == d.tri_material[t0]) ++t1;
    Mesh *mesh = new Mesh();
    mesh->setMaterial(by_id[d.tri_material[t0]]);
    std::vector<int> remap(d.num_vertices, -1);
    for (uint32_t t = t0; t < t1; ++t) {
      unsigned abc[3];
      for (int k = 0; k < 3; ++k) {
        const uint32_t v = d.indices[3 * t + k];
        if (remap[v] < 0) {
          remap[v] = (int)mesh->vertices.size();
          mesh->vertices.push_back(Vertex(d.vertices[3 * v], d.vertices[3 * v + 1], d.vertices[3 * v + 2]));
        }
        abc[k] = (unsigned)remap[v];
      }
      mesh->triangles.push_back(MeshTriangle{abc[0], abc[1], abc[2]});
    }
    r.shapes.push_back(mesh);
    r.scene.getShapes()->push_back(mesh);
    t0 = t1;
  }
  for (uint32_t q = 0; q < d.num_quads; ++q) {
    const float *Q = d.quads + 9 * q;
    const mpt_material &m = d.materials[d.quad_material[q]];
    Plane *pl = new Plane();
    const double ex[3] = {Q[3], Q[4], Q[5]}, ez[3] = {Q[6], Q[7], Q[8]};
    const double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
    Matrix4x4 M;
    for (int i = 0; i < 3; ++i) {
      M.m[4 * i] = ex[i];
      M.m[4 * i + 1] = ey[i];
      M.m[4 * i + 2] = ez[i];
      M.m[4 * i + 3] = Q[i];
    }
    pl->setTransToWorld(M);
    Material *mm = new Material();
    (*mm)["bsdf"] = std::string("diffuse");
    (*mm)["kd"] = rgb(m.kd);
    if (m.Le[0] > 0 || m.Le[1] > 0 || m.Le[2] > 0) {
      // power = Le pi A (Emitter.cpp:34-48 inverted)
      const double A = std::sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
      SpectralSampleSet P;
      for (int c = 0; c < 3; ++c) P[c].value = m.Le[c] * M_PI * A;
      (*mm)["emitter"] = std::string("oriented");
      (*mm)["power"] = P;
    }
    r.mats.push_back(mm);
    pl->setMaterial(mm);
    r.shapes.push_back(pl);
    r.scene.getShapes()->push_back(pl);
  }
  r.camera.setOrientation(Point3(cam.eye[0], cam.eye[1], cam.eye[2]), Vector3(cam.look[0], cam.look[1], cam.look[2]),
                          Vector3(cam.up[0], cam.up[1], cam.up[2]));
  r.camera["heightAngle"] = (real_t)cam.height_angle;
  r.camera["aspectRatio"] = (real_t)cam.aspect;
}

bool same_material(const mpt_material &a, const mpt_material &b) {
  return a.type == b.type && !std::memcmp(a.kd, b.kd, sizeof(a.kd)) && !std::memcmp(a.Le, b.Le, sizeof(a.Le)) &&
         (a.type == MPT_BSDF_DIFFUSE || !std::memcmp(a.ks, b.ks, sizeof(a.ks)));
}

}  // namespace

int main(int argc, char **argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
  mpt_gen *gen = nullptr;
  if (mpt_gen_create(MPT_SCENE_ROOM, 0, 1, &gen)) return fail("mpt_gen_create");
  mpt_scene_desc room;
  mpt_camera cam;
  mpt_gen_describe(gen, &room, &cam);
  MiltonRoom r;
  build(room, cam, r);

  // As Milton's loader and application drive a dynamic renderer.
  if (plugin_type() != MILTON_PLUGIN_TYPE_RENDERER) return fail("plugin_type");
  Renderer *ren = static_cast<Renderer *>(plugin_new());
  PointSampleRenderer *psr = dynamic_cast<PointSampleRenderer *>(ren);
  auto *amd = dynamic_cast<milton_amd_adapter::AmdPathTracer *>(ren);
  if (!psr || !amd) return fail("plugin_new does not return the adapter");
  const unsigned W = 48, H = 40;
  RenderOutput out(W, H);
  psr->setOutput(&out);
  ren->setScene(&r.scene);
  ren->setCamera(&r.camera);
  (*ren)["noSuperSamples"] = (real_t)16;
  (*ren)["maxDepth"] = (real_t)5;
  (*ren)["seed"] = (real_t)7;
  (*ren)["noDirectSamples"] = (real_t)4;
  ren->init();

  // The flattened scene against the room's own description.
  const milton_amd_adapter::Flat &f = amd->flat();
  if (!f.error.empty()) return fail(f.error.c_str());
  const mpt_scene_desc d = f.desc();
  if (d.num_triangles != room.num_triangles || d.num_quads != room.num_quads || d.num_spheres != 0)
    return fail("primitive counts");
  for (uint32_t t = 0; t < d.num_triangles; ++t) {
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c)
        if (d.vertices[3 * d.indices[3 * t + k] + c] != room.vertices[3 * room.indices[3 * t + k] + c])
          return fail("triangle vertices");
    if (!same_material(d.materials[d.tri_material[t]], room.materials[room.tri_material[t]]))
      return fail("triangle material");
  }
  for (uint32_t q = 0; q < d.num_quads; ++q) {
    for (int k = 0; k < 9; ++k)
      if (std::fabs(d.quads[9 * q + k] - room.quads[9 * q + k]) > 1e-7f) return fail("quad geometry");
    if (!same_material(d.materials[d.quad_material[q]], room.materials[room.quad_material[q]]))
      return fail("quad material (light Le)");
  }
  const mpt_camera c2 = milton_amd_adapter::to_camera(&r.camera);
  if (std::memcmp(c2.eye, cam.eye, sizeof(cam.eye)) || std::memcmp(c2.look, cam.look, sizeof(cam.look)) ||
      c2.height_angle != cam.height_angle || c2.aspect != cam.aspect || c2.aperture != 0)
    return fail("camera");
  // The same closest hits through the host build of the device tree.
  mpt_kd_tree *ta = nullptr, *tb = nullptr;
  if (mpt_kd_build(&d, nullptr, &ta) || mpt_kd_build(&room, nullptr, &tb)) return fail("mpt_kd_build");
  std::vector<mpt_ray> rays(20000);
  uint64_t s = 88172645463325252ull;
  for (mpt_ray &ry : rays) {
    double v[6];
    for (double &x : v) {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      x = (s >> 11) * (1.0 / 9007199254740992.0);
    }
    const double th = std::acos(1 - 2 * v[3]), ph = 2 * M_PI * v[4];
    for (int k = 0; k < 3; ++k) ry.o[k] = (float)(0.05 + 0.9 * v[k]);
    ry.d[0] = (float)(std::sin(th) * std::cos(ph));
    ry.d[1] = (float)(std::sin(th) * std::sin(ph));
    ry.d[2] = (float)std::cos(th);
  }
  std::vector<mpt_hit> ha(rays.size()), hb(rays.size());
  mpt_kd_trace(ta, rays.data(), rays.size(), ha.data(), nullptr, 2);
  mpt_kd_trace(tb, rays.data(), rays.size(), hb.data(), nullptr, 2);
  size_t diff = 0;
  for (size_t i = 0; i < rays.size(); ++i)
    if (!(ha[i].t == hb[i].t || std::fabs(ha[i].t - hb[i].t) <= 1e-6f * hb[i].t)) ++diff;
  mpt_kd_destroy(ta);
  mpt_kd_destroy(tb);
  if (diff) return fail("closest hits of the flattened scene");
  std::printf("flattened: %u triangles, %u quads, %u materials; %zu host rays agree\n", d.num_triangles,
              d.num_quads, d.num_materials, rays.size());

  if (!gpu) {
    // no device: init stops at mpt_open, render draws nothing
    ren->render();
    if (amd->ok() || out.finalized()) return fail("init without a device should fail");
    std::printf("no device: init failed cleanly (%s)\n", amd->facade()->lastError().c_str());
    plugin_delete(ren);
    mpt_gen_destroy(gen);
    return 0;
  }
  if (!amd->ok()) return fail(amd->facade()->lastError().c_str());
  ren->render();
  if (!out.finalized()) return fail("render");
  // the same job through the C ABI
  mpt_context *ctx = nullptr;
  mpt_scene *scene = nullptr;
  if (mpt_open(0, &ctx) || mpt_scene_create(ctx, &room, nullptr, &scene)) return fail("C ABI scene");
  mpt_job job{};
  job.camera = cam;
  job.width = W;
  job.height = H;
  job.spp = 16;
  job.max_depth = 5;
  job.seed = 7;
  job.direct_samples = 4;
  std::vector<float> ref(3 * W * H);
  if (mpt_render(scene, &job, ref.data(), nullptr)) return fail("mpt_render");
  size_t same = 0;
  double mean = 0;
  for (size_t i = 0; i < ref.size(); ++i) {
    same += out.pixels()[i] == (double)ref[i];
    mean += ref[i];
  }
  std::printf("adapter render: %zu of %zu values identical to the C ABI, mean %.6f\n", same, ref.size(),
              mean / ref.size());
  mpt_scene_destroy(scene);
  mpt_close(ctx);
  plugin_delete(ren);
  mpt_gen_destroy(gen);
  return same == ref.size() && mean > 0 ? 0 : 1;
}
