// Seeded benchmark scenes for the path-tracing core (host C++).
//
// The reference has no synthetic scenes; its Cornell box is a JSON scenefile
// (exec/scenes/cornellBox.js) whose light is a transformed unit "plane" with an
// "oriented" emitter (:52-67).  The room built here follows that construction:
// walls and boxes are mesh triangles with unshared per-face vertices (so
// Mesh::computeNormals, shapes/Mesh.cpp:335-377, leaves flat face normals),
// the light is a quad.  All coordinates are rounded to float so that the
// device path and the double-precision oracle intersect identical geometry.
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/milton_pt_scenes.h"
#include "rng.h"

namespace {

struct Gen {
  uint32_t kind = 0;
  std::vector<float> verts, spheres, quads;
  std::vector<uint32_t> idx, tri_mat, sph_mat, quad_mat;
  std::vector<mpt_material> mats;
  float env[3] = {0, 0, 0};
  mpt_camera cam{};
};

mpt_material diffuse(float r, float g, float b) {
  mpt_material m{};
  m.type = MPT_BSDF_DIFFUSE;
  m.kd[0] = r; m.kd[1] = g; m.kd[2] = b;
  return m;
}

// Appends the two triangles of the quad (p0, p1, p2, p3) in that winding.
// The face normal (p1 - p0) x (p2 - p0) follows the right-hand rule.
void add_quad_tris(Gen &g, const double p[4][3], uint32_t mat) {
  const uint32_t base = (uint32_t)(g.verts.size() / 3);
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) g.verts.push_back((float)p[i][k]);
  const uint32_t t[6] = {0, 1, 2, 0, 2, 3};
  for (int i = 0; i < 6; ++i) g.idx.push_back(base + t[i]);
  g.tri_mat.push_back(mat);
  g.tri_mat.push_back(mat);
}

// Axis-aligned box with outward-facing triangles.
void add_box(Gen &g, const double lo[3], const double hi[3], uint32_t mat) {
  const double x0 = lo[0], y0 = lo[1], z0 = lo[2], x1 = hi[0], y1 = hi[1], z1 = hi[2];
  const double faces[6][4][3] = {
      {{x0, y0, z0}, {x1, y0, z0}, {x1, y0, z1}, {x0, y0, z1}},   // -y
      {{x0, y1, z0}, {x0, y1, z1}, {x1, y1, z1}, {x1, y1, z0}},   // +y
      {{x0, y0, z0}, {x0, y1, z0}, {x1, y1, z0}, {x1, y0, z0}},   // -z
      {{x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}},   // +z
      {{x0, y0, z0}, {x0, y0, z1}, {x0, y1, z1}, {x0, y1, z0}},   // -x
      {{x1, y0, z0}, {x1, y1, z0}, {x1, y1, z1}, {x1, y0, z1}},   // +x
  };
  for (int f = 0; f < 6; ++f) add_quad_tris(g, faces[f], mat);
}

void set_camera(Gen &g, const double eye[3], const double look[3], float angle, float aspect) {
  for (int k = 0; k < 3; ++k) {
    g.cam.eye[k] = (float)eye[k];
    g.cam.look[k] = (float)look[k];
    g.cam.up[k] = k == 1 ? 1.0f : 0.0f;
  }
  g.cam.height_angle = angle;
  g.cam.aspect = aspect;
}

void build_room(Gen &g) {
  g.mats.push_back(diffuse(0.75f, 0.75f, 0.75f));   // 0 grey walls and boxes
  g.mats.push_back(diffuse(0.6f, 0.1f, 0.1f));      // 1 left wall
  g.mats.push_back(diffuse(0.1f, 0.6f, 0.1f));      // 2 right wall
  mpt_material light = diffuse(0.75f, 0.75f, 0.75f);
  light.Le[0] = light.Le[1] = light.Le[2] = 15.0f;
  g.mats.push_back(light);                          // 3 ceiling light
  // Walls, normals pointing into the room; the front (z = 0) is open.
  const double floor_[4][3] = {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}};
  const double ceil_[4][3] = {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
  const double back[4][3] = {{0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
  const double left[4][3] = {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}};
  const double right[4][3] = {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
  add_quad_tris(g, floor_, 0);
  add_quad_tris(g, ceil_, 0);
  add_quad_tris(g, back, 0);
  add_quad_tris(g, left, 1);
  add_quad_tris(g, right, 2);
  const double b1lo[3] = {0.15, 0.0, 0.45}, b1hi[3] = {0.45, 0.6, 0.75};
  const double b2lo[3] = {0.55, 0.0, 0.2}, b2hi[3] = {0.85, 0.3, 0.5};
  add_box(g, b1lo, b1hi, 0);
  add_box(g, b2lo, b2hi, 0);
  // Light: 0.25 x 0.25 quad just below the ceiling, facing -y:
  // normalize(ez x ex) = normalize((0,0,-.25) x (.25,0,0)) = (0,-1,0).
  const float q[9] = {0.5f, 0.999f, 0.5f, 0.25f, 0.0f, 0.0f, 0.0f, 0.0f, -0.25f};
  g.quads.assign(q, q + 9);
  g.quad_mat.push_back(3);
  const double eye[3] = {0.5, 0.5, -1.4}, look[3] = {0, 0, 1};
  set_camera(g, eye, look, 40.0f, 1.0f);
}

void build_soup(Gen &g, uint32_t n, uint32_t seed) {
  g.mats.push_back(diffuse(0.5f, 0.5f, 0.5f));
  g.env[0] = g.env[1] = g.env[2] = 1.0f;
  g.verts.resize(9ull * n);
  g.idx.resize(3ull * n);
  g.tri_mat.assign(n, 0);
  const double mu = std::log(0.005), sigma = 0.5;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t r[16];
    for (uint32_t w = 0; w < 4; ++w) mpt_rng::philox_words(seed, 0x534F5550u, i, w, r + 4 * w);   // "SOUP"
    const double c[3] = {mpt_rng::unit53(r[0], r[1]), mpt_rng::unit53(r[2], r[3]), mpt_rng::unit53(r[4], r[5])};
    // Box-Muller normal for the lognormal edge length.
    const double u1 = 1.0 - mpt_rng::unit32(r[6]), u2 = mpt_rng::unit32(r[7]);
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    const double L = std::exp(mu + sigma * z);
    for (int k = 0; k < 3; ++k) {
      // vertex k = centroid + L * (uniform direction on S^2)
      const double uz = 1.0 - 2.0 * mpt_rng::unit32(r[8 + k]);
      const double phi = 2.0 * M_PI * mpt_rng::unit32(r[12 + k]);
      const double s = std::sqrt(std::max(0.0, 1.0 - uz * uz));
      const double d[3] = {s * std::cos(phi), s * std::sin(phi), uz};
      for (int a = 0; a < 3; ++a) g.verts[9ull * i + 3 * k + a] = (float)(c[a] + L * d[a]);
    }
    for (int k = 0; k < 3; ++k) g.idx[3ull * i + k] = 3 * i + k;
  }
  const double eye[3] = {0.5, 0.5, -1.0}, look[3] = {0, 0, 1};
  set_camera(g, eye, look, 45.0f, 16.0f / 9.0f);
}

// Seeded value noise: smooth bilinear interpolation of hashed lattice values
// in [0, 1].
double lattice(uint32_t seed, uint32_t octave, int64_t ix, int64_t iz) {
  uint32_t r[4];
  mpt_rng::philox_words(seed, 0x4E4F4953u + octave, (uint32_t)ix, (uint32_t)iz, r);   // "NOIS"
  return mpt_rng::unit32(r[0]);
}
double value_noise(uint32_t seed, uint32_t octave, double x, double z) {
  const double fx = std::floor(x), fz = std::floor(z);
  const int64_t ix = (int64_t)fx, iz = (int64_t)fz;
  double tx = x - fx, tz = z - fz;
  tx = tx * tx * (3 - 2 * tx);
  tz = tz * tz * (3 - 2 * tz);
  const double a = lattice(seed, octave, ix, iz), b = lattice(seed, octave, ix + 1, iz);
  const double c = lattice(seed, octave, ix, iz + 1), d = lattice(seed, octave, ix + 1, iz + 1);
  return (a * (1 - tx) + b * tx) * (1 - tz) + (c * (1 - tx) + d * tx) * tz;
}
// World size of the terrain: 500 spheres of radius 0.01-0.05 then cover about
// 10% of it, the dielectric share of the materials.
const double kTerrainSize = 4.0;

double terrain_height(uint32_t seed, double x, double z) {
  // 6 octaves, amplitude 0.2, lacunarity 2, gain 0.5; the first octave has
  // 4 noise cells across the terrain.
  double h = 0, amp = 0.2, freq = 4.0 / kTerrainSize;
  for (uint32_t o = 0; o < 6; ++o) {
    h += amp * value_noise(seed, o, x * freq, z * freq);
    amp *= 0.5;
    freq *= 2.0;
  }
  return h;
}

void build_terrain(Gen &g, uint32_t side, uint32_t seed) {
  mpt_material lam = diffuse(0.55f, 0.5f, 0.45f);                 // 0 Lambertian
  mpt_material phong{};                                           // 1 modified Phong
  phong.type = MPT_BSDF_PHONG;
  phong.kd[0] = phong.kd[1] = phong.kd[2] = 0.2f;
  phong.ks[0] = phong.ks[1] = phong.ks[2] = 0.6f;
  phong.exponent = 50.0f;
  mpt_material glass{};                                           // 2 dielectric
  glass.type = MPT_BSDF_DIELECTRIC;
  glass.ks[0] = glass.ks[1] = glass.ks[2] = 1.0f;
  glass.ior = 1.5f;
  glass.transparency = 1.0f;
  mpt_material sun{};                                             // 3 distant light
  sun.type = MPT_BSDF_DIFFUSE;
  sun.Le[0] = sun.Le[1] = sun.Le[2] = 40000.0f;
  g.mats = {lam, phong, glass, sun};
  const uint32_t n = side;
  g.verts.resize(3ull * n * n);
  const double inv = kTerrainSize / (n - 1);
  for (uint32_t j = 0; j < n; ++j)
    for (uint32_t i = 0; i < n; ++i) {
      const double x = i * inv, z = j * inv;
      float *v = &g.verts[3ull * (j * (uint64_t)n + i)];
      v[0] = (float)x;
      v[1] = (float)terrain_height(seed, x, z);
      v[2] = (float)z;
    }
  const uint64_t cells = (uint64_t)(n - 1) * (n - 1);
  g.idx.resize(6 * cells);
  g.tri_mat.resize(2 * cells);
  for (uint32_t j = 0; j + 1 < n; ++j)
    for (uint32_t i = 0; i + 1 < n; ++i) {
      const uint64_t c = (uint64_t)j * (n - 1) + i;
      const uint32_t v00 = j * n + i, v10 = v00 + 1, v01 = v00 + n, v11 = v01 + 1;
      // Upward (+y) facing: (v01 - v00) x (v10 - v00) has +y.
      const uint32_t t[6] = {v00, v01, v10, v10, v01, v11};
      for (int k = 0; k < 6; ++k) g.idx[6 * c + k] = t[k];
      uint32_t r[4];
      mpt_rng::philox_words(seed, 0x4D415453u, (uint32_t)c, 0, r);   // "MATS"
      // 70% : 20% Lambertian : Phong over the terrain (the remaining 10% of
      // the materials are the dielectric spheres).
      const uint32_t m = mpt_rng::unit32(r[0]) < 7.0 / 9.0 ? 0 : 1;
      g.tri_mat[2 * c] = m;
      g.tri_mat[2 * c + 1] = m;
    }
  // 500 dielectric spheres resting on the terrain.
  for (uint32_t s = 0; s < 500; ++s) {
    uint32_t r[4];
    mpt_rng::philox_words(seed, 0x53504852u, s, 0, r);   // "SPHR"
    const double x = kTerrainSize * (0.05 + 0.9 * mpt_rng::unit32(r[0]));
    const double z = kTerrainSize * (0.05 + 0.9 * mpt_rng::unit32(r[1]));
    const double rad = 0.01 + 0.04 * mpt_rng::unit32(r[2]);
    const float sp[4] = {(float)x, (float)(terrain_height(seed, x, z) + rad), (float)z, (float)rad};
    g.spheres.insert(g.spheres.end(), sp, sp + 4);
    g.sph_mat.push_back(2);
  }
  // Distant area light: a square at distance 100 subtending 0.5 degrees,
  // facing the terrain centre (2, 0, 2).
  const double dir[3] = {0.3, 1.0, 0.2};
  const double dl = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  const double l[3] = {dir[0] / dl, dir[1] / dl, dir[2] / dl};
  const double side_len = 2.0 * 100.0 * std::tan(0.25 * M_PI / 180.0);
  // Orthonormal a, b with b = l x a, so a x b = l and
  // normalize(ez x ex) = normalize(b x a) = -l for ex = side * a, ez = side * b.
  double a[3] = {-l[1], l[0], 0.0};   // (0,0,1) x l
  const double al = std::sqrt(a[0] * a[0] + a[1] * a[1]);
  a[0] /= al;
  a[1] /= al;
  const double b[3] = {l[1] * a[2] - l[2] * a[1], l[2] * a[0] - l[0] * a[2], l[0] * a[1] - l[1] * a[0]};
  const double cx = 0.5 * kTerrainSize, cz = 0.5 * kTerrainSize;
  const float q[9] = {(float)(cx + 100.0 * l[0]), (float)(100.0 * l[1]), (float)(cz + 100.0 * l[2]),
                      (float)(side_len * a[0]), (float)(side_len * a[1]), (float)(side_len * a[2]),
                      (float)(side_len * b[0]), (float)(side_len * b[1]), (float)(side_len * b[2])};
  g.quads.assign(q, q + 9);
  g.quad_mat.push_back(3);
  const double eye[3] = {2.0, 1.4, -0.6};
  const double look[3] = {0.0, -0.55, 1.0};
  set_camera(g, eye, look, 45.0f, 1.0f);
}

}  // namespace

struct mpt_gen : Gen {};

extern "C" {

int mpt_gen_create(uint32_t kind, uint32_t count, uint32_t seed, mpt_gen **out) {
  if (!out) return MPT_E_INVALID;
  *out = nullptr;
  std::unique_ptr<mpt_gen> g(new (std::nothrow) mpt_gen());
  if (!g) return MPT_E_NOMEM;
  g->kind = kind;
  try {
    if (kind == MPT_SCENE_ROOM) {
      build_room(*g);
    } else if (kind == MPT_SCENE_SOUP) {
      build_soup(*g, count ? count : 1000000u, seed);
    } else if (kind == MPT_SCENE_TERRAIN) {
      if (count == 1) return MPT_E_INVALID;
      build_terrain(*g, count ? count : 2048u, seed);
    } else {
      return MPT_E_INVALID;
    }
  } catch (const std::bad_alloc &) {
    return MPT_E_NOMEM;
  }
  *out = g.release();
  return MPT_OK;
}

int mpt_gen_describe(const mpt_gen *g, mpt_scene_desc *d, mpt_camera *cam) {
  if (!g || !d) return MPT_E_INVALID;
  std::memset(d, 0, sizeof(*d));
  d->version = MPT_ABI_VERSION;
  d->num_vertices = (uint32_t)(g->verts.size() / 3);
  d->vertices = g->verts.data();
  d->num_triangles = (uint32_t)(g->idx.size() / 3);
  d->indices = g->idx.data();
  d->tri_material = g->tri_mat.data();
  d->num_spheres = (uint32_t)(g->spheres.size() / 4);
  d->spheres = g->spheres.data();
  d->sphere_material = g->sph_mat.data();
  d->num_quads = (uint32_t)(g->quads.size() / 9);
  d->quads = g->quads.data();
  d->quad_material = g->quad_mat.data();
  d->num_materials = (uint32_t)g->mats.size();
  d->materials = g->mats.data();
  for (int k = 0; k < 3; ++k) d->env_radiance[k] = g->env[k];
  if (cam) *cam = g->cam;
  return MPT_OK;
}

int mpt_gen_destroy(mpt_gen *g) {
  delete g;
  return MPT_OK;
}

int mpt_gen_rays(const mpt_gen *g, uint64_t n, uint32_t seed, uint32_t film_w, uint32_t film_h, mpt_ray *rays) {
  if (!g || !rays || film_w == 0 || film_h == 0) return MPT_E_INVALID;
  // Pinhole camera basis as in ThinLensCamera::setOrientation
  // (cameras/ThinLensCamera.cpp:241-256).
  const mpt_camera &c = g->cam;
  double nn[3] = {-c.look[0], -c.look[1], -c.look[2]};
  double u[3] = {c.up[1] * nn[2] - c.up[2] * nn[1], c.up[2] * nn[0] - c.up[0] * nn[2],
                 c.up[0] * nn[1] - c.up[1] * nn[0]};
  auto norm = [](double *v) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
  };
  norm(nn);
  norm(u);
  const double v[3] = {nn[1] * u[2] - nn[2] * u[1], nn[2] * u[0] - nn[0] * u[2], nn[0] * u[1] - nn[1] * u[0]};
  const double th = std::tan(c.height_angle * M_PI / 360.0), tw = c.aspect * th;
  const uint64_t half = n / 2;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r[4];
    mpt_rng::philox_words(seed, 0x52415953u, (uint32_t)i, (uint32_t)(i >> 32), r);   // "RAYS"
    double o[3], d[3];
    if (i < half) {
      const uint64_t pix = i % ((uint64_t)film_w * film_h);
      const double x = ((double)(pix % film_w) + mpt_rng::unit32(r[0])) / film_w;
      const double y = ((double)(pix / film_w) + mpt_rng::unit32(r[1])) / film_h;
      const double nx = 2 * x - 1, ny = 1 - 2 * y;
      for (int k = 0; k < 3; ++k) {
        o[k] = c.eye[k];
        d[k] = u[k] * nx * tw + v[k] * ny * th - nn[k];
      }
    } else {
      for (int k = 0; k < 3; ++k) o[k] = mpt_rng::unit32(r[k]);
      uint32_t r2[4];
      mpt_rng::philox_words(seed, 0x44495253u, (uint32_t)i, (uint32_t)(i >> 32), r2);   // "DIRS"
      const double uz = 1.0 - 2.0 * mpt_rng::unit32(r2[0]);
      const double phi = 2.0 * M_PI * mpt_rng::unit32(r2[1]);
      const double s = std::sqrt(std::max(0.0, 1.0 - uz * uz));
      d[0] = s * std::cos(phi);
      d[1] = s * std::sin(phi);
      d[2] = uz;
    }
    norm(d);
    for (int k = 0; k < 3; ++k) {
      rays[i].o[k] = (float)o[k];
      rays[i].d[k] = (float)d[k];
    }
  }
  return MPT_OK;
}

}  // extern "C"
