// Milton scene files (include/milton_pt.h, "Scene files").
//
// Reads the JSON-like scenefile format of loaders/MiltonJSONSceneLoader.cpp
// (comments, trailing commas and repeated keys allowed, as the reference's
// grammar does) and flattens its scene graph into an mpt_scene_desc plus the
// camera, film size and sampling settings of a job.  Supported subset, each
// mapped as the reference builds it:
//   scene / shapeSet   nested groups (MiltonJSONSceneLoader.cpp:1493-1525)
//   material           bsdf diffuse, modifiedPhong / phong, dielectric,
//                      specular, transmissive, absorbent (:901-969); emitter
//                      oriented or null (:977-1053); medium ior (:1055-1097);
//                      applies to every shape nested in it
//   transform          translate, rotate [px py pz ax ay az degrees], scale,
//                      arbitraryTransform (row-major 4x4), composed as
//                      transform * translation * rotation * scale per node and
//                      parent * child down the graph (SceneGraph.h:68-70)
//   plane              with "vertices": two one-sided triangles (v1 v2 v3),
//                      (v1 v3 v4) (:612-627); otherwise the transformed unit
//                      plane y = 0, |x|, |z| <= 0.5 as a quad
//   cube               the transformed unit cube as six quads with outward
//                      normals (shapes/Cube.cpp:17-52; not emitting)
//   triangle           one-sided triangle with its face normal (:526-567)
//   sphere             unit sphere of radius 0.5 under position / radius
//                      (:569-583) and the transform (similarities only)
//   mesh               OBJ or ASCII PLY file, optional "normalize" (:484-525)
//   background         emitter with a constant "power", or an "environment"
//                      image (.hdr / .rgbe / .pfm / .ppm) with its filter
//   camera             pinhole / thinlens (aperture, fstop, focalDistance,
//                      focalPoint), eye, look or focus, up, heightAngle,
//                      aspectRatio (:842-899)
//   output             width, height, size, dimensions (default 480 x 480),
//                      tonemap (:783; "default" = Reinhard), type and filter
//                      (box, triangle, gaussian, mitchell, lanczosSinc)
//   renderer           type (bidirectionalPathTracer / bidirPathTracer select
//                      the bidirectional renderer, :720-721; mlt / MLT
//                      Metropolis light transport, :731-738; any other type
//                      renders with the path tracer), noSuperSamples (spp),
//                      noDirectSamples, seed, maxDepth, efficientDirect
//                      (maxDepth: a cap the reference does not have; default
//                      8); for mlt noRenderThreads (chains), mltNoInitialPaths,
//                      mltMaxDepth, mltMaxConsequtiveRejections, the three
//                      mlt*MutationProb, mltFilterProposed, and mltMutations
//                      per chain (the reference runs until stopped; default
//                      noSuperSamples x pixels / chains, at most 32767)
//   point              with an "omni" emitter: a point light (PointShape,
//                      OmniEmitter)
// Anything else that would change the image (cone, cylinder, blob,
// dynamic plugins, emitting triangles, spheres or cubes, images Qt or
// OpenEXR would read, non-uniform sphere scales, sheared planes or cubes) is an error,
// not silently dropped.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/milton_pt.h"
#include "envmap.h"

namespace {

// ---------------------------------------------------------------------------
// Tolerant JSON.
// ---------------------------------------------------------------------------
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;   // in file order, repeats kept

  const JVal *find(const char *key) const {
    for (const auto &kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class JParser {
 public:
  explicit JParser(const std::string &text) : s_(text) {}

  bool parse(JVal &out, std::string &err) {
    if (!value(out, 0)) {
      err = err_;
      return false;
    }
    ws();
    if (i_ != s_.size()) return fail(err, "trailing characters after the root value");
    return true;
  }

 private:
  bool fail(std::string &err, const char *what) {
    err_ = msg(what);
    err = err_;
    return false;
  }
  bool fail(const char *what) {
    err_ = msg(what);
    return false;
  }
  std::string msg(const char *what) const {
    int line = 1;
    for (size_t k = 0; k < i_ && k < s_.size(); ++k) line += s_[k] == '\n';
    return std::string("line ") + std::to_string(line) + ": " + what;
  }
  void ws() {
    for (;;) {
      while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
      if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
        const size_t e = s_.find("*/", i_ + 2);
        i_ = e == std::string::npos ? s_.size() : e + 2;
      } else if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
        while (i_ < s_.size() && s_[i_] != '\n') ++i_;
      } else {
        return;
      }
    }
  }
  bool string(std::string &out) {
    if (s_[i_] != '"') return fail("expected a string");
    ++i_;
    out.clear();
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\' && i_ < s_.size()) {
        const char e = s_[i_++];
        c = e == 'n' ? '\n' : e == 't' ? '\t' : e;
      }
      out.push_back(c);
    }
    if (i_ >= s_.size()) return fail("unterminated string");
    ++i_;
    return true;
  }
  bool value(JVal &v, int depth) {
    if (depth > 200) return fail("nesting too deep");
    ws();
    if (i_ >= s_.size()) return fail("unexpected end of input");
    const char c = s_[i_];
    if (c == '{') {
      v.kind = JVal::OBJ;
      ++i_;
      for (;;) {
        ws();
        if (i_ < s_.size() && s_[i_] == '}') {
          ++i_;
          return true;
        }
        std::string key;
        if (i_ >= s_.size() || !string(key)) return fail("expected a key or '}'");
        ws();
        if (i_ >= s_.size() || s_[i_] != ':') return fail("expected ':'");
        ++i_;
        v.obj.emplace_back(key, JVal());
        if (!value(v.obj.back().second, depth + 1)) return false;
        ws();
        if (i_ < s_.size() && s_[i_] == ',') ++i_;
        else if (i_ < s_.size() && s_[i_] == '}') continue;
        else return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = JVal::ARR;
      ++i_;
      for (;;) {
        ws();
        if (i_ < s_.size() && s_[i_] == ']') {
          ++i_;
          return true;
        }
        v.arr.emplace_back();
        if (!value(v.arr.back(), depth + 1)) return false;
        ws();
        if (i_ < s_.size() && s_[i_] == ',') ++i_;
        else if (i_ < s_.size() && s_[i_] == ']') continue;
        else return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = JVal::STR;
      return string(v.str);
    }
    if (s_.compare(i_, 4, "true") == 0) {
      v.kind = JVal::BOOL;
      v.b = true;
      i_ += 4;
      return true;
    }
    if (s_.compare(i_, 5, "false") == 0) {
      v.kind = JVal::BOOL;
      i_ += 5;
      return true;
    }
    if (s_.compare(i_, 4, "null") == 0) {
      v.kind = JVal::NUL;
      i_ += 4;
      return true;
    }
    // numbers as JavaScript writes them (".75", "-5", "1e3")
    const char *b = s_.c_str() + i_;
    char *e = nullptr;
    const double x = std::strtod(b, &e);
    if (e == b) return fail("unexpected character");
    v.kind = JVal::NUM;
    v.num = x;
    i_ += (size_t)(e - b);
    return true;
  }

  const std::string &s_;
  size_t i_ = 0;
  std::string err_;
};

// ---------------------------------------------------------------------------
// 4x4 affine transforms in double (row-major, column vectors).
// ---------------------------------------------------------------------------
struct Mat {
  double m[4][4];
  static Mat identity() {
    Mat r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1;
    return r;
  }
  Mat operator*(const Mat &o) const {
    Mat r{};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) r.m[i][j] += m[i][k] * o.m[k][j];
    return r;
  }
  void point(const double p[3], double out[3]) const {
    for (int i = 0; i < 3; ++i) out[i] = m[i][0] * p[0] + m[i][1] * p[1] + m[i][2] * p[2] + m[i][3];
  }
  void vector(const double v[3], double out[3]) const {
    for (int i = 0; i < 3; ++i) out[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
  }
};

Mat translate(const double t[3]) {
  Mat r = Mat::identity();
  for (int i = 0; i < 3; ++i) r.m[i][3] = t[i];
  return r;
}

Mat scale(const double s[3]) {
  Mat r = Mat::identity();
  for (int i = 0; i < 3; ++i) r.m[i][i] = s[i];
  return r;
}

// Rotation by `deg` degrees about the axis `a` through the point `p`,
// right-handed (Rodrigues); the reference builds the same rotation from
// elementary rotations (common/math/Matrix.cpp:83-99).
Mat rotate(const double p[3], const double a[3], double deg) {
  const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double x = a[0] / l, y = a[1] / l, z = a[2] / l;
  const double th = deg * M_PI / 180.0, c = std::cos(th), s = std::sin(th), t = 1 - c;
  Mat r = Mat::identity();
  r.m[0][0] = t * x * x + c;     r.m[0][1] = t * x * y - s * z; r.m[0][2] = t * x * z + s * y;
  r.m[1][0] = t * x * y + s * z; r.m[1][1] = t * y * y + c;     r.m[1][2] = t * y * z - s * x;
  r.m[2][0] = t * x * z - s * y; r.m[2][1] = t * y * z + s * x; r.m[2][2] = t * z * z + c;
  const double mp[3] = {-p[0], -p[1], -p[2]};
  return translate(p) * r * translate(mp);
}

// ---------------------------------------------------------------------------
// The flattened scene.
// ---------------------------------------------------------------------------
struct Flat {
  std::vector<float> verts, normals;     // normals: one per vertex when has_normals
  std::vector<uint32_t> idx, tri_mat, tri_flags;
  std::vector<uint32_t> tri_shape;       // per triangle: the mesh or triangle shape it came from
  uint32_t shapes = 0;
  std::vector<float> sphere_frames;      // per sphere: images of the object axes (9 floats)
  bool any_normals = false;              // some mesh brought its own normals
  std::vector<uint8_t> vert_has_normal;  // per vertex: normal given by the file
  std::vector<float> spheres, quads;
  std::vector<float> points;             // point lights: position, power
  std::vector<uint32_t> sphere_mat, quad_mat;
  std::vector<mpt_material> mats;
  std::vector<int> mat_emits;            // per material: 1 oriented emitter, 2 omni emitter (with power)
  std::vector<double> mat_power;         // 3 per material
  float env[3] = {0, 0, 0};
  bool has_env_map = false;              // background is an environment image
  std::vector<float> env_rgb;
  mpt_env_map env_map{};
};

struct Settings {
  double eye[3] = {2, 2, 2}, look[3] = {-1, -1, -1}, up[3] = {0, 1, 0};
  double height_angle = 45, aspect = 1;
  // thin lens (metres); aperture 0 = pinhole
  double aperture = 0, focal_length = 0, focal_distance = 1, focal_point[2] = {0.5, 0.5};
  bool autofocus = false;
  uint32_t width = 480, height = 480, spp = 1, max_depth = 8, seed = 1, direct = 1;
  uint32_t direct_off = 0;   // the renderer's efficientDirect = false
  uint32_t renderer = MPT_RENDERER_PATH;
  // Metropolis light transport (MLTRenderer::init / render, renderers/mlt/
  // MLTRenderer.cpp:52-91; MLTAggregatePathMutation.cpp:41-61; RenderOutput::
  // init, renderers/RenderOutput.cpp:63): the reference's defaults
  uint32_t mlt_chains = 1, mlt_mutations = 0, mlt_seed_paths = 10000, mlt_max_depth = 10;
  uint32_t mlt_max_rejections = 500, mlt_flags = 0;
  float mlt_probs[3] = {1, 1, 1};
  uint32_t tonemap = MPT_TONEMAP_REINHARD;   // RenderOutput::init's "default"
  uint32_t filter = 0, filter_support = 0;    // mpt_job.filter: 0 pixel bin, else 1 + MPT_FILTER_*
  float filter_params[2] = {0, 0};
};

class Loader {
 public:
  Loader(std::string base_dir) : base_(std::move(base_dir)) {}

  bool load(const JVal &root, std::string &err) {
    const JVal *sf = root.kind == JVal::OBJ ? root.find("scenefile") : nullptr;
    if (!sf || sf->kind != JVal::OBJ) return fail(err, "missing \"scenefile\" object");
    for (const auto &kv : sf->obj) {
      const std::string &k = kv.first;
      bool ok = true;
      if (k == "renderer") ok = renderer(kv.second);
      else if (k == "camera") ok = camera(kv.second);
      else if (k == "output") ok = output(kv.second);
      else if (k == "scene") ok = group(kv.second, Mat::identity(), -1, true);
      if (!ok) return fail(err, err_.c_str());
    }
    if (flat_.idx.empty() && flat_.spheres.empty() && flat_.quads.empty())
      return fail(err, "the scene has no supported shapes");
    return true;
  }

  Flat flat_;
  Settings set_;

 private:
  bool fail(std::string &err, const char *what) {
    err = what;
    return false;
  }
  bool bad(const std::string &what) {
    err_ = what;
    return false;
  }

  static bool num(const JVal *v, double &out) {
    if (!v || v->kind != JVal::NUM) return false;
    out = v->num;
    return true;
  }
  static bool vec(const JVal *v, double *out, size_t n) {
    if (!v || v->kind != JVal::ARR || v->arr.size() != n) return false;
    for (size_t i = 0; i < n; ++i) {
      if (v->arr[i].kind != JVal::NUM) return false;
      out[i] = v->arr[i].num;
    }
    return true;
  }
  // A spectrum: one number for all channels or three.
  static bool spectrum(const JVal *v, double out[3]) {
    if (v && v->kind == JVal::NUM) {
      out[0] = out[1] = out[2] = v->num;
      return true;
    }
    return vec(v, out, 3);
  }

  bool renderer(const JVal &v) {
    if (v.kind != JVal::OBJ) return bad("renderer must be an object");
    double x;
    if (num(v.find("noSuperSamples"), x)) set_.spp = (uint32_t)std::max(1.0, x);
    if (num(v.find("maxDepth"), x)) set_.max_depth = (uint32_t)std::max(1.0, x);
    if (num(v.find("seed"), x)) set_.seed = (uint32_t)x;
    if (num(v.find("noDirectSamples"), x)) set_.direct = (uint32_t)std::max(1.0, x);
    // MiltonJSONSceneLoader::_parseRenderer (loaders/MiltonJSONSceneLoader.cpp:720-721)
    const JVal *type = v.find("type");
    if (type && type->kind == JVal::STR && (type->str == "bidirectionalPathTracer" || type->str == "bidirPathTracer"))
      set_.renderer = MPT_RENDERER_BIDIR;
    if (type && type->kind == JVal::STR && (type->str == "mlt" || type->str == "MLT")) {
      set_.renderer = MPT_RENDERER_MLT;
      if (num(v.find("noRenderThreads"), x)) set_.mlt_chains = (uint32_t)std::max(1.0, x);
      if (num(v.find("mltNoInitialPaths"), x)) set_.mlt_seed_paths = (uint32_t)std::max(1.0, x);
      if (num(v.find("mltMaxDepth"), x)) set_.mlt_max_depth = (uint32_t)std::max(0.0, x);
      if (num(v.find("mltMaxConsequtiveRejections"), x)) set_.mlt_max_rejections = (uint32_t)std::max(0.0, x);
      if (num(v.find("mltMutations"), x)) set_.mlt_mutations = (uint32_t)std::max(1.0, x);
      if (num(v.find("mltBidirPathMutationProb"), x)) set_.mlt_probs[0] = (float)x;
      if (num(v.find("mltLensSubpathMutationProb"), x)) set_.mlt_probs[1] = (float)x;
      if (num(v.find("mltPerturbationPathMutationProb"), x)) set_.mlt_probs[2] = (float)x;
      const JVal *fp = v.find("mltFilterProposed");
      if (fp) {
        if (fp->kind == JVal::BOOL) set_.mlt_flags = fp->b ? 0u : 1u;
        else if (fp->kind == JVal::NUM) set_.mlt_flags = fp->num == 0 ? 1u : 0u;
        else return bad("renderer.mltFilterProposed must be a boolean");
      }
    }
    // PathTracer::init (renderers/renderers/PathTracer.cpp:22-26)
    const JVal *ed = v.find("efficientDirect");
    if (ed) {
      if (ed->kind == JVal::BOOL) set_.direct_off = ed->b ? 0u : 1u;
      else if (ed->kind == JVal::NUM) set_.direct_off = ed->num == 0 ? 1u : 0u;
      else return bad("renderer.efficientDirect must be a boolean");
    }
    return true;
  }

  bool camera(const JVal &v) {
    if (v.kind != JVal::OBJ) return bad("camera must be an object");
    const JVal *type = v.find("type");
    if (!type || type->kind != JVal::STR || (type->str != "pinhole" && type->str != "thinlens"))
      return bad("camera.type must be pinhole or thinlens");
    vec(v.find("eye"), set_.eye, 3);
    double focus[3];
    if (!vec(v.find("look"), set_.look, 3)) {
      const double o[3] = {0, 0, 0};
      const double *f = vec(v.find("focus"), focus, 3) ? focus : o;
      for (int i = 0; i < 3; ++i) set_.look[i] = f[i] - set_.eye[i];
    }
    vec(v.find("up"), set_.up, 3);
    num(v.find("heightAngle"), set_.height_angle);
    num(v.find("aspectRatio"), set_.aspect);
    if (type->str == "thinlens") {
      // ThinLensCamera::init (cameras/ThinLensCamera.cpp:88-118): aperture in
      // mm (default 35), fstop (default 14), focal length = fstop * aperture;
      // autofocus through focalPoint (default the film centre) unless only a
      // focalDistance (default 1) is given.
      double mm = 35, fstop = 14;
      num(v.find("aperture"), mm);
      num(v.find("fstop"), fstop);
      if (!(mm >= 0) || !(fstop > 0)) return bad("camera aperture must be >= 0 and fstop > 0");
      set_.aperture = mm / 1000.0;
      set_.focal_length = fstop * set_.aperture;
      const JVal *fp = v.find("focalPoint");
      const JVal *fd = v.find("focalDistance");
      set_.autofocus = fp || !fd;
      if (fp && !vec(fp, set_.focal_point, 2)) return bad("focalPoint expects 2 numbers");
      if (fd && (!num(fd, set_.focal_distance) || !(set_.focal_distance > 0)))
        return bad("focalDistance must be a positive number");
    } else {
      set_.aperture = 0;   // PinholeCamera::init (cameras/PinholeCamera.cpp:21-25)
    }
    return true;
  }

  bool output(const JVal &v) {
    if (v.kind != JVal::OBJ) return bad("output must be an object");
    double wh[2], x;
    if (vec(v.find("size"), wh, 2) || vec(v.find("dimensions"), wh, 2)) {
      set_.width = (uint32_t)wh[0];
      set_.height = (uint32_t)wh[1];
    }
    if (num(v.find("width"), x)) set_.width = (uint32_t)x;
    if (num(v.find("height"), x)) set_.height = (uint32_t)x;
    if (const JVal *tm = v.find("tonemap")) {
      if (tm->kind != JVal::STR) return bad("output.tonemap must be a string");
      set_.tonemap = mpt_tonemap_from_name(tm->str.c_str());
    }
    // the reconstruction filter (MiltonJSONSceneLoader.cpp:771-821 and
    // _parse_filter, :1389-1440); a "naive" output ignores it
    std::string type = "reconstruction";
    if (const JVal *t = v.find("type")) {
      if (t->kind != JVal::STR) return bad("output.type must be a string");
      type = t->str;
    }
    if (type != "naive" && type != "default" && type != "reconstruction")
      return bad("output type '" + type + "' is not supported");
    const JVal *f = v.find("filter");
    if (f && type != "naive") {
      if (f->kind != JVal::OBJ) return bad("output.filter must be an object");
      const JVal *ft = f->find("type");
      if (!ft || ft->kind != JVal::STR) return bad("unable to infer filter type");
      const std::string &n = ft->str;
      uint32_t k;
      if (n == "box") k = MPT_FILTER_BOX;
      else if (n == "triangle") k = MPT_FILTER_TRIANGLE;
      else if (n == "gaussian") k = MPT_FILTER_GAUSSIAN;
      else if (n == "mitchell") k = MPT_FILTER_MITCHELL;
      else if (n == "lanczosSinc") k = MPT_FILTER_LANCZOS;
      else return bad("output filter '" + n + "' is not supported");
      set_.filter = 1 + k;
      if (num(f->find("support"), x)) {
        if (!(x >= 1 && x <= 64)) return bad("filter support must be in [1, 64]");
        set_.filter_support = (uint32_t)x;
      }
      if (k == MPT_FILTER_GAUSSIAN && num(f->find("sigma"), x)) set_.filter_params[0] = (float)x;
      if (k == MPT_FILTER_MITCHELL) {
        set_.filter_params[0] = set_.filter_params[1] = 1.0f / 3.0f;
        if (num(f->find("B"), x)) set_.filter_params[0] = (float)x;
        if (num(f->find("C"), x)) set_.filter_params[1] = (float)x;
      }
      if (k == MPT_FILTER_LANCZOS && num(f->find("tau"), x)) set_.filter_params[0] = (float)x;
    }
    return true;
  }

  // Node transform of a "transform" object (SceneGraph.h:68-70).
  bool node_transform(const JVal &v, Mat &out) {
    Mat arb = Mat::identity(), tr = Mat::identity(), rot = Mat::identity(), sc = Mat::identity();
    for (const auto &kv : v.obj) {
      double d[16];
      if (kv.first == "translate") {
        if (!vec(&kv.second, d, 3)) return bad("translate expects 3 numbers");
        tr = tr * translate(d);
      } else if (kv.first == "rotate") {
        if (!vec(&kv.second, d, 7)) return bad("rotate expects [px, py, pz, vx, vy, vz, degrees]");
        rot = rot * rotate(d, d + 3, d[6]);
      } else if (kv.first == "scale") {
        if (!vec(&kv.second, d, 3)) return bad("scale expects 3 numbers");
        sc = sc * scale(d);
      } else if (kv.first == "arbitraryTransform") {
        if (!vec(&kv.second, d, 16)) return bad("arbitraryTransform expects 16 numbers, row-major");
        Mat a;
        for (int i = 0; i < 16; ++i) a.m[i / 4][i % 4] = d[i];
        arb = arb * a;
      }
    }
    out = arb * tr * rot * sc;
    return true;
  }

  // Children of a group, material or transform object.
  bool group(const JVal &v, const Mat &M, int mat, bool root) {
    if (v.kind != JVal::OBJ) return bad("expected an object");
    for (const auto &kv : v.obj) {
      const std::string &k = kv.first;
      const JVal &c = kv.second;
      if (k == "material") {
        int m = -1;
        if (!material(c, m)) return false;
        if (!group(c, M, m, false)) return false;
      } else if (k == "transform") {
        if (c.kind != JVal::OBJ) return bad("transform must be an object");
        Mat T;
        if (!node_transform(c, T)) return false;
        if (!group(c, M * T, mat, false)) return false;
      } else if (k == "shapeSet") {
        if (!group(c, M, mat, false)) return false;
      } else if (k == "shape") {
        const JVal *type = c.find("type");
        if (!type || type->kind != JVal::STR) return bad("shape needs a type");
        if (!shape(type->str, c, M, mat)) return false;
      } else if (k == "plane" || k == "sphere" || k == "triangle" || k == "mesh" || k == "cube" || k == "point") {
        if (!shape(k, c, M, mat)) return false;
      } else if (k == "cone" || k == "cylinder" || k == "blob" ||
                 k == "meta" || k == "dynamic") {
        return bad("shape '" + k + "' is not supported by this path tracer");
      } else if (k == "background" && root) {
        if (!background(c)) return false;
      }
    }
    return true;
  }

  bool background(const JVal &v) {
    if (v.kind != JVal::OBJ) return bad("background must be an object");
    if (v.obj.empty()) return true;
    const JVal *e = v.find("emitter");
    if (!e || e->kind != JVal::OBJ) return bad("background needs one emitter");
    const JVal *type = e->find("type");
    if (type && type->kind == JVal::STR && type->str == "environment") return environment(*e);
    double p[3];
    if (!spectrum(e->find("power"), p)) return bad("background emitter needs a power");
    // a constant environment emits its power as radiance (1x1 image semantics)
    for (int i = 0; i < 3; ++i) flat_.env[i] = (float)p[i];
    return true;
  }

  // An "environment" background emitter (MiltonJSONSceneLoader.cpp:1025-1048):
  // the image at "path" (relative paths against the scene file's directory),
  // "repeatU" / "repeatV", and the material's kernel filter: "filter"
  // (Material::init default "triangle", materials/Material.cpp:52-56) with
  // "support", "sigma", "B", "C", "tau" (filters/KernelFilter.cpp:18-45).
  bool environment(const JVal &e) {
    const JVal *path = e.find("path");
    if (!path || path->kind != JVal::STR)
      return bad("environment map must contain a 'path' to an environment texture");
    const std::string &rel = path->str;
    const std::string file = (!rel.empty() && rel[0] == '/') || base_.empty() ? rel : base_ + "/" + rel;
    uint32_t w = 0, h = 0;
    if (mpt_image_read(file.c_str(), &w, &h, nullptr, 0) != MPT_E_INVALID)
      return bad("environment image '" + file + "' cannot be read (hdr, rgbe, pfm or ppm)");
    flat_.env_rgb.assign(3 * (size_t)w * h, 0.0f);
    if (mpt_image_read(file.c_str(), &w, &h, flat_.env_rgb.data(), flat_.env_rgb.size()) != MPT_OK)
      return bad("environment image '" + file + "' cannot be read");
    mpt_env_map &em = flat_.env_map;
    em = mpt_env_map{};
    em.width = w;
    em.height = h;
    em.filter = MPT_FILTER_TRIANGLE;
    double x;
    if (num(e.find("repeatU"), x)) em.repeat_u = (float)x;
    if (num(e.find("repeatV"), x)) em.repeat_v = (float)x;
    if (const JVal *f = e.find("filter")) {
      if (f->kind != JVal::STR) return bad("filter must be a string");
      const std::string &n = f->str;
      if (n == "box") em.filter = MPT_FILTER_BOX;
      else if (n == "triangle") em.filter = MPT_FILTER_TRIANGLE;
      else if (n == "gaussian") em.filter = MPT_FILTER_GAUSSIAN;
      else if (n == "mitchell") em.filter = MPT_FILTER_MITCHELL;
      else if (n == "lanczosSinc") em.filter = MPT_FILTER_LANCZOS;
      else return bad("encountered unexpected filter: " + n);
    }
    if (num(e.find("support"), x)) {
      if (!(x >= 1)) return bad("filter support must be at least 1");
      em.support = (uint32_t)x;
    }
    if (em.filter == MPT_FILTER_GAUSSIAN && num(e.find("sigma"), x)) em.params[0] = (float)x;
    if (em.filter == MPT_FILTER_MITCHELL) {
      em.params[0] = 1.0f / 3.0f;
      em.params[1] = 1.0f / 3.0f;
      if (num(e.find("B"), x)) em.params[0] = (float)x;
      if (num(e.find("C"), x)) em.params[1] = (float)x;
    }
    if (em.filter == MPT_FILTER_LANCZOS && num(e.find("tau"), x)) em.params[0] = (float)x;
    std::vector<double> k;
    uint32_t m;
    std::string err;
    if (!mpt_env::kernel(em, k, m, err)) return bad(err);
    flat_.has_env_map = true;
    return true;
  }

  // Material::init and the BSDF / emitter parsers; defaults as the BSDFs
  // read them (DiffuseBSDF.cpp:45, ModifiedPhongBSDF.cpp:51-73,
  // DielectricBSDF.cpp:49-51,136).
  bool material(const JVal &v, int &out) {
    mpt_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = MPT_BSDF_DIFFUSE;
    for (int i = 0; i < 3; ++i) m.kd[i] = 0.5f;
    m.exponent = 1;
    m.ior = 1.0003f;
    m.transparency = 1;
    double ior[3] = {1.0003, 1.0003, 1.0003};
    bool have_ior = false;
    if (v.kind == JVal::OBJ) {
      const JVal *med = v.find("medium");
      if (med && med->kind == JVal::OBJ && spectrum(med->find("ior"), ior)) have_ior = true;
      const JVal *bsdf = v.find("bsdf");
      if (bsdf) {
        if (bsdf->kind != JVal::OBJ) return bad("bsdf must be an object");
        const JVal *type = bsdf->find("type");
        if (!type || type->kind != JVal::STR) return bad("bsdf needs a type");
        const std::string &t = type->str;
        double kd[3], ks[3], n[3], tr[3];
        const bool hkd = spectrum(bsdf->find("kd"), kd), hks = spectrum(bsdf->find("ks"), ks);
        if (spectrum(bsdf->find("ior"), ior)) have_ior = true;
        if (t == "diffuse") {
          m.type = MPT_BSDF_DIFFUSE;
          for (int i = 0; i < 3; ++i) m.kd[i] = hkd ? (float)kd[i] : 0.5f;
        } else if (t == "modifiedPhong" || t == "phong") {
          m.type = MPT_BSDF_PHONG;
          for (int i = 0; i < 3; ++i) {
            m.kd[i] = hkd ? (float)kd[i] : (hks ? 0.0f : 0.5f);
            m.ks[i] = hks ? (float)ks[i] : (hkd ? 0.0f : 0.5f);
          }
          if (spectrum(bsdf->find("n"), n)) {
            if (n[0] != n[1] || n[1] != n[2]) return bad("per-channel Phong exponents are not supported");
            m.exponent = (float)n[0];
          }
        } else if (t == "dielectric" || t == "specular" || t == "transmissive") {
          m.type = MPT_BSDF_DIELECTRIC;
          for (int i = 0; i < 3; ++i) {
            m.kd[i] = 0;
            m.ks[i] = hks ? (float)ks[i] : 1.0f;
          }
          if (t == "specular") m.transparency = 0;
          else if (t == "transmissive") m.transparency = 1;
          else if (spectrum(bsdf->find("transparency"), tr)) {
            if (tr[0] != tr[1] || tr[1] != tr[2]) return bad("per-channel transparency is not supported");
            m.transparency = (float)tr[0];
          }
        } else if (t == "absorbent") {
          m.type = MPT_BSDF_DIFFUSE;
          for (int i = 0; i < 3; ++i) m.kd[i] = 0;
        } else {
          return bad("bsdf type '" + t + "' is not supported");
        }
      }
      if (have_ior) {
        m.ior = (float)ior[0];
        if (ior[0] != ior[1] || ior[1] != ior[2]) {   // a dispersive IOR spectrum
          m.ior_g = (float)ior[1];
          m.ior_b = (float)ior[2];
        }
      }
      int emits = 0;
      double power[3] = {0, 0, 0};
      const JVal *em = v.find("emitter");
      if (em) {
        if (em->kind != JVal::OBJ) return bad("emitter must be an object");
        const JVal *type = em->find("type");
        if (!type || type->kind != JVal::STR) return bad("emitter needs a type");
        if (type->str == "oriented" || type->str == "omni") {
          if (spectrum(em->find("power"), power) && (power[0] > 0 || power[1] > 0 || power[2] > 0))
            emits = type->str == "omni" ? 2 : 1;
        } else if (type->str != "null") {
          return bad("emitter type '" + type->str + "' is not supported (oriented, omni or null)");
        }
      }
      flat_.mat_emits.push_back(emits);
      for (int i = 0; i < 3; ++i) flat_.mat_power.push_back(power[i]);
    } else {
      flat_.mat_emits.push_back(0);
      for (int i = 0; i < 3; ++i) flat_.mat_power.push_back(0);
    }
    out = (int)flat_.mats.size();
    flat_.mats.push_back(m);
    return true;
  }

  uint32_t need_mat(int mat) {
    if (mat >= 0) return (uint32_t)mat;
    // a shape outside every material gets Material's defaults (diffuse 0.5)
    if (default_mat_ < 0) {
      JVal empty;
      material(empty, default_mat_);
    }
    return (uint32_t)default_mat_;
  }

  void add_vertex(const double p[3]) {
    for (int i = 0; i < 3; ++i) flat_.verts.push_back((float)p[i]);
    for (int i = 0; i < 3; ++i) flat_.normals.push_back(0.0f);
    flat_.vert_has_normal.push_back(0);
  }

  // One triangle with its own three vertices: the standalone Triangle's
  // normal is its face normal (Triangle.cpp:305-311).
  bool one_sided(const double a[3], const double b[3], const double c[3], uint32_t mat) {
    // a standalone Triangle samples its light by texture coordinates
    // (Triangle::getPosition, shapes/Triangle.cpp:330-353)
    if (flat_.mat_emits[mat]) return bad("emitting triangles are not supported (planes, spheres and meshes may emit)");
    const uint32_t base = (uint32_t)(flat_.verts.size() / 3);
    add_vertex(a);
    add_vertex(b);
    add_vertex(c);
    flat_.idx.insert(flat_.idx.end(), {base, base + 1, base + 2});
    flat_.tri_mat.push_back(mat);
    flat_.tri_flags.push_back(MPT_TRI_ONE_SIDED);
    flat_.tri_shape.push_back(flat_.shapes++);
    return true;
  }

  // The unit plane through o spanned by x and z (|u|, |v| <= 0.5) under M, as
  // a quad whose normal z x x is M's normal transform of the object's
  // (Transformable: the inverse transpose), so a mirroring transform
  // (negative determinant) swaps the spanning vectors.
  bool quad(const Mat &M, const double o[3], const double x[3], const double z[3], uint32_t m, const char *what) {
    if (flat_.mat_emits[m] == 2) return bad("an omni emitter needs a point shape");
    double c[3], ex[3], ez[3];
    M.point(o, c);
    M.vector(x, ex);
    M.vector(z, ez);
    const double dotxz = ex[0] * ez[0] + ex[1] * ez[1] + ex[2] * ez[2];
    const double lx = std::sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
    const double lz = std::sqrt(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
    if (!(lx > 0 && lz > 0) || std::fabs(dotxz) > 1e-9 * lx * lz)
      return bad(std::string("a sheared ") + what + " is not supported (its transform must keep the face axes orthogonal)");
    const double det = M.m[0][0] * (M.m[1][1] * M.m[2][2] - M.m[1][2] * M.m[2][1]) -
                       M.m[0][1] * (M.m[1][0] * M.m[2][2] - M.m[1][2] * M.m[2][0]) +
                       M.m[0][2] * (M.m[1][0] * M.m[2][1] - M.m[1][1] * M.m[2][0]);
    const double *qx = det < 0 ? ez : ex, *qz = det < 0 ? ex : ez;
    for (int i = 0; i < 3; ++i) flat_.quads.push_back((float)c[i]);
    for (int i = 0; i < 3; ++i) flat_.quads.push_back((float)qx[i]);
    for (int i = 0; i < 3; ++i) flat_.quads.push_back((float)qz[i]);
    flat_.quad_mat.push_back(m);
    return true;
  }

  bool shape(const std::string &type, const JVal &v, const Mat &M, int mat_in) {
    if (v.kind != JVal::OBJ) return bad(type + " must be an object");
    // a material nested in the shape object applies to the shape
    int mat = mat_in;
    if (const JVal *mv = v.find("material")) {
      if (!material(*mv, mat)) return false;
    }
    const uint32_t m = need_mat(mat);
    if (type == "plane") {
      double d[12];
      if (v.find("vertices")) {
        if (!vec(v.find("vertices"), d, 12)) return bad("plane.vertices expects 12 numbers");
        double w[4][3];
        for (int i = 0; i < 4; ++i) M.point(d + 3 * i, w[i]);
        return one_sided(w[0], w[1], w[2], m) && one_sided(w[0], w[2], w[3], m);
      }
      const double o[3] = {0, 0, 0}, x[3] = {1, 0, 0}, z[3] = {0, 0, 1};
      return quad(M, o, x, z, m, "plane");
    }
    if (type == "point") {
      // PointShape (shapes/PointShape.cpp): the transformed origin; it is
      // never hit, so only an omni emitter on it changes the image
      if (flat_.mat_emits[m] == 1) return bad("oriented emitters on points are not supported (omni only)");
      if (flat_.mat_emits[m] == 2) {
        const double o[3] = {0, 0, 0};
        double w[3];
        M.point(o, w);
        for (int i = 0; i < 3; ++i) flat_.points.push_back((float)w[i]);
        for (int i = 0; i < 3; ++i) flat_.points.push_back((float)flat_.mat_power[3 * m + i]);
      }
      return true;
    }
    if (type == "cube") {
      // Cube::getIntersection (shapes/Cube.cpp:17-52): the unit cube
      // [-0.5, 0.5]^3 hit on its six faces p[u] = +-0.5 with |p[v]|, |p[w]|
      // <= 0.5 from either side, i.e. six unit planes with outward normals.
      if (flat_.mat_emits[m]) return bad("emitting cubes are not supported (planes, spheres and meshes may emit)");
      for (int u = 0; u < 3; ++u) {
        const int a = (u + 1) % 3, b = (u + 2) % 3;
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          double o[3] = {0, 0, 0}, x[3] = {0, 0, 0}, z[3] = {0, 0, 0};
          o[u] = 0.5 * sgn;
          // the quad's normal is z x x: e_u for (x, z) = (e_b, e_a), -e_u swapped
          x[sgn > 0 ? b : a] = 1;
          z[sgn > 0 ? a : b] = 1;
          if (!quad(M, o, x, z, m, "cube")) return false;
        }
      }
      return true;
    }
    if (type == "triangle") {
      double d[9] = {0, 0, 0, 0, 0, 0.5, 0.5, 0, 0};   // the reference defaults
      if (v.find("vertices") && !vec(v.find("vertices"), d, 9)) return bad("triangle.vertices expects 9 numbers");
      double w[3][3];
      for (int i = 0; i < 3; ++i) M.point(d + 3 * i, w[i]);
      return one_sided(w[0], w[1], w[2], m);
    }
    if (type == "sphere") {
      if (flat_.mat_emits[m] == 2) return bad("an omni emitter needs a point shape");
      Mat T = M;
      double pos[3], r;
      if (vec(v.find("position"), pos, 3)) T = T * translate(pos);
      if (num(v.find("radius"), r)) {
        const double s[3] = {2 * r, 2 * r, 2 * r};
        T = T * scale(s);
      }
      const double o[3] = {0, 0, 0};
      double c[3], ax[3][3];
      T.point(o, c);
      for (int k = 0; k < 3; ++k) {
        const double e[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
        T.vector(e, ax[k]);
      }
      double len[3];
      for (int k = 0; k < 3; ++k) len[k] = std::sqrt(ax[k][0] * ax[k][0] + ax[k][1] * ax[k][1] + ax[k][2] * ax[k][2]);
      for (int k = 0; k < 3; ++k)
        for (int j = k + 1; j < 3; ++j) {
          const double dp = ax[k][0] * ax[j][0] + ax[k][1] * ax[j][1] + ax[k][2] * ax[j][2];
          if (std::fabs(dp) > 1e-9 * len[k] * len[j]) return bad("a sheared sphere is not supported");
        }
      if (std::fabs(len[0] - len[1]) > 1e-9 * len[0] || std::fabs(len[0] - len[2]) > 1e-9 * len[0])
        return bad("a non-uniformly scaled sphere (ellipsoid) is not supported");
      for (int i = 0; i < 3; ++i) flat_.spheres.push_back((float)c[i]);
      flat_.spheres.push_back((float)(0.5 * len[0]));
      flat_.sphere_mat.push_back(m);
      // the transform's linear part places an emitting sphere's light
      // samples (Sphere::getPosition, shapes/Sphere.cpp:117-124)
      for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) flat_.sphere_frames.push_back((float)ax[k][i]);
      return true;
    }
    if (type == "mesh") {
      if (flat_.mat_emits[m] == 2) return bad("an omni emitter needs a point shape");
      const JVal *path = v.find("path");
      if (!path || path->kind != JVal::STR) return bad("mesh needs a path");
      const JVal *norm = v.find("normalize");
      return mesh(path->str, norm && norm->kind == JVal::BOOL && norm->b, M, m);
    }
    return bad("shape type '" + type + "' is not supported by this path tracer");
  }

  // OBJ (v, vn, f with v/vt/vn, polygons as fans) or ASCII PLY.
  bool mesh(const std::string &rel, bool normalize, const Mat &M, uint32_t mat) {
    const std::string path = (!rel.empty() && rel[0] == '/') || base_.empty() ? rel : base_ + "/" + rel;
    std::ifstream in(path);
    if (!in) return bad("mesh '" + path + "' cannot be opened");
    std::vector<double> P, N;
    std::vector<uint32_t> F, FN;   // triangle vertex indices and their normal indices (or ~0u)
    const std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : "";
    std::string line;
    if (ext == ".obj" || ext == ".OBJ") {
      while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") {
          double x, y, z;
          ls >> x >> y >> z;
          P.insert(P.end(), {x, y, z});
        } else if (tag == "vn") {
          double x, y, z;
          ls >> x >> y >> z;
          N.insert(N.end(), {x, y, z});
        } else if (tag == "f") {
          std::vector<long> vi, ni;
          std::string tok;
          while (ls >> tok) {
            long a = 0, c = 0;
            const size_t s1 = tok.find('/');
            a = std::strtol(tok.c_str(), nullptr, 10);
            if (s1 != std::string::npos) {
              const size_t s2 = tok.find('/', s1 + 1);
              if (s2 != std::string::npos) c = std::strtol(tok.c_str() + s2 + 1, nullptr, 10);
            }
            const long nv = (long)(P.size() / 3), nn = (long)(N.size() / 3);
            vi.push_back(a < 0 ? nv + a : a - 1);
            ni.push_back(c == 0 ? -1 : (c < 0 ? nn + c : c - 1));
          }
          for (size_t k = 1; k + 1 < vi.size(); ++k) {
            const long t[3] = {vi[0], vi[k], vi[k + 1]}, tn[3] = {ni[0], ni[k], ni[k + 1]};
            for (int j = 0; j < 3; ++j) {
              if (t[j] < 0 || t[j] >= (long)(P.size() / 3)) return bad("mesh '" + path + "': face index out of range");
              F.push_back((uint32_t)t[j]);
              FN.push_back(tn[j] >= 0 && tn[j] < (long)(N.size() / 3) ? (uint32_t)tn[j] : ~0u);
            }
          }
        }
      }
    } else if (ext == ".ply" || ext == ".PLY") {
      size_t nv = 0, nf = 0;
      std::vector<std::string> props;
      bool in_vertex = false;
      if (!std::getline(in, line) || line.compare(0, 3, "ply") != 0) return bad("mesh '" + path + "' is not PLY");
      while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string a, b, c;
        ls >> a;
        if (a == "format") {
          ls >> b;
          if (b != "ascii") return bad("mesh '" + path + "': only ASCII PLY is supported");
        } else if (a == "element") {
          ls >> b >> c;
          in_vertex = b == "vertex";
          if (b == "vertex") nv = std::strtoul(c.c_str(), nullptr, 10);
          if (b == "face") nf = std::strtoul(c.c_str(), nullptr, 10);
        } else if (a == "property" && in_vertex) {
          std::string last;
          while (ls >> b) last = b;
          props.push_back(last);
        } else if (a == "end_header") {
          break;
        }
      }
      int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1;
      for (size_t k = 0; k < props.size(); ++k) {
        const std::string &p = props[k];
        if (p == "x") ix = (int)k;
        if (p == "y") iy = (int)k;
        if (p == "z") iz = (int)k;
        if (p == "nx") inx = (int)k;
        if (p == "ny") iny = (int)k;
        if (p == "nz") inz = (int)k;
      }
      if (ix < 0 || iy < 0 || iz < 0) return bad("mesh '" + path + "': PLY vertices need x, y, z");
      const bool has_n = inx >= 0 && iny >= 0 && inz >= 0;
      for (size_t i = 0; i < nv; ++i) {
        std::vector<double> vals(props.size());
        for (double &x : vals) in >> x;
        P.insert(P.end(), {vals[ix], vals[iy], vals[iz]});
        if (has_n) N.insert(N.end(), {vals[inx], vals[iny], vals[inz]});
      }
      for (size_t f = 0; f < nf; ++f) {
        size_t k;
        in >> k;
        std::vector<long> vi(k);
        for (long &x : vi) in >> x;
        for (size_t j = 1; j + 1 < k; ++j) {
          const long t[3] = {vi[0], vi[j], vi[j + 1]};
          for (int q = 0; q < 3; ++q) {
            if (t[q] < 0 || t[q] >= (long)nv) return bad("mesh '" + path + "': face index out of range");
            F.push_back((uint32_t)t[q]);
            FN.push_back(has_n ? (uint32_t)t[q] : ~0u);
          }
        }
      }
      if (!in) return bad("mesh '" + path + "': truncated PLY data");
    } else {
      return bad("mesh '" + path + "': unknown format (.obj or .ply)");
    }
    if (F.empty()) return bad("mesh '" + path + "' has no faces");
    const size_t nv = P.size() / 3;
    if (normalize) {
      // into the unit cube [-0.5, 0.5]^3 per axis (MiltonJSONSceneLoader.cpp:496-524)
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (size_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], P[3 * i + k]);
          hi[k] = std::max(hi[k], P[3 * i + k]);
        }
      for (size_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
          double dg = hi[k] - lo[k];
          dg += dg == 0;
          P[3 * i + k] = (P[3 * i + k] - 0.5 * (lo[k] + hi[k])) / dg;
        }
    }
    // A vertex keeps one normal: the file's (from its first face corner that
    // names one), else the averaged face normals computed at scene creation.
    const uint32_t base = (uint32_t)(flat_.verts.size() / 3);
    for (size_t i = 0; i < nv; ++i) {
      double w[3];
      M.point(&P[3 * i], w);
      add_vertex(w);
    }
    // normals transform by the inverse transpose; for the similarity
    // transforms of scene files that is the linear part up to scale
    Mat NM = M;
    for (size_t f = 0; f < F.size(); ++f) {
      const uint32_t v = base + F[f];
      if (FN[f] != ~0u && !flat_.vert_has_normal[v]) {
        double n[3];
        NM.vector(&N[3 * FN[f]], n);
        const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0) {
          for (int k = 0; k < 3; ++k) flat_.normals[3 * v + k] = (float)(n[k] / l);
          flat_.vert_has_normal[v] = 1;
          flat_.any_normals = true;
        }
      }
    }
    const uint32_t shape = flat_.shapes++;
    for (size_t f = 0; f + 2 < F.size(); f += 3) {
      flat_.idx.insert(flat_.idx.end(), {base + F[f], base + F[f + 1], base + F[f + 2]});
      flat_.tri_mat.push_back(mat);
      flat_.tri_flags.push_back(0);
      flat_.tri_shape.push_back(shape);
    }
    return true;
  }

  std::string base_;
  std::string err_;
  int default_mat_ = -1;
};

}  // namespace

struct mpt_scenefile {
  Flat flat;
  Settings set;
  std::vector<uint8_t> vert_normal_mask;
  mpt_scene_desc desc;
};

namespace {

int finish(Loader &ld, mpt_scenefile **out, char *err, size_t err_len) {
  std::unique_ptr<mpt_scenefile> sf(new (std::nothrow) mpt_scenefile());
  if (!sf) return MPT_E_NOMEM;
  sf->flat = std::move(ld.flat_);
  sf->set = ld.set_;
  Flat &f = sf->flat;
  // emitting quads: Le = power / (pi * area) (Emitter.cpp:37-47,
  // OrientedEmitter.cpp:17-23); one material per emitting quad
  std::vector<mpt_material> mats = f.mats;
  for (size_t q = 0; q < f.quad_mat.size(); ++q) {
    const uint32_t m = f.quad_mat[q];
    if (!f.mat_emits[m]) continue;
    const float *ex = &f.quads[9 * q + 3], *ez = &f.quads[9 * q + 6];
    const double nx = (double)ez[1] * ex[2] - (double)ez[2] * ex[1], ny = (double)ez[2] * ex[0] - (double)ez[0] * ex[2],
                 nz = (double)ez[0] * ex[1] - (double)ez[1] * ex[0];
    const double area = std::sqrt(nx * nx + ny * ny + nz * nz);
    mpt_material em = f.mats[m];
    for (int i = 0; i < 3; ++i) em.Le[i] = (float)(f.mat_power[3 * m + i] / (M_PI * area));
    f.quad_mat[q] = (uint32_t)mats.size();
    mats.push_back(em);
  }
  // emitting spheres: area pi whatever the scale (Sphere::_getSurfaceArea,
  // shapes/Sphere.cpp:138-141)
  for (size_t i = 0; i < f.sphere_mat.size(); ++i) {
    const uint32_t m = f.sphere_mat[i];
    if (!f.mat_emits[m]) continue;
    mpt_material em = f.mats[m];
    for (int k = 0; k < 3; ++k) em.Le[k] = (float)(f.mat_power[3 * m + k] / (M_PI * M_PI));
    f.sphere_mat[i] = (uint32_t)mats.size();
    mats.push_back(em);
  }
  // emitting meshes: their world-space area (Mesh::_getSurfaceArea,
  // shapes/Mesh.cpp:452-466), one material per mesh
  {
    std::vector<double> area(f.shapes, 0.0);
    for (size_t t = 0; t < f.tri_mat.size(); ++t) {
      if (!f.mat_emits[f.tri_mat[t]]) continue;
      const float *a = &f.verts[3 * f.idx[3 * t]], *b = &f.verts[3 * f.idx[3 * t + 1]], *c = &f.verts[3 * f.idx[3 * t + 2]];
      const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
      const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
      const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
      area[f.tri_shape[t]] += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    }
    std::vector<uint32_t> shape_mat(f.shapes, ~0u);
    for (size_t t = 0; t < f.tri_mat.size(); ++t) {
      const uint32_t m = f.tri_mat[t], sh = f.tri_shape[t];
      if (!f.mat_emits[m]) continue;
      if (shape_mat[sh] == ~0u) {
        mpt_material em = f.mats[m];
        for (int k = 0; k < 3; ++k) em.Le[k] = (float)(area[sh] > 0 ? f.mat_power[3 * m + k] / (M_PI * area[sh]) : 0);
        shape_mat[sh] = (uint32_t)mats.size();
        mats.push_back(em);
      }
      f.tri_mat[t] = shape_mat[sh];
    }
  }
  f.mats = std::move(mats);
  mpt_scene_desc &d = sf->desc;
  std::memset(&d, 0, sizeof(d));
  d.version = MPT_ABI_VERSION;
  d.num_vertices = (uint32_t)(f.verts.size() / 3);
  d.vertices = f.verts.data();
  d.num_triangles = (uint32_t)f.tri_mat.size();
  d.indices = f.idx.data();
  d.tri_material = f.tri_mat.data();
  d.num_spheres = (uint32_t)f.sphere_mat.size();
  d.spheres = f.spheres.data();
  d.sphere_material = f.sphere_mat.data();
  d.num_quads = (uint32_t)f.quad_mat.size();
  d.quads = f.quads.data();
  d.quad_material = f.quad_mat.data();
  d.num_materials = (uint32_t)f.mats.size();
  d.materials = f.mats.data();
  for (int i = 0; i < 3; ++i) d.env_radiance[i] = f.env[i];
  if (f.has_env_map) {
    f.env_map.rgb = f.env_rgb.data();
    d.env_map = &f.env_map;
  }
  d.tri_flags = f.tri_flags.data();
  d.tri_shape = f.tri_shape.data();
  d.sphere_frames = f.sphere_frames.data();
  d.vertex_normals = f.any_normals ? f.normals.data() : nullptr;
  d.vertex_normal_mask = f.any_normals ? f.vert_has_normal.data() : nullptr;
  d.num_points = (uint32_t)(f.points.size() / 6);
  d.points = f.points.data();
  *out = sf.release();
  (void)err;
  (void)err_len;
  return MPT_OK;
}

void set_err(char *err, size_t err_len, const std::string &msg) {
  if (err && err_len) {
    std::snprintf(err, err_len, "%s", msg.c_str());
  }
}

}  // namespace

extern "C" {

int mpt_scenefile_parse(const char *text, const char *base_dir, mpt_scenefile **out, char *err, size_t err_len) {
  if (!text || !out) return MPT_E_INVALID;
  *out = nullptr;
  try {
    const std::string s(text);
    JVal root;
    std::string e;
    JParser p(s);
    if (!p.parse(root, e)) {
      set_err(err, err_len, "scene file syntax: " + e);
      return MPT_E_INVALID;
    }
    Loader ld(base_dir ? base_dir : "");
    if (!ld.load(root, e)) {
      set_err(err, err_len, e);
      return MPT_E_INVALID;
    }
    return finish(ld, out, err, err_len);
  } catch (const std::bad_alloc &) {
    set_err(err, err_len, "out of host memory");
    return MPT_E_NOMEM;
  }
}

int mpt_scenefile_load(const char *path, mpt_scenefile **out, char *err, size_t err_len) {
  if (!path || !out) return MPT_E_INVALID;
  *out = nullptr;
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    set_err(err, err_len, std::string("cannot open ") + path);
    return MPT_E_INVALID;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  std::string dir(path);
  const size_t slash = dir.rfind('/');
  dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
  return mpt_scenefile_parse(ss.str().c_str(), dir.c_str(), out, err, err_len);
}

const mpt_scene_desc *mpt_scenefile_desc(const mpt_scenefile *sf) { return sf ? &sf->desc : nullptr; }

uint32_t mpt_scenefile_tonemap(const mpt_scenefile *sf) { return sf ? sf->set.tonemap : MPT_TONEMAP_REINHARD; }

int mpt_scenefile_job(const mpt_scenefile *sf, mpt_job *job) {
  if (!sf || !job) return MPT_E_INVALID;
  std::memset(job, 0, sizeof(*job));
  const Settings &s = sf->set;
  for (int i = 0; i < 3; ++i) {
    job->camera.eye[i] = (float)s.eye[i];
    job->camera.look[i] = (float)s.look[i];
    job->camera.up[i] = (float)s.up[i];
  }
  job->camera.height_angle = (float)s.height_angle;
  job->camera.aspect = (float)s.aspect;
  job->camera.aperture = (float)s.aperture;
  job->camera.focal_length = (float)s.focal_length;
  job->camera.focal_distance = (float)s.focal_distance;
  job->camera.autofocus = s.autofocus ? 1u : 0u;
  job->camera.focal_point[0] = (float)s.focal_point[0];
  job->camera.focal_point[1] = (float)s.focal_point[1];
  job->width = s.width;
  job->height = s.height;
  job->spp = s.spp;
  job->max_depth = s.max_depth;
  job->seed = s.seed;
  job->direct_samples = s.direct;
  job->direct_off = s.direct_off;
  job->renderer = s.renderer;
  job->mlt_chains = s.mlt_chains;
  job->mlt_seed_paths = s.mlt_seed_paths;
  job->mlt_max_depth = s.mlt_max_depth;
  job->mlt_max_rejections = s.mlt_max_rejections;
  job->mlt_flags = s.mlt_flags;
  for (int k = 0; k < 3; ++k) job->mlt_probs[k] = s.mlt_probs[k];
  job->mlt_mutations = s.mlt_mutations;
  if (!job->mlt_mutations) {
    const double per_chain = (double)s.spp * s.width * s.height / std::max(1u, s.mlt_chains);
    job->mlt_mutations = (uint32_t)std::max(1.0, std::min(32767.0, per_chain));
  }
  job->filter = s.filter;
  job->filter_support = s.filter_support;
  job->filter_params[0] = s.filter_params[0];
  job->filter_params[1] = s.filter_params[1];
  return MPT_OK;
}

int mpt_scenefile_destroy(mpt_scenefile *sf) {
  delete sf;
  return MPT_OK;
}

}  // extern "C"
