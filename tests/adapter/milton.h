/*
 * Test fixture, not part of the library: the small subset of Milton's C++
 * interface that the Milton-side adapter (milton_amd/adapter/
 * milton_amd_renderer.cpp) uses, declared with Milton's class and member
 * names so the adapter compiles and runs here without Milton's own headers
 * (which need boost and Qt4).  Bodies are minimal stand-ins written for the
 * test: property maps hold std::any values, matrices are 4x4 doubles, and
 * shapes carry only the data the adapter reads.  Members mirror
 * renderers/Renderer.h, renderers/PointSampleRenderer.h,
 * renderers/RenderOutput.h, core/Scene.h, shapes/{ShapeSet,Shape,
 * Transformable,Mesh,MeshTriangle,Triangle,Sphere,Plane,PointShape}.h,
 * materials/{Material,Emitter}.h and cameras/{Camera,ThinLensCamera,
 * PinholeCamera}.h of the reference (paths under mlt/src/milton/).
 */
#ifndef ADAPTER_TEST_MILTON_H_
#define ADAPTER_TEST_MILTON_H_

#include <any>
#include <map>
#include <string>
#include <vector>

namespace milton {

typedef double real_t;

struct Vector3 {
  double data[3] = {0, 0, 0};
  Vector3() = default;
  Vector3(double x, double y, double z) : data{x, y, z} {}
  double operator[](int i) const { return data[i]; }
  double &operator[](int i) { return data[i]; }
};
typedef Vector3 Vertex;
typedef Vector3 Normal;

struct Point3 {   // homogeneous, w = 1
  double data[4] = {0, 0, 0, 1};
  Point3() = default;
  Point3(double x, double y, double z) : data{x, y, z, 1} {}
  double operator[](int i) const { return data[i]; }
};

struct Matrix4x4 {
  double m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  Point3 operator*(const Point3 &p) const {
    Point3 r;
    for (int i = 0; i < 3; ++i) r.data[i] = m[4 * i] * p[0] + m[4 * i + 1] * p[1] + m[4 * i + 2] * p[2] + m[4 * i + 3];
    return r;
  }
  Vector3 operator*(const Vector3 &v) const {
    Vector3 r;
    for (int i = 0; i < 3; ++i) r.data[i] = m[4 * i] * v[0] + m[4 * i + 1] * v[1] + m[4 * i + 2] * v[2];
    return r;
  }
};

struct SpectralSample {
  double value = 0;
};
struct SpectralSampleSet {
  SpectralSample s[3];
  unsigned getN() const { return 3; }
  const SpectralSample &operator[](unsigned i) const { return s[i]; }
  SpectralSample &operator[](unsigned i) { return s[i]; }
  static SpectralSampleSet fill(double v) {
    SpectralSampleSet r;
    for (auto &x : r.s) x.value = v;
    return r;
  }
  static SpectralSampleSet black() { return fill(0); }
};

class PropertyMap {
 public:
  virtual ~PropertyMap() = default;
  template <class T>
  T getValue(const std::string &key, const T &fallback) const {
    const auto it = props_.find(key);
    if (it == props_.end()) return fallback;
    return std::any_cast<T>(it->second);
  }
  bool contains(const std::string &key) const { return props_.count(key) != 0; }
  std::any &operator[](const std::string &key) { return props_[key]; }

 private:
  std::map<std::string, std::any> props_;
};

class Material : public PropertyMap {};

class Emitter {
 public:
  explicit Emitter(Material *parent) : m_parent(parent) {}
  Material *getParent() { return m_parent; }

 private:
  Material *m_parent;
};

class Shape {
 public:
  virtual ~Shape() = default;
  Material *getMaterial() { return m_material; }
  void setMaterial(Material *material) { m_material = material; }

 private:
  Material *m_material = nullptr;
};

class Transformable : public Shape {
 public:
  const Matrix4x4 &getTransToWorld() const { return m_transToWorld; }
  void setTransToWorld(const Matrix4x4 &m) { m_transToWorld = m; }

 private:
  Matrix4x4 m_transToWorld;
};

class ShapeSet : public Shape {
 public:
  unsigned size() const { return (unsigned)m_shapes.size(); }
  Shape *operator[](unsigned i) { return m_shapes[i]; }
  void push_back(Shape *s) { m_shapes.push_back(s); }

 private:
  std::vector<Shape *> m_shapes;
};

struct MeshTriangle {
  unsigned A, B, C;
};

class Mesh : public Transformable, public PropertyMap {
 public:
  std::vector<Vertex> vertices;
  std::vector<MeshTriangle> triangles;
  std::vector<Normal> normals;   // per vertex, or empty
  Vertex *getVertices(unsigned &n) const {
    n = (unsigned)vertices.size();
    return const_cast<Vertex *>(vertices.data());
  }
  MeshTriangle *getTriangles(unsigned &n) const {
    n = (unsigned)triangles.size();
    return const_cast<MeshTriangle *>(triangles.data());
  }
  Normal *getNormals(unsigned &n) const {
    n = (unsigned)normals.size();
    return const_cast<Normal *>(normals.data());
  }
};

class Triangle : public Transformable {
 public:
  Vertex vertices[3];
};

class Implicit : public Transformable {};
class Sphere : public Implicit {};   // unit sphere of radius 0.5 in object space
class Plane : public Implicit {};    // unit square y = 0, |x|, |z| <= 0.5 in object space
class PointShape : public Transformable {};

class Scene {
 public:
  ShapeSet *getShapes() { return &m_shapes; }
  Emitter *getBackground() { return m_background; }
  void setBackground(Emitter *e) { m_background = e; }

 private:
  ShapeSet m_shapes;
  Emitter *m_background = nullptr;
};

class Camera : public Shape, public PropertyMap {};

class ThinLensCamera : public Camera {
 public:
  Point3 getEye() const { return m_eye; }
  Vector3 getLook() const { return m_look; }
  Vector3 getUp() const { return m_up; }
  void setOrientation(const Point3 &eye, const Vector3 &look, const Vector3 &up) {
    m_eye = eye;
    m_look = look;
    m_up = up;
  }

 private:
  Point3 m_eye;
  Vector3 m_look{0, 0, -1}, m_up{0, 1, 0};
};

class PinholeCamera : public ThinLensCamera {};

struct Point2 {
  double data[3] = {0, 0, 1};
  double &operator[](int i) { return data[i]; }
  double operator[](int i) const { return data[i]; }
};

struct PointSample {
  Point2 position;
  SpectralSampleSet value;
};

class RenderOutput : public PropertyMap {
 public:
  RenderOutput(unsigned w, unsigned h) : m_w(w), m_h(h), m_rgb(3ull * w * h, 0.0) {}
  unsigned getWidth() const { return m_w; }
  unsigned getHeight() const { return m_h; }
  // pixel-bin film: the sample's value lands in the pixel its position is in
  virtual void addSample(PointSample &s) {
    const unsigned x = (unsigned)(s.position[0] * m_w), y = (unsigned)(s.position[1] * m_h);
    for (int k = 0; k < 3; ++k) m_rgb[3ull * (y * m_w + x) + k] = s.value[k].value;
  }
  virtual void finalize() { m_finalized = true; }
  const std::vector<double> &pixels() const { return m_rgb; }
  bool finalized() const { return m_finalized; }

 private:
  unsigned m_w, m_h;
  std::vector<double> m_rgb;
  bool m_finalized = false;
};

class Renderer : public PropertyMap {
 public:
  virtual ~Renderer() = default;
  virtual void init() {}
  virtual void render() = 0;
  void setCamera(Camera *c) { m_camera = c; }
  void setScene(Scene *s) { m_scene = s; }

 protected:
  Camera *m_camera = nullptr;
  Scene *m_scene = nullptr;
};

class PointSampleRenderer : public Renderer {
 public:
  virtual RenderOutput *getOutput() { return m_output; }
  virtual void setOutput(RenderOutput *output) { m_output = output; }

 protected:
  RenderOutput *m_output = nullptr;
};

}  // namespace milton

enum MILTON_PLUGIN_TYPE { MILTON_PLUGIN_TYPE_RENDERER = 0 };

#endif
