// refcpu — the oracle. TEST INFRASTRUCTURE ONLY.
//
// A plain C++17 f64 restatement of the zigtrace CPU path-tracing core
// (src/vec.zig, ray.zig, aabb.zig, rng.zig, geometry.zig, material.zig,
// camera.zig, bvh.zig, light.zig, scene.zig, render.zig, scenefile.zig).
// It exists to check the product (the HIP library under zigtrace_amd/) and
// to serve as the CPU throughput baseline. Only tests/, the smoke check and
// the benchmark's CPU-baseline leg may load it; the product never does.
//
// Numeric conventions copied from the reference (Zig 0.14, strict float mode):
//   * no FMA contraction (built with -ffp-contract=off, no fast-math);
//   * @min/@max on floats are minnum/maxnum -> fmin/fmax (NaN-ignoring);
//   * Vec3.div(s) multiplies by the reciprocal (vec.zig:38-40);
//   * @intFromFloat truncates; std.mem.sort is stable (std::stable_sort);
//   * comptime constants such as 1 + 1e-9 are folded exactly, then rounded;
//   * std.math.pow(f64, x, 5) uses frexp-normalised repeated squaring, whose
//     rounding equals x * ((x*x) * (x*x)) for normal inputs.
// Transcendentals (sin, cos, tan, sqrt) come from the C library here and
// from Zig's compiler_rt in the reference: images are therefore not
// bit-pinned against the Zig binary (sqrt is correctly rounded in both).
//
// Every function cites the reference file:line it restates.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- vec.zig:5-125
struct V3 {
  double x, y, z;
};
static inline V3 v3(double x, double y, double z) { return {x, y, z}; }
static inline V3 splat(double v) { return {v, v, v}; }
static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 vdiv(V3 a, double s) { return scale(a, 1.0 / s); }  // vec.zig:38-40
static inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline double len2(V3 a) { return dot(a, a); }
static inline double length(V3 a) { return std::sqrt(len2(a)); }
static inline V3 normalize(V3 a) { return vdiv(a, length(a)); }  // vec.zig:66-70
static inline V3 lerp(V3 a, V3 b, double t) { return add(scale(a, 1.0 - t), scale(b, t)); }
static inline V3 vmin(V3 a, V3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
static inline V3 vmax(V3 a, V3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
static inline double maxComponent(V3 a) { return std::fmax(a.x, std::fmax(a.y, a.z)); }
static inline double get(V3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }
static inline bool nearZero(V3 a) {  // vec.zig:91-94
  const double eps = 1e-8;
  return std::fabs(a.x) < eps && std::fabs(a.y) < eps && std::fabs(a.z) < eps;
}
static inline bool isFinite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
static inline V3 reflect(V3 v, V3 n) { return sub(v, scale(n, 2.0 * dot(v, n))); }  // vec.zig:112-114
static inline V3 refract(V3 uv, V3 n, double eta) {                              // vec.zig:119-125
  const double cos_theta = std::fmin(dot(neg(uv), n), 1.0);
  const V3 r_perp = scale(add(uv, scale(n, cos_theta)), eta);
  const V3 r_par = scale(n, -std::sqrt(std::fabs(1.0 - len2(r_perp))));
  return add(r_perp, r_par);
}

// ---------------------------------------------------------------- ray.zig:4-14
struct Ray {
  V3 o, d;
};
static inline V3 at(const Ray& r, double t) { return add(r.o, scale(r.d, t)); }

// ---------------------------------------------------------------- aabb.zig:7-88
struct Aabb {
  V3 mn, mx;
};
static const double kInf = std::numeric_limits<double>::infinity();
static inline Aabb aabbEmpty() { return {splat(kInf), splat(-kInf)}; }                 // aabb.zig:12-15
static inline Aabb fromPoints(V3 a, V3 b) { return {vmin(a, b), vmax(a, b)}; }          // aabb.zig:17-19
static inline Aabb merge(const Aabb& a, const Aabb& b) { return {vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }
static inline Aabb include(const Aabb& a, V3 p) { return {vmin(a.mn, p), vmax(a.mx, p)}; }
static inline Aabb pad(Aabb a, double delta) {  // aabb.zig:30-39
  if (a.mx.x - a.mn.x < delta) { a.mn.x -= delta / 2; a.mx.x += delta / 2; }
  if (a.mx.y - a.mn.y < delta) { a.mn.y -= delta / 2; a.mx.y += delta / 2; }
  if (a.mx.z - a.mn.z < delta) { a.mn.z -= delta / 2; a.mx.z += delta / 2; }
  return a;
}
static inline bool isEmpty(const Aabb& a) { return a.mn.x > a.mx.x || a.mn.y > a.mx.y || a.mn.z > a.mx.z; }
static inline V3 centroid(const Aabb& a) { return scale(add(a.mn, a.mx), 0.5); }
static inline double surfaceArea(const Aabb& a) {  // aabb.zig:53-57
  if (isEmpty(a)) return 0;
  const V3 e = sub(a.mx, a.mn);
  return 2.0 * (e.x * e.y + e.y * e.z + e.z * e.x);
}
static inline int longestAxis(const Aabb& a) {  // aabb.zig:59-63
  const V3 e = sub(a.mx, a.mn);
  if (e.x >= e.y && e.x >= e.z) return 0;
  return e.y >= e.z ? 1 : 2;
}
// aabb.zig:68-84. Returns false for a miss; NaN comparisons leave the
// interval unchanged exactly as the reference's `if (t0 > lo)` form does.
static inline bool aabbHit(const Aabb& a, const Ray& r, V3 inv, double tmin, double tmax, double* out) {
  double lo = tmin, hi = tmax;
  for (int ax = 0; ax < 3; ++ax) {
    const double iv = get(inv, ax), o = get(r.o, ax);
    double t0 = (get(a.mn, ax) - o) * iv;
    double t1 = (get(a.mx, ax) - o) * iv;
    if (iv < 0) std::swap(t0, t1);
    if (t0 > lo) lo = t0;
    if (t1 < hi) hi = t1;
    if (hi < lo) return false;
  }
  *out = lo;
  return true;
}
static inline V3 inverse(V3 d) { return v3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z); }  // aabb.zig:86-88

// ---------------------------------------------------------------- rng.zig:10-88
static inline uint64_t splitMix64(uint64_t* state) {  // rng.zig:82-88
  *state += 0x9E3779B97F4A7C15ull;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
struct Rng {
  uint64_t s[4];
  static Rng init(uint64_t seed) {  // rng.zig:13-18
    uint64_t sm = seed;
    Rng r;
    for (int i = 0; i < 4; ++i) r.s[i] = splitMix64(&sm);
    return r;
  }
  static Rng forStream(uint64_t seed, uint64_t stream) {  // rng.zig:22-25
    uint64_t sm = seed ^ (stream * 0x9E3779B97F4A7C15ull);
    return init(splitMix64(&sm));
  }
  uint64_t next() {  // rng.zig:27-38 (xoshiro256++)
    const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }             // rng.zig:41-43
  double range(double lo, double hi) { return lo + (hi - lo) * uniform(); }  // rng.zig:46-48
  uint64_t below(uint64_t n) {                                               // rng.zig:51-55
    return (uint64_t)(((unsigned __int128)next() * n) >> 64);
  }
  V3 unitVector() {  // rng.zig:58-63
    const double z = range(-1.0, 1.0);
    const double phi = range(0.0, 2.0 * M_PI);
    const double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    return v3(r * std::cos(phi), r * std::sin(phi), z);
  }
  V3 inUnitSphere() {  // rng.zig:66-71
    for (;;) {
      const double a = range(-1, 1);
      const double b = range(-1, 1);
      const double c = range(-1, 1);
      const V3 p = v3(a, b, c);
      if (len2(p) < 1.0) return p;
    }
  }
  V3 inUnitDisk() {  // rng.zig:74-79
    for (;;) {
      const double a = range(-1, 1);
      const double b = range(-1, 1);
      const V3 p = v3(a, b, 0);
      if (len2(p) < 1.0) return p;
    }
  }
};

// ---------------------------------------------------------------- geometry.zig:11-139
enum ShapeKind : uint32_t { kSphere = 0, kTriangle = 1, kPlane = 2 };
struct Hit {
  double t;
  V3 p, n;
  bool front;
  uint32_t material;
  double curv = 0;  // 1/radius of a sphere hit, 0 for flat shapes (used only by the rounding perturbation)
};
static inline Hit makeHit(const Ray& r, double t, V3 outward, uint32_t material) {  // geometry.zig:20-29
  const bool front = dot(r.d, outward) < 0;
  return {t, at(r, t), front ? outward : neg(outward), front, material};
}
struct Shape {
  uint32_t kind, material;
  V3 a, b, c;  // sphere: a=center, b.x=radius; triangle: v0 v1 v2; plane: a=point, b=normal
};
static inline bool sphereIntersect(const Shape& s, const Ray& r, double tmin, double tmax, Hit* h) {
  // geometry.zig:37-52
  const double radius = s.b.x;
  const V3 oc = sub(s.a, r.o);
  const double a = len2(r.d);
  const double hh = dot(r.d, oc);
  const double c = len2(oc) - radius * radius;
  const double disc = hh * hh - a * c;
  if (disc < 0) return false;
  const double sq = std::sqrt(disc);
  double t = (hh - sq) / a;
  if (t <= tmin || t >= tmax) {
    t = (hh + sq) / a;
    if (t <= tmin || t >= tmax) return false;
  }
  const V3 outward = vdiv(sub(at(r, t), s.a), radius);
  *h = makeHit(r, t, outward, s.material);
  h->curv = 1.0 / radius;
  return true;
}
static inline bool planeIntersect(const Shape& s, const Ray& r, double tmin, double tmax, Hit* h) {
  // geometry.zig:67-73
  const double denom = dot(s.b, r.d);
  if (std::fabs(denom) < 1e-12) return false;
  const double t = dot(sub(s.a, r.o), s.b) / denom;
  if (t <= tmin || t >= tmax) return false;
  *h = makeHit(r, t, s.b, s.material);
  return true;
}
static inline bool triangleIntersect(const Shape& s, const Ray& r, double tmin, double tmax, Hit* h) {
  // geometry.zig:83-99 (Moller-Trumbore)
  const V3 e1 = sub(s.b, s.a);
  const V3 e2 = sub(s.c, s.a);
  const V3 p = cross(r.d, e2);
  const double det = dot(e1, p);
  if (std::fabs(det) < 1e-12) return false;
  const double inv_det = 1.0 / det;
  const V3 sv = sub(r.o, s.a);
  const double u = dot(sv, p) * inv_det;
  if (u < 0 || u > 1) return false;
  const V3 q = cross(sv, e1);
  const double v = dot(r.d, q) * inv_det;
  if (v < 0 || u + v > 1) return false;
  const double t = dot(e2, q) * inv_det;
  if (t <= tmin || t >= tmax) return false;
  *h = makeHit(r, t, normalize(cross(e1, e2)), s.material);
  return true;
}
static inline bool shapeIntersect(const Shape& s, const Ray& r, double tmin, double tmax, Hit* h) {
  switch (s.kind) {
    case kSphere: return sphereIntersect(s, r, tmin, tmax, h);
    case kTriangle: return triangleIntersect(s, r, tmin, tmax, h);
    default: return planeIntersect(s, r, tmin, tmax, h);
  }
}
static inline bool shapeBounds(const Shape& s, Aabb* out) {  // geometry.zig:54-57,101-103,118-124
  if (s.kind == kSphere) {
    const V3 r = splat(std::fabs(s.b.x));
    *out = {sub(s.a, r), add(s.a, r)};
    return true;
  }
  if (s.kind == kTriangle) {
    *out = pad(include(fromPoints(s.a, s.b), s.c), 1e-6);
    return true;
  }
  return false;
}

struct Counters {  // instrumentation for the algorithmic-bytes model
  uint64_t extend_rays = 0, shadow_rays = 0, node_pops = 0, leaf_pops = 0;
  uint64_t tri_tests = 0, sph_tests = 0, plane_tests = 0, paths = 0;
  void addTest(const Shape& s) {
    if (s.kind == kTriangle) ++tri_tests;
    else if (s.kind == kSphere) ++sph_tests;
    else ++plane_tests;
  }
  void merge(const Counters& o) {
    extend_rays += o.extend_rays; shadow_rays += o.shadow_rays; node_pops += o.node_pops;
    leaf_pops += o.leaf_pops; tri_tests += o.tri_tests; sph_tests += o.sph_tests;
    plane_tests += o.plane_tests; paths += o.paths;
  }
};

static inline bool intersectAll(const std::vector<Shape>& shapes, const Ray& r, double tmin, double tmax,
                                Hit* best, Counters* ctr) {  // geometry.zig:129-139
  double closest = tmax;
  bool found = false;
  Hit h;
  for (const Shape& s : shapes) {
    if (ctr) ctr->addTest(s);
    if (shapeIntersect(s, r, tmin, closest, &h)) {
      closest = h.t;
      *best = h;
      found = true;
    }
  }
  return found;
}

// ---------------------------------------------------------------- material.zig:15-70
enum MatKind : uint32_t { kLambertian = 0, kMetal = 1, kDielectric = 2, kEmissive = 3 };
struct Material {
  uint32_t kind;
  V3 color;      // albedo or emitted colour
  double param;  // fuzz or ior
};
static inline double pow5(double x) {  // std.math.pow(f64, x, 5) by repeated squaring
  const double x2 = x * x;
  const double x4 = x2 * x2;
  return x * x4;
}
static inline double schlick(double cosine, double eta) {  // material.zig:66-70
  double r0 = (1.0 - eta) / (1.0 + eta);
  r0 *= r0;
  return r0 + (1.0 - r0) * pow5(1.0 - cosine);
}
static inline V3 emitted(const Material& m, const Hit& h) {  // material.zig:26-32
  if (m.kind == kEmissive) return h.front ? m.color : splat(0);
  return splat(0);
}
static inline bool scatter(const Material& m, const Ray& rin, const Hit& h, Rng* rng, V3* att, Ray* out) {
  // material.zig:34-62
  switch (m.kind) {
    case kLambertian: {
      V3 dir = add(h.n, rng->unitVector());
      if (nearZero(dir)) dir = h.n;
      *att = m.color;
      *out = {h.p, normalize(dir)};
      return true;
    }
    case kMetal: {
      V3 dir = reflect(normalize(rin.d), h.n);
      if (m.param > 0) dir = normalize(add(dir, scale(rng->unitVector(), m.param)));
      if (dot(dir, h.n) <= 0) return false;
      *att = m.color;
      *out = {h.p, dir};
      return true;
    }
    case kDielectric: {
      const double eta = h.front ? 1.0 / m.param : m.param;
      const V3 unit = normalize(rin.d);
      const double cos_theta = std::fmin(dot(neg(unit), h.n), 1.0);
      const double sin_theta = std::sqrt(std::fmax(0.0, 1.0 - cos_theta * cos_theta));
      const bool cannot_refract = eta * sin_theta > 1.0;
      const V3 dir = (cannot_refract || schlick(cos_theta, eta) > rng->uniform()) ? reflect(unit, h.n)
                                                                                : refract(unit, h.n, eta);
      *att = splat(1);
      *out = {h.p, dir};
      return true;
    }
    default: return false;
  }
}

// ---------------------------------------------------------------- camera.zig:8-67
struct CameraConfig {
  V3 from = v3(0, 0, 0), at = v3(0, 0, -1), up = v3(0, 1, 0);
  double vfov = 60, aperture = 0;
  bool has_focus = false;
  double focus = 0;
};
struct Camera {
  V3 origin, upper_left, horizontal, vertical, u, v;
  double lens_radius;
  static Camera init(const CameraConfig& cfg, double aspect) {  // camera.zig:29-55
    const double theta = cfg.vfov * 0.017453292519943295;       // std.math.degreesToRadians
    const double half_h = std::tan(theta / 2);
    const double viewport_h = 2.0 * half_h;
    const double viewport_w = aspect * viewport_h;
    const V3 w = normalize(sub(cfg.from, cfg.at));
    const V3 u = normalize(cross(cfg.up, w));
    const V3 v = cross(w, u);
    const double focus = cfg.has_focus ? cfg.focus : length(sub(cfg.from, cfg.at));
    const V3 horizontal = scale(u, viewport_w * focus);
    const V3 vertical = scale(v, -viewport_h * focus);
    const V3 upper_left = sub(sub(sub(cfg.from, scale(w, focus)), scale(horizontal, 0.5)), scale(vertical, 0.5));
    return {cfg.from, upper_left, horizontal, vertical, u, v, cfg.aperture / 2};
  }
  Ray ray(double s, double t, Rng* rng) const {  // camera.zig:59-67
    V3 o = origin;
    if (lens_radius > 0) {
      const V3 d = scale(rng->inUnitDisk(), lens_radius);
      o = add(add(o, scale(u, d.x)), scale(v, d.y));
    }
    const V3 target = add(add(upper_left, scale(horizontal, s)), scale(vertical, t));
    return {o, normalize(sub(target, o))};
  }
};

// ---------------------------------------------------------------- bvh.zig:16-262
struct Node {
  Aabb bounds;
  uint32_t first, count;
};
struct Bvh {
  std::vector<Node> nodes;
  std::vector<uint32_t> indices;
};
enum BuildError { kBuildOk = 0, kUnboundedShape = 1, kTooManyPrimitives = 2 };

struct Builder {  // bvh.zig:140-262
  std::vector<Node>* nodes;
  std::vector<uint32_t>* indices;
  const std::vector<Aabb>* boxes;
  const std::vector<V3>* centroids;
  struct Split {
    int axis;
    int bin;
    double cost;
  };
  static int binOf(double c, double lo, double extent) {  // bvh.zig:258-262
    const double f = (c - lo) / extent * 12.0;
    return (int)std::fmax(0.0, std::fmin(f, 11.0));
  }
  bool findSplit(uint32_t first, uint32_t count, const Aabb& bounds, const Aabb& cbounds, Split* best) {
    // bvh.zig:211-256
    bool have = false;
    const double parent_area = surfaceArea(bounds);
    if (parent_area <= 0) return false;
    for (int axis = 0; axis < 3; ++axis) {
      const double lo = get(cbounds.mn, axis);
      const double extent = get(cbounds.mx, axis) - lo;
      if (!(extent > 0)) continue;
      Aabb bin_bounds[12];
      uint64_t bin_counts[12];
      for (int i = 0; i < 12; ++i) { bin_bounds[i] = aabbEmpty(); bin_counts[i] = 0; }
      for (uint32_t j = 0; j < count; ++j) {
        const uint32_t idx = (*indices)[first + j];
        const int b = binOf(get((*centroids)[idx], axis), lo, extent);
        bin_counts[b] += 1;
        bin_bounds[b] = merge(bin_bounds[b], (*boxes)[idx]);
      }
      double right_area[12];
      uint64_t right_count[12];
      Aabb acc = aabbEmpty();
      uint64_t acc_n = 0;
      for (int i = 11; i >= 0; --i) {
        acc = merge(acc, bin_bounds[i]);
        acc_n += bin_counts[i];
        right_area[i] = surfaceArea(acc);
        right_count[i] = acc_n;
      }
      acc = aabbEmpty();
      acc_n = 0;
      for (int b = 0; b < 11; ++b) {
        acc = merge(acc, bin_bounds[b]);
        acc_n += bin_counts[b];
        const uint64_t rn = right_count[b + 1];
        if (acc_n == 0 || rn == 0) continue;
        const double cost = 1.0 + (surfaceArea(acc) * (double)acc_n + right_area[b + 1] * (double)rn) / parent_area;
        if (!have || cost < best->cost) {
          *best = {axis, b, cost};
          have = true;
        }
      }
    }
    return have;
  }
  void makeChildren(size_t node_idx, uint32_t first, uint32_t left_count, uint32_t count, int level) {
    // bvh.zig:187-195
    const uint32_t left = (uint32_t)nodes->size();
    nodes->push_back({aabbEmpty(), first, left_count});
    nodes->push_back({aabbEmpty(), first + left_count, count - left_count});
    (*nodes)[node_idx].first = left;
    (*nodes)[node_idx].count = 0;
    subdivide(left, level + 1);
    subdivide(left + 1, level + 1);
  }
  void splitMedian(size_t node_idx, const Aabb& cbounds, int level) {  // bvh.zig:197-209
    const int axis = longestAxis(cbounds);
    const Node node = (*nodes)[node_idx];
    const std::vector<V3>& cs = *centroids;
    std::stable_sort(indices->begin() + node.first, indices->begin() + node.first + node.count,
                     [&](uint32_t a, uint32_t b) { return get(cs[a], axis) < get(cs[b], axis); });
    makeChildren(node_idx, node.first, node.count / 2, node.count, level);
  }
  void subdivide(size_t node_idx, int level) {  // bvh.zig:148-185
    const uint32_t first = (*nodes)[node_idx].first;
    const uint32_t count = (*nodes)[node_idx].count;
    Aabb bounds = aabbEmpty(), cbounds = aabbEmpty();
    for (uint32_t j = 0; j < count; ++j) {
      const uint32_t i = (*indices)[first + j];
      bounds = merge(bounds, (*boxes)[i]);
      cbounds = include(cbounds, (*centroids)[i]);
    }
    (*nodes)[node_idx].bounds = bounds;
    if (count <= 1 || level >= 64) return;
    const double leaf_cost = (double)count;
    Split split;
    if (!findSplit(first, count, bounds, cbounds, &split)) {
      if (count <= 4) return;
      splitMedian(node_idx, cbounds, level);
      return;
    }
    if (count <= 4 && split.cost >= leaf_cost) return;
    const double lo = get(cbounds.mn, split.axis);
    const double extent = get(cbounds.mx, split.axis) - lo;
    uint32_t mid = 0;
    for (uint32_t j = 0; j < count; ++j) {
      const uint32_t idx = (*indices)[first + j];
      if (binOf(get((*centroids)[idx], split.axis), lo, extent) <= split.bin) {
        std::swap((*indices)[first + mid], (*indices)[first + j]);
        mid += 1;
      }
    }
    makeChildren(node_idx, first, mid, count, level);
  }
};

static int bvhBuild(const std::vector<Shape>& shapes, Bvh* out) {  // bvh.zig:41-67
  const size_t n = shapes.size();
  out->nodes.clear();
  out->indices.clear();
  if (n == 0) return kBuildOk;
  if (n > 0xFFFFFFFFull / 2) return kTooManyPrimitives;
  std::vector<Aabb> boxes(n);
  std::vector<V3> cents(n);
  out->indices.resize(n);
  for (size_t i = 0; i < n; ++i) {
    out->indices[i] = (uint32_t)i;
    if (!shapeBounds(shapes[i], &boxes[i])) {
      out->indices.clear();
      return kUnboundedShape;
    }
    cents[i] = centroid(boxes[i]);
  }
  out->nodes.reserve(2 * n - 1);
  out->nodes.push_back({aabbEmpty(), 0, (uint32_t)n});
  Builder b{&out->nodes, &out->indices, &boxes, &cents};
  b.subdivide(0, 0);
  return kBuildOk;
}

static bool bvhIntersect(const Bvh& bvh, const std::vector<Shape>& shapes, const Ray& r, double tmin, double tmax,
                         Hit* best, Counters* ctr) {  // bvh.zig:80-126
  if (bvh.nodes.empty()) return false;
  const V3 inv = inverse(r.d);
  double closest = tmax;
  bool found = false;
  struct Entry {
    uint32_t node;
    double t;
  } stack[66];
  int sp = 0;
  double t;
  if (aabbHit(bvh.nodes[0].bounds, r, inv, tmin, closest, &t)) {
    stack[0] = {0, t};
    sp = 1;
  }
  Hit h;
  while (sp > 0) {
    sp -= 1;
    const Entry e = stack[sp];
    if (e.t > closest) continue;
    const Node& node = bvh.nodes[e.node];
    if (node.count > 0) {
      if (ctr) ++ctr->leaf_pops;
      for (uint32_t j = 0; j < node.count; ++j) {
        const Shape& s = shapes[bvh.indices[node.first + j]];
        if (ctr) ctr->addTest(s);
        if (shapeIntersect(s, r, tmin, closest, &h)) {
          closest = h.t;
          *best = h;
          found = true;
        }
      }
      continue;
    }
    if (ctr) ++ctr->node_pops;
    const uint32_t l = node.first;
    double tl, tr;
    const bool hl = aabbHit(bvh.nodes[l].bounds, r, inv, tmin, closest, &tl);
    const bool hr = aabbHit(bvh.nodes[l + 1].bounds, r, inv, tmin, closest, &tr);
    if (hl && hr) {
      const bool near_left = tl <= tr;
      stack[sp] = near_left ? Entry{l + 1, tr} : Entry{l, tl};
      stack[sp + 1] = near_left ? Entry{l, tl} : Entry{l + 1, tr};
      sp += 2;
    } else if (hl) {
      stack[sp++] = {l, tl};
    } else if (hr) {
      stack[sp++] = {l + 1, tr};
    }
  }
  return found;
}

static size_t bvhDepthFrom(const Bvh& b, uint32_t i) {  // bvh.zig:128-137
  const Node& n = b.nodes[i];
  if (n.count > 0) return 1;
  return 1 + std::max(bvhDepthFrom(b, n.first), bvhDepthFrom(b, n.first + 1));
}

// ---------------------------------------------------------------- light.zig:20-198
struct Light {
  Shape shape;
  uint32_t shape_index;
  V3 radiance;
  double weight;
};
static inline double luminance(V3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }  // light.zig:38-40
static inline double triangleArea(const Shape& t) {  // light.zig:42-44
  return 0.5 * length(cross(sub(t.b, t.a), sub(t.c, t.a)));
}
static inline bool outside(const Shape& s, V3 origin) {  // light.zig:47-50
  const double d = length(sub(s.a, origin));
  return d > std::fabs(s.b.x) * 1.000000001;
}
static inline void basis(V3 w, V3* u, V3* v) {  // light.zig:53-57
  const V3 a = std::fabs(w.x) > 0.9 ? v3(0, 1, 0) : v3(1, 0, 0);
  *v = normalize(cross(w, a));
  *u = cross(*v, w);
}
struct LightSample {
  V3 dir;
  double distance;
  V3 radiance;
  double pdf;
  size_t index = 0;  // which light was picked (not part of the reference's struct)
};
struct Lights {
  std::vector<Light> items;
  std::vector<double> cdf;
  double total = 0;

  void build(const std::vector<Shape>& shapes, const std::vector<Material>& materials) {  // light.zig:68-94
    items.clear();
    cdf.clear();
    for (size_t i = 0; i < shapes.size(); ++i) {
      const Shape& s = shapes[i];
      double area;
      if (s.kind == kTriangle) area = triangleArea(s);
      else if (s.kind == kSphere) area = s.b.x > 0 ? 4 * M_PI * s.b.x * s.b.x : 0;
      else continue;
      const Material& m = materials[s.material];
      if (m.kind != kEmissive) continue;
      const double weight = luminance(m.color) * area;
      if (!(weight > 0) || !std::isfinite(weight)) continue;
      items.push_back({s, (uint32_t)i, m.color, weight});
    }
    double sum = 0;
    for (const Light& l : items) {
      sum += l.weight;
      cdf.push_back(sum);
    }
    total = sum;
  }
  const Light& pick(Rng* rng) const {  // light.zig:106-115
    const double x = rng->uniform() * total;
    size_t lo = 0, hi = cdf.size() - 1;
    while (lo < hi) {
      const size_t mid = (lo + hi) / 2;
      if (cdf[mid] > x) hi = mid;
      else lo = mid + 1;
    }
    return items[lo];
  }
  bool sample(V3 origin, Rng* rng, LightSample* out) const {  // light.zig:120-166
    if (items.empty()) return false;
    const Light& light = pick(rng);
    out->index = (size_t)(&light - items.data());
    const double select = light.weight / total;
    const Shape& s = light.shape;
    if (s.kind == kTriangle) {
      const double su = std::sqrt(rng->uniform());
      const double b2 = rng->uniform();
      const V3 p = add(add(scale(s.a, 1 - su), scale(s.b, su * (1 - b2))), scale(s.c, su * b2));
      const V3 to = sub(p, origin);
      const double dist = length(to);
      if (dist < 1e-9) return false;
      const V3 dir = scale(to, 1 / dist);
      const V3 n = normalize(cross(sub(s.b, s.a), sub(s.c, s.a)));
      const double cos_l = -dot(dir, n);
      if (cos_l <= 1e-9) return false;
      *out = {dir, dist, light.radiance, select * dist * dist / (cos_l * triangleArea(s)), out->index};
      return true;
    }
    // sphere
    if (!outside(s, origin)) return false;
    const V3 axis = sub(s.a, origin);
    const double d2 = len2(axis);
    const V3 w = scale(axis, 1 / std::sqrt(d2));
    const double cos_max = std::sqrt(std::fmax(0.0, 1 - s.b.x * s.b.x / d2));
    const double cos_t = 1 - rng->uniform() * (1 - cos_max);
    const double sin_t = std::sqrt(std::fmax(0.0, 1 - cos_t * cos_t));
    const double phi = 2 * M_PI * rng->uniform();
    V3 bu, bv;
    basis(w, &bu, &bv);
    const V3 dir =
        normalize(add(add(scale(bu, sin_t * std::cos(phi)), scale(bv, sin_t * std::sin(phi))), scale(w, cos_t)));
    Hit h;
    if (!sphereIntersect(s, Ray{origin, dir}, 0, kInf, &h)) return false;
    if (!h.front) return false;
    *out = {dir, h.t, light.radiance, select / (2 * M_PI * (1 - cos_max)), out->index};
    return true;
  }
  double pdfHit(V3 origin, const Hit& hit) const {  // light.zig:171-198
    const V3 to = sub(hit.p, origin);
    const double dist = length(to);
    if (dist < 1e-9) return 0;
    const V3 dir = scale(to, 1 / dist);
    const Ray ray{origin, dir};
    double pdf = 0;
    for (const Light& light : items) {
      const double select = light.weight / total;
      const Shape& s = light.shape;
      Hit h;
      if (s.kind == kTriangle) {
        if (!triangleIntersect(s, ray, 1e-9, kInf, &h)) continue;
        if (std::fabs(h.t - dist) > 1e-6 * dist || !h.front) continue;
        pdf += select * dist * dist / (-dot(dir, h.n) * triangleArea(s));
      } else if (s.kind == kSphere) {
        if (!outside(s, origin)) continue;
        if (!sphereIntersect(s, ray, 1e-9, kInf, &h)) continue;
        if (std::fabs(h.t - dist) > 1e-6 * dist || !h.front) continue;
        const double d2 = len2(sub(s.a, origin));
        const double cos_max = std::sqrt(std::fmax(0.0, 1 - s.b.x * s.b.x / d2));
        pdf += select / (2 * M_PI * (1 - cos_max));
      }
    }
    return pdf;
  }
};

// ---------------------------------------------------------------- scene.zig:15-120
struct Background {
  uint32_t kind = 1;  // 0 solid, 1 sky
  V3 a = v3(1, 1, 1), b = v3(0.5, 0.7, 1.0);
  V3 radiance(V3 dir) const {  // scene.zig:20-28
    if (kind == 0) return a;
    const double t = std::fmax(0.0, std::fmin(normalize(dir).y, 1.0));
    return lerp(a, b, t);
  }
};
struct Settings {
  uint64_t width = 400, height = 225, spp = 32, max_depth = 16;
};
struct Scene {
  std::vector<Material> materials;
  std::vector<Shape> shapes, planes;
  CameraConfig camera;
  Background background;
  Settings settings;
  bool built = false;
  Bvh bvh;
  Lights lights;

  void add(const Shape& s) {  // scene.zig:71-78
    built = false;
    lights = Lights();
    if (s.kind == kPlane) planes.push_back(s);
    else shapes.push_back(s);
  }
  void addQuad(V3 corner, V3 u, V3 v, uint32_t mat) {  // scene.zig:82-88
    const V3 p1 = ::orc::add(corner, u);
    const V3 p2 = ::orc::add(::orc::add(corner, u), v);
    const V3 p3 = ::orc::add(corner, v);
    add(Shape{kTriangle, mat, corner, p1, p2});
    add(Shape{kTriangle, mat, corner, p2, p3});
  }
  int build() {  // scene.zig:92-96
    built = false;
    const int e = bvhBuild(shapes, &bvh);
    if (e != kBuildOk) return e;
    lights.build(shapes, materials);
    built = true;
    return kBuildOk;
  }
  bool intersect(const Ray& r, double tmin, double tmax, Hit* out, Counters* ctr) const {  // scene.zig:108-116
    bool found = built ? bvhIntersect(bvh, shapes, r, tmin, tmax, out, ctr)
                       : intersectAll(shapes, r, tmin, tmax, out, ctr);
    const double limit = found ? out->t : tmax;
    Hit h;
    if (intersectAll(planes, r, tmin, limit, &h, ctr)) {
      *out = h;
      found = true;
    }
    return found;
  }
};

// ---------------------------------------------------------------- render.zig:12-180
struct Options {
  uint64_t width, height, spp, max_depth, threads, seed, tile_size;
  uint32_t light_sampling, _pad;
};
static const double kRayEpsilon = 1e-4;  // render.zig:26
static const uint64_t kRouletteDepth = 4;  // render.zig:28
static inline double powerWeight(double a, double b) {  // render.zig:38-41
  const double a2 = a * a;
  return a2 / (a2 + b * b);
}
static V3 directLight(const Scene& scene, V3 albedo, const Hit& hit, Rng* rng, Counters* ctr) {  // render.zig:46-56
  LightSample ls;
  if (!scene.lights.sample(hit.p, rng, &ls)) return splat(0);
  const double cos_s = dot(hit.n, ls.dir);
  if (cos_s <= 0 || !(ls.pdf > 0)) return splat(0);
  const Ray shadow{hit.p, ls.dir};
  if (ctr) ++ctr->shadow_rays;
  Hit tmp;
  if (scene.intersect(shadow, kRayEpsilon, ls.distance * 0.999999, &tmp, ctr)) return splat(0);
  const double bsdf_pdf = cos_s / M_PI;
  const double sc = cos_s / (M_PI * ls.pdf) * powerWeight(ls.pdf, bsdf_pdf);
  return scale(mul(albedo, ls.radiance), sc);
}
// Rounding-scale perturbation (off unless a test sets it): every scattered
// direction is multiplied component-wise by 1 + e, and every hit point moved
// by that much within its tangent plane, e uniform in (-rel, rel) from a hash
// of the path's RNG state and depth (no draw is consumed, so the path keeps
// its stream); on a sphere the normal follows the surface to the moved point.
// The component along the normal is left out: an offset off the surface is
// what t_min (and, on the device, skipping the primitive a ray leaves)
// absorbs, not a change of the path. With rel at the scale of
// fp32 rounding (2^-24 .. 2^-22) it measures how far rounding alone moves a
// converged image with common random numbers: the noise floor any fp32
// renderer of the same estimator has on that scene.
static double g_perturb_rel = 0;
static inline double perturbed(uint64_t key, double v) {
  uint64_t z = key;
  const double u = (double)(splitMix64(&z) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  return v * (1.0 + u * g_perturb_rel);
}
static inline V3 perturbed3(uint64_t key, V3 v) {
  return v3(perturbed(key, v.x), perturbed(key + 1, v.y), perturbed(key + 2, v.z));
}

static V3 traceWith(const Scene& scene, Ray r, Rng* rng, uint64_t max_depth, bool light_sampling,
                    Counters* ctr) {  // render.zig:61-102
  const bool use_nee = light_sampling && !scene.lights.items.empty();
  V3 radiance = splat(0), throughput = splat(1);
  V3 prev_origin = splat(0);
  double prev_pdf = 0;
  for (uint64_t depth = 0; depth < max_depth; ++depth) {
    Hit hit;
    if (ctr) ++ctr->extend_rays;
    if (!scene.intersect(r, kRayEpsilon, kInf, &hit, ctr))
      return add(radiance, mul(throughput, scene.background.radiance(r.d)));
    if (g_perturb_rel > 0) {  // within the tangent plane: off-surface error is what t_min absorbs
      V3 d = sub(perturbed3(rng->s[0] ^ (depth * 0x9E3779B97F4A7C15ull), hit.p), hit.p);
      d = sub(d, scale(hit.n, dot(d, hit.n)));
      hit.p = add(hit.p, d);
      // On a sphere the normal follows the surface to the moved point.
      if (hit.curv != 0) hit.n = normalize(add(hit.n, scale(d, hit.front ? hit.curv : -hit.curv)));
    }
    const Material& mat = scene.materials[hit.material];
    const V3 le = emitted(mat, hit);
    if (maxComponent(le) > 0) {
      const double w = prev_pdf > 0 ? powerWeight(prev_pdf, scene.lights.pdfHit(prev_origin, hit)) : 1.0;
      radiance = add(radiance, scale(mul(throughput, le), w));
    }
    if (mat.kind == kLambertian && use_nee)
      radiance = add(radiance, mul(throughput, directLight(scene, mat.color, hit, rng, ctr)));
    V3 att;
    Ray next;
    if (!scatter(mat, r, hit, rng, &att, &next)) break;
    if (g_perturb_rel > 0) next.d = perturbed3(rng->s[1] ^ (depth * 0xD1B54A32D192ED03ull), next.d);
    throughput = mul(throughput, att);
    prev_pdf = (use_nee && mat.kind == kLambertian) ? std::fmax(dot(next.d, hit.n), 0.0) / M_PI : 0;
    prev_origin = hit.p;
    r = next;
    if (depth + 1 >= kRouletteDepth) {
      const double p = std::fmin(maxComponent(throughput), 0.95);
      if (p <= 0 || rng->uniform() >= p) break;
      throughput = vdiv(throughput, p);
    }
  }
  return radiance;
}

// ---------------------------------------------------------------- gradient (dual numbers)
// Forward-mode derivative of traceWith (render.zig:61-102) with respect to
// the material colours: the albedo of Lambertian/metal materials and the
// radiance of emissive ones (3 parameters each, offsets per material). The
// value part repeats traceWith operation for operation. Geometry, visibility,
// MIS weights and light-selection probabilities are constants; the Russian
// roulette probability is used by value only (the estimator T/p*[u<p] has
// expectation T for any p, so its derivative is taken with p fixed).
static const int kMaxGradParams = 96;
struct DV3 {
  V3 v;
  V3 d[kMaxGradParams];
};
struct GradCtx {
  const int32_t* offsets;  // per material: first parameter index or -1
  int nparams;
};
static inline void dvZero(DV3& a, int np, V3 v) {
  a.v = v;
  for (int k = 0; k < np; ++k) a.d[k] = splat(0);
}
// a += b (colour contribution with derivatives)
static inline void dvAddInto(DV3& a, const DV3& b, int np) {
  a.v = add(a.v, b.v);
  for (int k = 0; k < np; ++k) a.d[k] = add(a.d[k], b.d[k]);
}
// unit derivative of material m's colour channel c
static inline V3 unitChan(int c, double x) { return c == 0 ? v3(x, 0, 0) : (c == 1 ? v3(0, x, 0) : v3(0, 0, x)); }

static V3 traceWithGrad(const Scene& scene, Ray r, Rng* rng, uint64_t max_depth, bool light_sampling,
                        const GradCtx& g, DV3* out) {
  const int np = g.nparams;
  const bool use_nee = light_sampling && !scene.lights.items.empty();
  DV3 radiance, throughput, tmp;
  dvZero(radiance, np, splat(0));
  dvZero(throughput, np, splat(1));
  V3 prev_origin = splat(0);
  double prev_pdf = 0;
  for (uint64_t depth = 0; depth < max_depth; ++depth) {
    Hit hit;
    if (!scene.intersect(r, kRayEpsilon, kInf, &hit, nullptr)) {
      const V3 bg = scene.background.radiance(r.d);
      radiance.v = add(radiance.v, mul(throughput.v, bg));
      for (int k = 0; k < np; ++k) radiance.d[k] = add(radiance.d[k], mul(throughput.d[k], bg));
      *out = radiance;
      return radiance.v;
    }
    const Material& mat = scene.materials[hit.material];
    const V3 le = emitted(mat, hit);
    if (maxComponent(le) > 0) {
      const double w = prev_pdf > 0 ? powerWeight(prev_pdf, scene.lights.pdfHit(prev_origin, hit)) : 1.0;
      radiance.v = add(radiance.v, scale(mul(throughput.v, le), w));
      for (int k = 0; k < np; ++k) radiance.d[k] = add(radiance.d[k], scale(mul(throughput.d[k], le), w));
      const int off = g.offsets[hit.material];
      if (off >= 0)
        for (int c = 0; c < 3; ++c)
          radiance.d[off + c] = add(radiance.d[off + c], scale(mul(throughput.v, unitChan(c, 1.0)), w));
    }
    if (mat.kind == kLambertian && use_nee) {  // directLight, render.zig:46-56
      LightSample ls;
      if (scene.lights.sample(hit.p, rng, &ls)) {
        const double cos_s = dot(hit.n, ls.dir);
        Hit hh;
        if (cos_s > 0 && ls.pdf > 0 &&
            !scene.intersect(Ray{hit.p, ls.dir}, kRayEpsilon, ls.distance * 0.999999, &hh, nullptr)) {
          const double bsdf_pdf = cos_s / M_PI;
          const double sc = cos_s / (M_PI * ls.pdf) * powerWeight(ls.pdf, bsdf_pdf);
          const V3 al = scale(mul(mat.color, ls.radiance), sc);
          radiance.v = add(radiance.v, mul(throughput.v, al));
          for (int k = 0; k < np; ++k) radiance.d[k] = add(radiance.d[k], mul(throughput.d[k], al));
          const int offa = g.offsets[hit.material];
          if (offa >= 0)
            for (int c = 0; c < 3; ++c)
              radiance.d[offa + c] = add(radiance.d[offa + c],
                                         mul(throughput.v, scale(mul(unitChan(c, 1.0), ls.radiance), sc)));
          const int offl = g.offsets[scene.lights.items[ls.index].shape.material];
          if (offl >= 0)
            for (int c = 0; c < 3; ++c)
              radiance.d[offl + c] = add(radiance.d[offl + c],
                                         mul(throughput.v, scale(mul(mat.color, unitChan(c, 1.0)), sc)));
        }
      }
    }
    V3 att;
    Ray next;
    if (!scatter(mat, r, hit, rng, &att, &next)) break;
    tmp = throughput;
    throughput.v = mul(tmp.v, att);
    for (int k = 0; k < np; ++k) throughput.d[k] = mul(tmp.d[k], att);
    if (mat.kind == kLambertian || mat.kind == kMetal) {
      const int off = g.offsets[hit.material];
      if (off >= 0)
        for (int c = 0; c < 3; ++c) throughput.d[off + c] = add(throughput.d[off + c], mul(tmp.v, unitChan(c, 1.0)));
    }
    prev_pdf = (use_nee && mat.kind == kLambertian) ? std::fmax(dot(next.d, hit.n), 0.0) / M_PI : 0;
    prev_origin = hit.p;
    r = next;
    if (depth + 1 >= kRouletteDepth) {
      const double p = std::fmin(maxComponent(throughput.v), 0.95);
      if (p <= 0 || rng->uniform() >= p) break;
      throughput.v = vdiv(throughput.v, p);
      const double ip = 1.0 / p;
      for (int k = 0; k < np; ++k) throughput.d[k] = scale(throughput.d[k], ip);
    }
  }
  *out = radiance;
  return radiance.v;
}

struct RenderCtx {  // render.zig:104-148
  const Scene* scene;
  Camera camera;
  Options opts;
  double* image;  // W*H*3, row-major y*W+x
  uint64_t tiles_x, tile_count;
  uint64_t pixel_stride;  // 1 = every pixel; k = only pixels with (y*W+x) % k == 0
  std::atomic<uint64_t> next_tile{0};
  std::atomic<uint64_t> totals[8];

  void renderTile(uint64_t tile, Counters* ctr) {
    const uint64_t ts = opts.tile_size;
    const uint64_t x0 = (tile % tiles_x) * ts, y0 = (tile / tiles_x) * ts;
    const uint64_t x1 = std::min(x0 + ts, opts.width), y1 = std::min(y0 + ts, opts.height);
    const double w = (double)opts.width, h = (double)opts.height;
    const double inv_spp = 1.0 / (double)opts.spp;
    for (uint64_t y = y0; y < y1; ++y) {
      for (uint64_t x = x0; x < x1; ++x) {
        const uint64_t pix = y * opts.width + x;
        if (pixel_stride > 1 && pix % pixel_stride != 0) continue;
        Rng rng = Rng::forStream(opts.seed, pix);
        V3 sum = splat(0);
        for (uint64_t k = 0; k < opts.spp; ++k) {
          const double s = ((double)x + rng.uniform()) / w;
          const double t = ((double)y + rng.uniform()) / h;
          const Ray cr = camera.ray(s, t, &rng);
          const V3 c = traceWith(*scene, cr, &rng, opts.max_depth, opts.light_sampling != 0, ctr);
          if (ctr) ++ctr->paths;
          if (isFinite(c)) sum = add(sum, c);
        }
        const V3 px = scale(sum, inv_spp);
        image[3 * pix + 0] = px.x;
        image[3 * pix + 1] = px.y;
        image[3 * pix + 2] = px.z;
      }
    }
  }
  void work(Counters* ctr) {
    for (;;) {
      const uint64_t tile = next_tile.fetch_add(1, std::memory_order_relaxed);
      if (tile >= tile_count) return;
      renderTile(tile, ctr);
    }
  }
};

}  // namespace orc

// =============================================================== C ABI for ctypes
using namespace orc;

extern "C" {

struct OrcMaterial {
  uint32_t kind, _pad;
  double color[3];
  double param;
};
struct OrcShape {
  uint32_t kind, material;
  double v[9];
};
struct OrcCamera {
  double from[3], at[3], up[3];
  double vfov, aperture, focus;
  uint32_t has_focus, _pad;
};
struct OrcBackground {
  uint32_t kind, _pad;
  double a[3], b[3];
};
struct OrcNode {
  double mn[3], mx[3];
  uint32_t first, count;
};
struct OrcRay {
  double o[3], d[3];
};
struct OrcHit {
  double t, p[3], n[3];
  uint32_t front, material;
  int32_t hit, _pad;
};
struct OrcCounters {
  uint64_t extend_rays, shadow_rays, node_pops, leaf_pops, tri_tests, sph_tests, plane_tests, paths;
};
struct OrcLightSample {
  double dir[3], distance, radiance[3], pdf;
  int32_t valid, _pad;
};

}  // extern "C"

static V3 fromArr(const double* a) { return v3(a[0], a[1], a[2]); }
static void toArr(V3 v, double* a) { a[0] = v.x; a[1] = v.y; a[2] = v.z; }
static Shape toShape(const OrcShape& s) {
  Shape r;
  r.kind = s.kind;
  r.material = s.material;
  r.a = v3(s.v[0], s.v[1], s.v[2]);
  if (s.kind == kSphere) r.b = v3(s.v[3], 0, 0), r.c = v3(0, 0, 0);
  else r.b = v3(s.v[3], s.v[4], s.v[5]), r.c = v3(s.v[6], s.v[7], s.v[8]);
  return r;
}
static OrcShape fromShape(const Shape& s) {
  OrcShape r{};
  r.kind = s.kind;
  r.material = s.material;
  r.v[0] = s.a.x; r.v[1] = s.a.y; r.v[2] = s.a.z;
  if (s.kind == kSphere) {
    r.v[3] = s.b.x;
  } else {
    r.v[3] = s.b.x; r.v[4] = s.b.y; r.v[5] = s.b.z;
    r.v[6] = s.c.x; r.v[7] = s.c.y; r.v[8] = s.c.z;
  }
  return r;
}
static Ray toRay(const OrcRay& r) { return Ray{fromArr(r.o), fromArr(r.d)}; }
static void toOrcHit(bool found, const Hit& h, OrcHit* o) {
  std::memset(o, 0, sizeof(*o));
  o->hit = found ? 1 : 0;
  if (!found) return;
  o->t = h.t;
  toArr(h.p, o->p);
  toArr(h.n, o->n);
  o->front = h.front ? 1 : 0;
  o->material = h.material;
}
static Hit fromOrcHit(const OrcHit& o) {
  Hit h;
  h.t = o.t;
  h.p = fromArr(o.p);
  h.n = fromArr(o.n);
  h.front = o.front != 0;
  h.material = o.material;
  return h;
}
static void toOrcCounters(const Counters& c, OrcCounters* o) {
  o->extend_rays = c.extend_rays; o->shadow_rays = c.shadow_rays; o->node_pops = c.node_pops;
  o->leaf_pops = c.leaf_pops; o->tri_tests = c.tri_tests; o->sph_tests = c.sph_tests;
  o->plane_tests = c.plane_tests; o->paths = c.paths;
}

// ------------------------------------------------------------- scene-file reader
// scenefile.zig:35-203: same directives, defaults, validation order and
// diagnostic messages.
namespace {
struct ParseFail {
  const char* msg;
};
struct LineParser {
  std::vector<std::string> toks;
  size_t pos = 0;
  bool more() const { return pos < toks.size(); }
  [[noreturn]] void fail(const char* m) { throw ParseFail{m}; }
  std::string word() {  // scenefile.zig:167-169
    if (!more()) fail("unexpected end of line");
    return toks[pos++];
  }
  // std.fmt.parseFloat(f64): optional sign, decimal or hex digits (with
  // single underscores between digits), optional exponent, or inf/nan.
  static bool parseFloatTok(const std::string& tok, double* out) {
    std::string s;
    size_t i = 0;
    if (tok.empty()) return false;
    if (tok[0] == '+' || tok[0] == '-') s.push_back(tok[i++]);
    std::string rest = tok.substr(i);
    std::string lower;
    for (char c : rest) lower.push_back((char)std::tolower((unsigned char)c));
    if (lower == "inf" || lower == "infinity") { *out = s == "-" ? -kInf : kInf; return true; }
    if (lower == "nan") { *out = std::numeric_limits<double>::quiet_NaN(); return true; }
    bool hex = lower.size() > 2 && lower[0] == '0' && lower[1] == 'x';
    bool any_digit = false;
    char prev = 0;
    for (size_t k = 0; k < rest.size(); ++k) {
      const char c = rest[k];
      if (c == '_') {
        const char nx = k + 1 < rest.size() ? rest[k + 1] : 0;
        const bool dprev = hex ? std::isxdigit((unsigned char)prev) : std::isdigit((unsigned char)prev);
        const bool dnext = hex ? std::isxdigit((unsigned char)nx) : std::isdigit((unsigned char)nx);
        if (!dprev || !dnext) return false;
        prev = c;
        continue;
      }
      if (std::isdigit((unsigned char)c)) any_digit = true;
      s.push_back(c);
      prev = c;
    }
    if (!any_digit) return false;
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == nullptr || *end != '\0') return false;
    *out = v;
    return true;
  }
  double number() {  // scenefile.zig:171-176
    const std::string tok = word();
    double v;
    if (!parseFloatTok(tok, &v)) fail("expected a number");
    if (!std::isfinite(v)) fail("number must be finite");
    return v;
  }
  uint64_t positiveInt() {  // scenefile.zig:178-183 (std.fmt.parseInt(usize, tok, 10))
    const std::string tok = word();
    size_t i = 0;
    if (!tok.empty() && tok[0] == '+') i = 1;
    if (i >= tok.size()) fail("expected a positive integer");
    unsigned __int128 v = 0;
    char prev = 0;
    for (size_t k = i; k < tok.size(); ++k) {
      const char c = tok[k];
      if (c == '_') {
        const char nx = k + 1 < tok.size() ? tok[k + 1] : 0;
        if (!std::isdigit((unsigned char)prev) || !std::isdigit((unsigned char)nx)) fail("expected a positive integer");
        prev = c;
        continue;
      }
      if (!std::isdigit((unsigned char)c)) fail("expected a positive integer");
      v = v * 10 + (unsigned)(c - '0');
      if (v > 0xFFFFFFFFFFFFFFFFull) fail("expected a positive integer");
      prev = c;
    }
    if (v == 0) fail("expected a positive integer");
    return (uint64_t)v;
  }
  V3 vec() {
    const double a = number();
    const double b = number();
    const double c = number();
    return v3(a, b, c);
  }
  V3 color() {  // scenefile.zig:189-193
    const V3 c = vec();
    if (c.x < 0 || c.y < 0 || c.z < 0) fail("colour components must be non-negative");
    return c;
  }
  uint32_t materialRef(const std::map<std::string, uint32_t>& names) {
    const std::string n = word();
    auto it = names.find(n);
    if (it == names.end()) fail("undefined material");
    return it->second;
  }
  void end() {
    if (more()) fail("unexpected extra tokens");
  }
};

Material parseMaterial(LineParser& p) {  // scenefile.zig:140-156
  const std::string kind = p.word();
  if (kind == "lambertian") return Material{kLambertian, p.color(), 0};
  if (kind == "metal") {
    const V3 albedo = p.color();
    const double fuzz = p.more() ? p.number() : 0;
    if (fuzz < 0 || fuzz > 1) p.fail("metal fuzz must be in [0, 1]");
    return Material{kMetal, albedo, fuzz};
  }
  if (kind == "dielectric") {
    const double ior = p.number();
    if (!(ior > 0)) p.fail("index of refraction must be positive");
    return Material{kDielectric, splat(0), ior};
  }
  if (kind == "emissive") return Material{kEmissive, p.color(), 0};
  p.fail("unknown material type");
}

void parseCamera(LineParser& p, Scene& scene) {  // scenefile.zig:107-138
  CameraConfig cfg = scene.camera;
  bool saw_from = false, saw_to = false;
  while (p.more()) {
    const std::string key = p.word();
    if (key == "from") { cfg.from = p.vec(); saw_from = true; }
    else if (key == "to") { cfg.at = p.vec(); saw_to = true; }
    else if (key == "up") { cfg.up = p.vec(); }
    else if (key == "fov") {
      cfg.vfov = p.number();
      if (!(cfg.vfov > 0 && cfg.vfov < 180)) p.fail("fov must be between 0 and 180 degrees");
    } else if (key == "aperture") {
      cfg.aperture = p.number();
      if (cfg.aperture < 0) p.fail("aperture must be non-negative");
    } else if (key == "focus") {
      const double f = p.number();
      if (!(f > 0)) p.fail("focus distance must be positive");
      cfg.focus = f;
      cfg.has_focus = true;
    } else p.fail("unknown camera parameter");
  }
  if (!saw_from || !saw_to) p.fail("camera needs 'from' and 'to'");
  const V3 view = sub(cfg.from, cfg.at);
  if (nearZero(view)) p.fail("camera 'from' and 'to' must differ");
  if (nearZero(cross(cfg.up, view))) p.fail("camera 'up' is parallel to the view direction");
  scene.camera = cfg;
}

// Returns 0 on success, 1 on a syntax error (line/message set), 2+ on a
// build error (kUnboundedShape/kTooManyPrimitives + 1).
int parseScene(const char* src, size_t len, Scene* scene, uint64_t* line_out, std::string* msg) {
  std::map<std::string, uint32_t> names;
  uint64_t line_no = 0;
  size_t start = 0;
  try {
    for (;;) {
      size_t nl = start;
      while (nl < len && src[nl] != '\n') ++nl;
      std::string raw(src + start, nl - start);
      line_no += 1;
      *line_out = line_no;
      const size_t hash = raw.find('#');
      if (hash != std::string::npos) raw.resize(hash);
      LineParser p;
      {
        size_t i = 0;
        while (i < raw.size()) {
          while (i < raw.size() && (raw[i] == ' ' || raw[i] == '\t' || raw[i] == '\r')) ++i;
          size_t j = i;
          while (j < raw.size() && !(raw[j] == ' ' || raw[j] == '\t' || raw[j] == '\r')) ++j;
          if (j > i) p.toks.push_back(raw.substr(i, j - i));
          i = j;
        }
      }
      if (p.more()) {
        const std::string kw = p.word();
        if (kw == "image") {
          scene->settings.width = p.positiveInt();
          scene->settings.height = p.positiveInt();
        } else if (kw == "samples") {
          scene->settings.spp = p.positiveInt();
        } else if (kw == "depth") {
          scene->settings.max_depth = p.positiveInt();
        } else if (kw == "camera") {
          parseCamera(p, *scene);
        } else if (kw == "background") {
          const std::string kind = p.word();
          if (kind == "color") {
            scene->background.kind = 0;
            scene->background.a = p.vec();
          } else if (kind == "sky") {
            const V3 h = p.vec();
            const V3 z = p.vec();
            scene->background.kind = 1;
            scene->background.a = h;
            scene->background.b = z;
          } else p.fail("background must be 'color' or 'sky'");
        } else if (kw == "material") {
          const std::string name = p.word();
          if (names.count(name)) p.fail("material already defined");
          const Material m = parseMaterial(p);
          scene->materials.push_back(m);
          names[name] = (uint32_t)(scene->materials.size() - 1);
        } else if (kw == "sphere") {
          const V3 c = p.vec();
          const double r = p.number();
          if (r == 0) p.fail("sphere radius must be non-zero");
          const uint32_t m = p.materialRef(names);
          scene->add(Shape{kSphere, m, c, v3(r, 0, 0), v3(0, 0, 0)});
        } else if (kw == "plane") {
          const V3 pt = p.vec();
          const V3 n = p.vec();
          if (nearZero(n)) p.fail("plane normal must be non-zero");
          const uint32_t m = p.materialRef(names);
          scene->add(Shape{kPlane, m, pt, normalize(n), v3(0, 0, 0)});
        } else if (kw == "triangle") {
          const V3 a = p.vec();
          const V3 b = p.vec();
          const V3 c = p.vec();
          if (nearZero(cross(sub(b, a), sub(c, a)))) p.fail("degenerate triangle");
          const uint32_t m = p.materialRef(names);
          scene->add(Shape{kTriangle, m, a, b, c});
        } else if (kw == "quad") {
          const V3 c = p.vec();
          const V3 u = p.vec();
          const V3 v = p.vec();
          if (nearZero(cross(u, v))) p.fail("degenerate quad");
          const uint32_t m = p.materialRef(names);
          scene->addQuad(c, u, v, m);
        } else {
          p.fail("unknown directive");
        }
        p.end();
      }
      if (nl >= len) break;
      start = nl + 1;
    }
  } catch (const ParseFail& f) {
    *msg = f.msg;
    return 1;
  }
  *line_out = 0;
  msg->clear();
  const int e = scene->build();
  return e == kBuildOk ? 0 : 1 + e;
}
}  // namespace

extern "C" {

// Test-only: the rounding-scale perturbation of traceWith (0 = off, the default).
void orc_set_perturbation(double rel) { orc::g_perturb_rel = rel; }

// ---- RNG (rng.zig)

void orc_splitmix64(uint64_t state, uint64_t n, uint64_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = splitMix64(&state);
}
// mode 0: Rng.init(seed); mode 1: Rng.forStream(seed, stream)
void orc_rng_state(uint64_t seed, uint64_t stream, int mode, uint64_t* s4) {
  const Rng r = mode == 0 ? Rng::init(seed) : Rng::forStream(seed, stream);
  std::memcpy(s4, r.s, 32);
}
// kind: 0 next() u64, 1 float() f64, 2 below(bound) u64, 3 unitVector (3 f64),
// 4 inUnitSphere (3 f64), 5 inUnitDisk (3 f64). State is read and written back.
void orc_rng_draw(uint64_t* s4, int kind, uint64_t bound, uint64_t n, void* out) {
  Rng r;
  std::memcpy(r.s, s4, 32);
  for (uint64_t i = 0; i < n; ++i) {
    switch (kind) {
      case 0: ((uint64_t*)out)[i] = r.next(); break;
      case 1: ((double*)out)[i] = r.uniform(); break;
      case 2: ((uint64_t*)out)[i] = r.below(bound); break;
      case 3: toArr(r.unitVector(), (double*)out + 3 * i); break;
      case 4: toArr(r.inUnitSphere(), (double*)out + 3 * i); break;
      default: toArr(r.inUnitDisk(), (double*)out + 3 * i); break;
    }
  }
  std::memcpy(s4, r.s, 32);
}

// ---- geometry / aabb / vec
int orc_shape_intersect(const OrcShape* s, const OrcRay* r, double tmin, double tmax, OrcHit* out) {
  Hit h;
  const bool f = shapeIntersect(toShape(*s), toRay(*r), tmin, tmax, &h);
  toOrcHit(f, h, out);
  return f ? 1 : 0;
}
int orc_shape_bounds(const OrcShape* s, double* mn, double* mx) {
  Aabb b;
  if (!shapeBounds(toShape(*s), &b)) return 0;
  toArr(b.mn, mn);
  toArr(b.mx, mx);
  return 1;
}
int orc_aabb_hit(const double* mn, const double* mx, const OrcRay* r, double tmin, double tmax, double* t) {
  const Ray ray = toRay(*r);
  return aabbHit(Aabb{fromArr(mn), fromArr(mx)}, ray, inverse(ray.d), tmin, tmax, t) ? 1 : 0;
}
// op 0: surfaceArea(box a..b) ; 1: pad(a..b, delta) -> out6 ; 2: longestAxis
double orc_aabb_op(int op, const double* mn, const double* mx, double delta, double* out6) {
  Aabb a{fromArr(mn), fromArr(mx)};
  if (op == 0) return surfaceArea(a);
  if (op == 1) {
    a = pad(a, delta);
    toArr(a.mn, out6);
    toArr(a.mx, out6 + 3);
    return 0;
  }
  return (double)longestAxis(a);
}
void orc_vec_refract(const double* uv, const double* n, double eta, double* out) {
  toArr(refract(fromArr(uv), fromArr(n), eta), out);
}
void orc_vec_reflect(const double* v, const double* n, double* out) { toArr(reflect(fromArr(v), fromArr(n)), out); }

// ---- materials
double orc_schlick(double c, double eta) { return schlick(c, eta); }
int orc_scatter(const OrcMaterial* m, const OrcRay* rin, const OrcHit* hit, uint64_t* s4, OrcRay* out, double* att) {
  Material mat{m->kind, fromArr(m->color), m->param};
  Rng r;
  std::memcpy(r.s, s4, 32);
  V3 a;
  Ray o;
  const bool ok = scatter(mat, toRay(*rin), fromOrcHit(*hit), &r, &a, &o);
  std::memcpy(s4, r.s, 32);
  if (ok) {
    toArr(o.o, out->o);
    toArr(o.d, out->d);
    toArr(a, att);
  }
  return ok ? 1 : 0;
}
void orc_emitted(const OrcMaterial* m, const OrcHit* hit, double* out) {
  Material mat{m->kind, fromArr(m->color), m->param};
  toArr(emitted(mat, fromOrcHit(*hit)), out);
}

// ---- camera
static CameraConfig toCam(const OrcCamera& c) {
  CameraConfig cfg;
  cfg.from = fromArr(c.from);
  cfg.at = fromArr(c.at);
  cfg.up = fromArr(c.up);
  cfg.vfov = c.vfov;
  cfg.aperture = c.aperture;
  cfg.has_focus = c.has_focus != 0;
  cfg.focus = c.focus;
  return cfg;
}
// out: origin, upper_left, horizontal, vertical, u, v (18) + lens_radius
void orc_camera_init(const OrcCamera* c, double aspect, double* out19) {
  const Camera cam = Camera::init(toCam(*c), aspect);
  toArr(cam.origin, out19);
  toArr(cam.upper_left, out19 + 3);
  toArr(cam.horizontal, out19 + 6);
  toArr(cam.vertical, out19 + 9);
  toArr(cam.u, out19 + 12);
  toArr(cam.v, out19 + 15);
  out19[18] = cam.lens_radius;
}
void orc_camera_rays(const OrcCamera* c, double aspect, const double* s, const double* t, uint64_t n, uint64_t* s4,
                     OrcRay* out) {
  const Camera cam = Camera::init(toCam(*c), aspect);
  Rng r;
  std::memcpy(r.s, s4, 32);
  for (uint64_t i = 0; i < n; ++i) {
    const Ray ray = cam.ray(s[i], t[i], &r);
    toArr(ray.o, out[i].o);
    toArr(ray.d, out[i].d);
  }
  std::memcpy(s4, r.s, 32);
}

// ---- scenes
void* orc_scene_new(const OrcMaterial* mats, uint64_t nm, const OrcShape* shapes, uint64_t ns, const OrcCamera* cam,
                    const OrcBackground* bg, int build, int* err) {
  Scene* s = new Scene();
  for (uint64_t i = 0; i < nm; ++i) s->materials.push_back({mats[i].kind, fromArr(mats[i].color), mats[i].param});
  for (uint64_t i = 0; i < ns; ++i) s->add(toShape(shapes[i]));
  if (cam) s->camera = toCam(*cam);
  if (bg) {
    s->background.kind = bg->kind;
    s->background.a = fromArr(bg->a);
    s->background.b = fromArr(bg->b);
  }
  *err = build ? s->build() : 0;
  return s;
}
void orc_scene_free(void* s) { delete (Scene*)s; }
// returns 0 ok, 1 syntax error, 2 unbounded, 3 too many primitives
int orc_scene_parse(const char* src, uint64_t len, void** out, uint64_t* diag_line, char* msg, int msgcap) {
  Scene* s = new Scene();
  std::string m;
  const int rc = parseScene(src, len, s, diag_line, &m);
  if (msgcap > 0) {
    std::snprintf(msg, (size_t)msgcap, "%s", m.c_str());
  }
  if (rc != 0) {
    delete s;
    *out = nullptr;
    return rc;
  }
  *out = s;
  return 0;
}
// info[0..]: materials, shapes, planes, nodes, lights, bvh depth, width, height, spp, max_depth, built
void orc_scene_info(void* sp, uint64_t* info) {
  const Scene* s = (const Scene*)sp;
  info[0] = s->materials.size();
  info[1] = s->shapes.size();
  info[2] = s->planes.size();
  info[3] = s->bvh.nodes.size();
  info[4] = s->lights.items.size();
  info[5] = s->bvh.nodes.empty() ? 0 : bvhDepthFrom(s->bvh, 0);
  info[6] = s->settings.width;
  info[7] = s->settings.height;
  info[8] = s->settings.spp;
  info[9] = s->settings.max_depth;
  info[10] = s->built ? 1 : 0;
}
void orc_scene_flat(void* sp, OrcMaterial* mats, OrcShape* shapes, OrcShape* planes, OrcCamera* cam,
                    OrcBackground* bg) {
  const Scene* s = (const Scene*)sp;
  for (size_t i = 0; i < s->materials.size(); ++i) {
    mats[i] = OrcMaterial{};
    mats[i].kind = s->materials[i].kind;
    toArr(s->materials[i].color, mats[i].color);
    mats[i].param = s->materials[i].param;
  }
  for (size_t i = 0; i < s->shapes.size(); ++i) shapes[i] = fromShape(s->shapes[i]);
  for (size_t i = 0; i < s->planes.size(); ++i) planes[i] = fromShape(s->planes[i]);
  *cam = OrcCamera{};
  toArr(s->camera.from, cam->from);
  toArr(s->camera.at, cam->at);
  toArr(s->camera.up, cam->up);
  cam->vfov = s->camera.vfov;
  cam->aperture = s->camera.aperture;
  cam->focus = s->camera.focus;
  cam->has_focus = s->camera.has_focus ? 1 : 0;
  *bg = OrcBackground{};
  bg->kind = s->background.kind;
  toArr(s->background.a, bg->a);
  toArr(s->background.b, bg->b);
}
void orc_scene_bvh(void* sp, OrcNode* nodes, uint32_t* indices) {
  const Scene* s = (const Scene*)sp;
  for (size_t i = 0; i < s->bvh.nodes.size(); ++i) {
    const Node& n = s->bvh.nodes[i];
    toArr(n.bounds.mn, nodes[i].mn);
    toArr(n.bounds.mx, nodes[i].mx);
    nodes[i].first = n.first;
    nodes[i].count = n.count;
  }
  std::memcpy(indices, s->bvh.indices.data(), s->bvh.indices.size() * 4);
}
void orc_scene_lights(void* sp, uint32_t* shape_index, double* radiance, double* weight, double* cdf, double* total) {
  const Scene* s = (const Scene*)sp;
  for (size_t i = 0; i < s->lights.items.size(); ++i) {
    shape_index[i] = s->lights.items[i].shape_index;
    toArr(s->lights.items[i].radiance, radiance + 3 * i);
    weight[i] = s->lights.items[i].weight;
    cdf[i] = s->lights.cdf[i];
  }
  *total = s->lights.total;
}
// Standalone Bvh.build over a shape list (planes included, to test
// UnboundedShape). Returns 0 / kUnboundedShape / kTooManyPrimitives.
int orc_bvh_build(const OrcShape* shapes, uint64_t n, OrcNode* nodes, uint64_t* nnodes, uint32_t* indices) {
  std::vector<Shape> v;
  for (uint64_t i = 0; i < n; ++i) v.push_back(toShape(shapes[i]));
  Bvh b;
  const int e = bvhBuild(v, &b);
  if (e != kBuildOk) return e;
  *nnodes = b.nodes.size();
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    toArr(b.nodes[i].bounds.mn, nodes[i].mn);
    toArr(b.nodes[i].bounds.mx, nodes[i].mx);
    nodes[i].first = b.nodes[i].first;
    nodes[i].count = b.nodes[i].count;
  }
  std::memcpy(indices, b.indices.data(), b.indices.size() * 4);
  return 0;
}
// mode 0: Scene.intersect; 1: brute force over bounded shapes then planes
// (geometry.intersectAll); 2: Bvh.intersect only (bounded shapes).
void orc_scene_intersect(void* sp, const OrcRay* rays, uint64_t n, double tmin, double tmax, int mode, OrcHit* out,
                         OrcCounters* ctr) {
  const Scene* s = (const Scene*)sp;
  Counters c;
  for (uint64_t i = 0; i < n; ++i) {
    const Ray r = toRay(rays[i]);
    Hit h;
    bool f;
    if (mode == 0) f = s->intersect(r, tmin, tmax, &h, &c);
    else if (mode == 2) f = bvhIntersect(s->bvh, s->shapes, r, tmin, tmax, &h, &c);
    else {
      f = intersectAll(s->shapes, r, tmin, tmax, &h, &c);
      Hit hp;
      if (intersectAll(s->planes, r, tmin, f ? h.t : tmax, &hp, &c)) { h = hp; f = true; }
    }
    toOrcHit(f, h, &out[i]);
  }
  if (ctr) toOrcCounters(c, ctr);
}
void orc_background(void* sp, const double* dir, double* out) {
  toArr(((const Scene*)sp)->background.radiance(fromArr(dir)), out);
}
int orc_light_sample(void* sp, const double* origin, uint64_t* s4, OrcLightSample* out) {
  const Scene* s = (const Scene*)sp;
  Rng r;
  std::memcpy(r.s, s4, 32);
  LightSample ls;
  const bool ok = s->lights.sample(fromArr(origin), &r, &ls);
  std::memcpy(s4, r.s, 32);
  std::memset(out, 0, sizeof(*out));
  out->valid = ok ? 1 : 0;
  if (ok) {
    toArr(ls.dir, out->dir);
    out->distance = ls.distance;
    toArr(ls.radiance, out->radiance);
    out->pdf = ls.pdf;
  }
  return ok ? 1 : 0;
}
double orc_light_pdf_hit(void* sp, const double* origin, const OrcHit* hit) {
  return ((const Scene*)sp)->lights.pdfHit(fromArr(origin), fromOrcHit(*hit));
}
// traceWith for n rays, each with its own RNG state (4 u64, read and written back).
void orc_trace(void* sp, const OrcRay* rays, uint64_t n, uint64_t* states, uint64_t max_depth, int light_sampling,
               double* out, OrcCounters* ctr) {
  const Scene* s = (const Scene*)sp;
  Counters c;
  for (uint64_t i = 0; i < n; ++i) {
    Rng r;
    std::memcpy(r.s, states + 4 * i, 32);
    const V3 v = traceWith(*s, toRay(rays[i]), &r, max_depth, light_sampling != 0, &c);
    std::memcpy(states + 4 * i, r.s, 32);
    toArr(v, out + 3 * i);
  }
  if (ctr) toOrcCounters(c, ctr);
}
struct OrcOptions {
  uint64_t width, height, spp, max_depth, threads, seed, tile_size;
  uint32_t light_sampling, _pad;
};
// render.zig:152-180. Returns 0 or 1 (InvalidOptions). `pixel_stride` > 1
// renders only pixels whose raster index is a multiple of it (a bounded
// sample for timing); the others stay zero.
int orc_render(void* sp, const OrcOptions* o, double* out, uint64_t pixel_stride, OrcCounters* ctr) {
  const Scene* s = (const Scene*)sp;
  if (o->width == 0 || o->height == 0 || o->spp == 0 || o->tile_size == 0) return 1;
  std::memset(out, 0, o->width * o->height * 3 * sizeof(double));
  Options opts{o->width, o->height, o->spp, o->max_depth, o->threads, o->seed, o->tile_size, o->light_sampling, 0};
  const double aspect = (double)opts.width / (double)opts.height;
  const uint64_t tiles_x = (opts.width + opts.tile_size - 1) / opts.tile_size;
  const uint64_t tiles_y = (opts.height + opts.tile_size - 1) / opts.tile_size;
  RenderCtx ctx;
  ctx.scene = s;
  ctx.camera = Camera::init(s->camera, aspect);
  ctx.opts = opts;
  ctx.image = out;
  ctx.tiles_x = tiles_x;
  ctx.tile_count = tiles_x * tiles_y;
  ctx.pixel_stride = pixel_stride == 0 ? 1 : pixel_stride;
  const uint64_t nthreads = std::min<uint64_t>(std::max<uint64_t>(opts.threads, 1), ctx.tile_count);
  std::vector<Counters> per(nthreads);
  std::vector<std::thread> pool;
  for (uint64_t i = 1; i < nthreads; ++i) pool.emplace_back([&ctx, &per, i] { ctx.work(&per[i]); });
  ctx.work(&per[0]);
  for (auto& t : pool) t.join();
  if (ctr) {
    Counters all;
    for (const Counters& c : per) all.merge(c);
    toOrcCounters(all, ctr);
  }
  return 0;
}
// Renders like orc_render and returns dLoss/dparams = sum over pixels of
// dimg . dpixel/dparams (dimg: width*height*3 adjoint of the image) for the
// material-colour parameters given by `offsets` (per material, -1 = none).
int orc_render_grad(void* sp, const OrcOptions* o, const int32_t* offsets, int nparams, const double* dimg,
                    double* image, double* grad) {
  const Scene* s = (const Scene*)sp;
  if (o->width == 0 || o->height == 0 || o->spp == 0 || o->tile_size == 0) return 1;
  if (nparams > kMaxGradParams) return 2;
  const uint64_t W = o->width, H = o->height;
  const Camera cam = Camera::init(s->camera, (double)W / (double)H);
  const GradCtx g{offsets, nparams};
  const uint64_t nthreads = std::max<uint64_t>(1, std::min<uint64_t>(o->threads, H));
  std::vector<std::vector<double>> part(nthreads, std::vector<double>(nparams, 0.0));
  std::atomic<uint64_t> next_row{0};
  auto work = [&](uint64_t t) {
    DV3* dv = new DV3;
    for (;;) {
      const uint64_t y = next_row.fetch_add(1);
      if (y >= H) break;
      for (uint64_t x = 0; x < W; ++x) {
        const uint64_t pix = y * W + x;
        Rng rng = Rng::forStream(o->seed, pix);
        V3 sum = splat(0);
        const double inv_spp = 1.0 / (double)o->spp;
        for (uint64_t k = 0; k < o->spp; ++k) {
          const double su = ((double)x + rng.uniform()) / (double)W;
          const double tv = ((double)y + rng.uniform()) / (double)H;
          const Ray cr = cam.ray(su, tv, &rng);
          const V3 c = traceWithGrad(*s, cr, &rng, o->max_depth, o->light_sampling != 0, g, dv);
          if (!isFinite(c)) continue;
          sum = add(sum, c);
          for (int q = 0; q < nparams; ++q)
            part[t][q] += (dimg[3 * pix] * dv->d[q].x + dimg[3 * pix + 1] * dv->d[q].y + dimg[3 * pix + 2] * dv->d[q].z) *
                          inv_spp;
        }
        const V3 px = scale(sum, inv_spp);
        image[3 * pix] = px.x;
        image[3 * pix + 1] = px.y;
        image[3 * pix + 2] = px.z;
      }
    }
    delete dv;
  };
  std::vector<std::thread> pool;
  for (uint64_t t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (int q = 0; q < nparams; ++q) {
    double acc = 0;
    for (uint64_t t = 0; t < nthreads; ++t) acc += part[t][q];
    grad[q] = acc;
  }
  return 0;
}

uint64_t orc_hardware_threads() { return std::max(1u, std::thread::hardware_concurrency()); }

}  // extern "C"
