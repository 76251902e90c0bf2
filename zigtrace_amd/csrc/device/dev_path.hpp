// Path state and the per-event integrator steps (render.zig:38-102,
// material.zig:26-70, light.zig:106-198, camera.zig:59-67, scene.zig:20-28),
// shared by the megakernel and the wavefront kernels so that both produce the
// same bits for the same seed.
//
// A path advances one ray query at a time: an extension ray (closest hit) or
// a pending next-event shadow ray (any hit). All RNG draws of a bounce
// (light pick, light point, scatter, roulette) happen in `shade`, in the
// reference's order; the shadow ray only decides whether its precomputed
// contribution is added, before the next extension ray is shaded, so the
// floating-point order of additions into the path radiance is the
// reference's as well.
#pragma once

#include "dev_trace.hpp"
#include "zt_types.h"

namespace ztd {

enum Phase : int { kCamera = 0, kExtend = 1, kShadow = 2, kFinish = 3 };

template <typename R>
struct Path {
  V3<R> ro, rd;  // extension ray; ro is also the shadow-ray origin
  V3<R> radiance, throughput;
  R prev_pdf;  // BSDF pdf of the bounce that cast rd from ro (0: no MIS weight)
  V3<R> sh_d;        // pending shadow ray direction
  R sh_tmax;         // distance * (1 - 1e-6)
  V3<R> sh_contrib;  // throughput * directLight, added if unoccluded
  int self;          // primitive the current rays leave from (-1: camera)
  int sh_excl;       // emitter primitive of the shadow ray
  int depth;
  int terminated;  // path ends after the pending shadow ray
};

struct Params {  // per-render constants (render.zig:12-23)
  uint64_t seed;
  uint32_t width, height, spp, max_depth;
  uint32_t use_nee;  // light_sampling && the scene has lights (render.zig:62)
  uint32_t _pad;
};

// Status words every render kernel receives (`int* status`, 32 bytes,
// cleared before the launch): word 0 holds error bits, words 2-3 and 4-5 the
// 64-bit counts of extension rays (each Scene.intersect of render.zig:73) and
// next-event shadow rays (render.zig:51) the launch traced.
constexpr size_t kStatusBytes = 32;
ZT_D void addRayCounts(int* status, uint64_t ext, uint64_t sh) {
  if (ext) atomicAdd((unsigned long long*)(status + 2), (unsigned long long)ext);
  if (sh) atomicAdd((unsigned long long*)(status + 4), (unsigned long long)sh);
}

// Rays traced by one lane, for kernels whose lanes trace independently.
struct RayTally {
  uint32_t ext = 0, sh = 0;
  ZT_D void count(bool shadow) {
    if (shadow) ++sh;
    else ++ext;
  }
  // Adds the wave's lanes' counts to the status words with one atomic pair
  // per wave. Every lane of the wave must call it together.
  ZT_D void flush(int* status) const {
    uint64_t e = ext, s = sh;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      e += __shfl_xor(e, o);
      s += __shfl_xor(s, o);
    }
    if (__lane_id() == 0) addRayCounts(status, e, s);
  }
};

// ---------------------------------------------------------------- camera.zig:59-67 (+ render.zig:138-140)
// Pixel jitter, lens sample and ray in f64 for both precisions.
template <typename R>
ZT_D void cameraRay(const CamD& cam, uint32_t x, uint32_t y, double w, double h, Rng& rng, V3<R>* o, V3<R>* d) {
  const double s = ((double)x + rng.u64f()) / w;
  const double t = ((double)y + rng.u64f()) / h;
  double ox = cam.origin[0], oy = cam.origin[1], oz = cam.origin[2];
  if (cam.lens_radius > 0) {
    double a, b;
    for (;;) {  // rng.zig:74-79
      a = -1.0 + 2.0 * rng.u64f();
      b = -1.0 + 2.0 * rng.u64f();
      if (a * a + b * b + 0.0 * 0.0 < 1.0) break;
    }
    const double dx = a * cam.lens_radius, dy = b * cam.lens_radius;
    ox = (ox + cam.u[0] * dx) + cam.v[0] * dy;
    oy = (oy + cam.u[1] * dx) + cam.v[1] * dy;
    oz = (oz + cam.u[2] * dx) + cam.v[2] * dy;
  }
  const double tx = (cam.upper_left[0] + cam.horizontal[0] * s) + cam.vertical[0] * t;
  const double ty = (cam.upper_left[1] + cam.horizontal[1] * s) + cam.vertical[1] * t;
  const double tz = (cam.upper_left[2] + cam.horizontal[2] * s) + cam.vertical[2] * t;
  const V3<double> dir = normalize(mk(tx - ox, ty - oy, tz - oz));
  *o = mk((R)ox, (R)oy, (R)oz);
  *d = cvt<R>(dir);
}

// ---------------------------------------------------------------- scene.zig:20-28
template <typename R>
ZT_D V3<R> background(const SceneView<R>& sc, V3<R> dir) {
  if (sc.bg_kind == 0) return sc.bg_a;
  const R t = fmax(R(0), fmin(normalize(dir).y, R(1)));
  return lerp(sc.bg_a, sc.bg_b, t);
}

// ---------------------------------------------------------------- render.zig:38-41
template <typename R>
ZT_D R powerWeight(R a, R b) {
  const R a2 = a * a;
  return a2 / (a2 + b * b);
}

// ---------------------------------------------------------------- material.zig:66-70
template <typename R>
ZT_D R schlick(R cosine, R eta) {
  R r0 = (R(1) - eta) / (R(1) + eta);
  r0 *= r0;
  const R x = R(1) - cosine;
  const R x2 = x * x;
  return r0 + (R(1) - r0) * (x * (x2 * x2));  // std.math.pow(x, 5) by repeated squaring
}

// ---------------------------------------------------------------- light.zig:47-57
template <typename R>
ZT_D bool outside(V3<R> center, R radius, V3<R> origin) {
  const R d = length(sub(center, origin));
  if constexpr (K<R>::exact) return d > fabs(radius) * 1.000000001;
  else return d > fabs(radius);
}
template <typename R>
ZT_D void basis(V3<R> w, V3<R>* u, V3<R>* v) {
  const V3<R> a = fabs(w.x) > R(0.9) ? mk(R(0), R(1), R(0)) : mk(R(1), R(0), R(0));
  *v = normalize(cross(w, a));
  *u = cross(*v, w);
}
// 1 - cos(theta_max) of the cone a sphere of radius r subtends at squared
// distance d2: the reference's formula in f64, the cancellation-free
// (r^2/d2) / (1 + sqrt(1 - r^2/d2)) in fp32.
template <typename R>
ZT_D R oneMinusCosMax(R radius, R d2) {
  if constexpr (K<R>::exact) {
    const R cos_max = sqrt(fmax(R(0), 1 - radius * radius / d2));
    return 1 - cos_max;
  } else {
    const R x = radius * radius / d2;
    return x / (R(1) + sqrt(fmax(R(0), R(1) - x)));
  }
}

template <typename R>
struct LightSample {
  V3<R> dir;
  R distance;
  V3<R> radiance;
  R pdf;
  int prim;
  int material;
};

// Receives the contributions a path adds to its radiance, for the adjoint
// (path replay). Rendering uses NoObserver (with a null pointer), whose hooks
// are compiled out. The hooks are selected at compile time, not by testing
// the pointer: comparing a lane-private observer's address with null keeps
// the compiler from holding the observer in registers.
struct NoObserver {
  template <typename R>
  ZT_D void miss(V3<R>) {}
  template <typename R>
  ZT_D void emission(V3<R>, V3<R>, int) {}
  ZT_D void scattered(int) {}
  template <typename R>
  ZT_D void neePending(V3<R>, V3<R>, int) {}
  ZT_D void neeCommitted() {}
};
template <typename Obs>
constexpr bool kObserved = !std::is_same<Obs, NoObserver>::value;

// Lights.pick + Lights.sample (light.zig:106-166). The pick is done on the
// f64 CDF with the exact f64 uniform in both precisions, so it selects the
// same emitter as the reference for the same draw.
template <typename R>
ZT_D bool sampleLight(const SceneView<R>& sc, V3<R> origin, Rng& rng, LightSample<R>* out) {
  if (sc.num_lights == 0) return false;
  const double x = rng.u64f() * sc.total;
  int lo = 0, hi = sc.num_lights - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sc.cdf[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  const Light<R> L = sc.lights[lo];
  const R select = L.select;
  if (L.kind == (int)kKindTriangle) {
    const R su = sqrt(uniform<R>(rng));
    const R b2 = uniform<R>(rng);
    const V3<R> p = add(add(scale(L.p0, R(1) - su), scale(L.p1, su * (R(1) - b2))), scale(L.p2, su * b2));
    const V3<R> to = sub(p, origin);
    const R dist = length(to);
    if (dist < R(1e-9)) return false;
    const V3<R> dir = scale(to, R(1) / dist);
    const R cos_l = -dot(dir, L.normal);
    if (cos_l <= R(1e-9)) return false;
    *out = {dir, dist, L.radiance, select * dist * dist / (cos_l * L.area), L.prim, L.material};
    return true;
  }
  if (!outside(L.p0, L.radius, origin)) return false;
  const V3<R> axis = sub(L.p0, origin);
  const R d2 = len2(axis);
  const V3<R> w = scale(axis, R(1) / sqrt(d2));
  const R omc = oneMinusCosMax(L.radius, d2);
  const R u1 = uniform<R>(rng);
  const R cos_t = R(1) - u1 * omc;
  R sin_t;
  if constexpr (K<R>::exact) sin_t = sqrt(fmax(R(0), 1 - cos_t * cos_t));
  else sin_t = sqrt(fmax(R(0), (u1 * omc) * (R(2) - u1 * omc)));
  const R phi = K<R>::two_pi * uniform<R>(rng);
  V3<R> bu, bv;
  basis(w, &bu, &bv);
  R sp, cp;
  sincosT(phi, &sp, &cp);
  const V3<R> dir = normalize(add(add(scale(bu, sin_t * cp), scale(bv, sin_t * sp)), scale(w, cos_t)));
  // The cone is sampled, not the surface: find where the ray lands.
  Prim<R> sph;
  sph.a = L.p0;
  sph.b = mk(L.radius, R(0), R(0));
  sph.c = mk(R(0), R(0), R(0));
  sph.meta = 0;
  R t;
  if constexpr (K<R>::exact) {
    if (!sphereHit<R>(sph, origin, dir, R(0), R(INFINITY), &t)) return false;
    const V3<R> hp = add(origin, scale(dir, t));
    const V3<R> outward = vdiv(sub(hp, L.p0), L.radius);
    if (!(dot(dir, outward) < 0)) return false;
  } else {
    // Every direction inside the cone meets the sphere's near side; clamp
    // the discriminant so fp32 rounding at the rim cannot lose the sample.
    const V3<R> f = sub(origin, L.p0);
    const R h = -dot(f, dir);
    const V3<R> l = add(f, scale(dir, h / len2(dir)));
    const R disc = fmax(R(0), L.radius * L.radius - len2(l));
    t = h - sqrt(disc * len2(dir));
    t = t / len2(dir);
    if (!(t > 0)) return false;
  }
  *out = {dir, t, L.radiance, select / (K<R>::two_pi * omc), L.prim, L.material};
  return true;
}

// Lights.pdfHit (light.zig:171-198) for a BSDF-sampled ray from `origin`
// that hit emitter primitive `prim` at `p`. The reference re-intersects every
// light; the emitter hit is known here, so only its own term is evaluated
// (other lights could only contribute where two emitters coincide).
template <typename R>
ZT_D R pdfHit(const SceneView<R>& sc, V3<R> origin, V3<R> p, int prim) {
  if (prim < 0 || prim >= sc.num_prims) return R(0);
  const int li = sc.prim_light[prim];
  if (li < 0) return R(0);
  const V3<R> to = sub(p, origin);
  const R dist = length(to);
  if (dist < R(1e-9)) return R(0);
  const V3<R> dir = scale(to, R(1) / dist);
  const Light<R> L = sc.lights[li];
  if (L.kind == (int)kKindTriangle) {
    const R c = dot(dir, L.normal);
    if (!(c < 0)) return R(0);  // the emitting side faces the origin
    return L.select * dist * dist / (-c * L.area);
  }
  if (!outside(L.p0, L.radius, origin)) return R(0);
  const R d2 = len2(sub(L.p0, origin));
  return L.select / (K<R>::two_pi * oneMinusCosMax(L.radius, d2));
}

// ---------------------------------------------------------------- material.zig:34-62
template <typename R>
ZT_D bool scatter(const Mat<R>& m, V3<R> rd, const Surface<R>& s, Rng& rng, V3<R>* att, V3<R>* dir) {
  switch (m.kind) {
    case ZT_LAMBERTIAN: {
      V3<R> dv = add(s.n, unitVector<R>(rng));
      if (nearZero(dv)) dv = s.n;
      *att = m.color;
      *dir = normalize(dv);
      return true;
    }
    case ZT_METAL: {
      V3<R> dv = reflect(normalize(rd), s.n);
      if (m.param > 0) dv = normalize(add(dv, scale(unitVector<R>(rng), m.param)));
      if (dot(dv, s.n) <= 0) return false;
      *att = m.color;
      *dir = dv;
      return true;
    }
    case ZT_DIELECTRIC: {
      const R eta = s.front ? R(1) / m.param : m.param;
      const V3<R> unit = normalize(rd);
      const R cos_theta = fmin(dot(neg(unit), s.n), R(1));
      const R sin_theta = sqrt(fmax(R(0), R(1) - cos_theta * cos_theta));
      const bool cannot_refract = eta * sin_theta > R(1);
      const bool refl = cannot_refract || schlick(cos_theta, eta) > uniform<R>(rng);
      *dir = refl ? reflect(unit, s.n) : refract(unit, s.n, eta);
      *att = splat(R(1));
      return true;
    }
    default:
      return false;
  }
}

// Start sample `k` of a pixel: camera ray and a fresh path (render.zig:63-71).
template <typename R>
ZT_D int beginPath(const SceneView<R>& sc, const Params& pr, uint32_t x, uint32_t y, Rng& rng, Path<R>& st) {
  cameraRay<R>(sc.cam, x, y, (double)pr.width, (double)pr.height, rng, &st.ro, &st.rd);
  st.radiance = splat(R(0));
  st.throughput = splat(R(1));
  st.prev_pdf = R(0);
  st.self = -1;
  st.sh_excl = -1;
  st.depth = 0;
  st.terminated = 0;
  return pr.max_depth == 0 ? kFinish : kExtend;
}

// One bounce after the extension ray's query (render.zig:73-99). Returns the
// next phase: kShadow when a next-event ray is pending, kExtend to continue,
// kFinish when the path is complete.
template <typename R, typename Obs = NoObserver>
ZT_D int shade(const SceneView<R>& sc, const Params& pr, HitRec<R> h, Rng& rng, Path<R>& st, Obs* obs = nullptr) {
  if (h.prim < 0) {
    const V3<R> x = mul(st.throughput, background(sc, st.rd));
    st.radiance = add(st.radiance, x);
    if constexpr (kObserved<Obs>) obs->miss(x);
    return kFinish;
  }
  const Surface<R> s = surfaceAt(sc, st.ro, st.rd, h);
  const Mat<R> m = sc.mats[s.material];
  const bool emissive_front = m.kind == ZT_EMISSIVE && s.front;
  if (emissive_front && maxComponent(m.color) > 0) {
    // The previous vertex is this ray's origin (shade sets ro to the hit point).
    const R w = st.prev_pdf > 0 ? powerWeight(st.prev_pdf, pdfHit(sc, st.ro, s.p, h.prim)) : R(1);
    const V3<R> x = scale(mul(st.throughput, m.color), w);
    st.radiance = add(st.radiance, x);
    if constexpr (kObserved<Obs>) obs->emission(x, scale(st.throughput, w), s.material);
  }
  bool pending = false;
  const bool lambertian = m.kind == ZT_LAMBERTIAN;
  if (lambertian && pr.use_nee) {  // directLight, render.zig:46-56
    LightSample<R> ls;
    if (sampleLight(sc, s.p, rng, &ls)) {
      const R cos_s = dot(s.n, ls.dir);
      if (cos_s > 0 && ls.pdf > 0) {
        const R bsdf_pdf = cos_s / K<R>::pi;
        const R sc_ = cos_s / (K<R>::pi * ls.pdf) * powerWeight(ls.pdf, bsdf_pdf);
        st.sh_contrib = mul(st.throughput, scale(mul(m.color, ls.radiance), sc_));
        st.sh_d = ls.dir;
        st.sh_tmax = ls.distance * R(0.999999);
        st.sh_excl = ls.prim;
        pending = true;
        if constexpr (kObserved<Obs>)
          obs->neePending(st.sh_contrib, scale(mul(st.throughput, m.color), sc_), ls.material);
      }
    }
  }
  V3<R> att, dir;
  const bool scattered = scatter(m, st.rd, s, rng, &att, &dir);
  st.ro = s.p;
  st.self = h.prim;
  bool alive = scattered;
  if (scattered) {
    if constexpr (kObserved<Obs>)
      if (m.kind == ZT_LAMBERTIAN || m.kind == ZT_METAL) obs->scattered(s.material);
    st.throughput = mul(st.throughput, att);
    st.prev_pdf = (pr.use_nee && lambertian) ? fmax(dot(dir, s.n), R(0)) / K<R>::pi : R(0);
    st.rd = dir;
    if ((uint32_t)st.depth + 1 >= 4) {  // Russian roulette, render.zig:94-99
      const R p = fmin(maxComponent(st.throughput), R(0.95));
      if (p <= 0 || uniform<R>(rng) >= p) alive = false;
      else st.throughput = vdiv(st.throughput, p);
    }
    st.depth += 1;
    if ((uint32_t)st.depth >= pr.max_depth) alive = false;
  }
  st.terminated = alive ? 0 : 1;
  if (pending) return kShadow;
  return alive ? kExtend : kFinish;
}

// The shadow ray's outcome (render.zig:51): add the contribution if nothing
// blocks the segment, then continue or finish.
template <typename R, typename Obs = NoObserver>
ZT_D int afterShadow(HitRec<R> h, Path<R>& st, Obs* obs = nullptr) {
  if (h.prim < 0) {
    st.radiance = add(st.radiance, st.sh_contrib);
    if constexpr (kObserved<Obs>) obs->neeCommitted();
  }
  return st.terminated ? kFinish : kExtend;
}

}  // namespace ztd
