// Device-side math, RNG and scene layout shared by every kernel.
//
// Templated on the arithmetic type R: float is the product path, double the
// verification path that repeats the reference's f64 arithmetic operation
// for operation (the library is compiled with -ffp-contract=off, so no FMA is
// formed behind our back and the two instantiations of one source line stay
// deterministic across kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>

#define ZT_D __device__ __forceinline__
#define ZT_HD __host__ __device__ __forceinline__

namespace ztd {

// Overloads that pick the float or double routine on both host and device.
ZT_HD float sqrtT(float x) { return sqrtf(x); }
ZT_HD double sqrtT(double x) { return sqrt(x); }
ZT_HD float fminT(float a, float b) { return fminf(a, b); }
ZT_HD double fminT(double a, double b) { return fmin(a, b); }
ZT_HD float fmaxT(float a, float b) { return fmaxf(a, b); }
ZT_HD double fmaxT(double a, double b) { return fmax(a, b); }
ZT_HD float fabsT(float a) { return fabsf(a); }
ZT_HD double fabsT(double a) { return fabs(a); }

template <typename R>
struct V3 {
  R x, y, z;
};

template <typename R> ZT_HD V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> ZT_HD V3<R> splat(R v) { return V3<R>{v, v, v}; }
template <typename R> ZT_HD V3<R> add(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> ZT_HD V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> ZT_HD V3<R> mul(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename R> ZT_HD V3<R> scale(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
// vec.zig:38-40: division is a multiply by the reciprocal.
template <typename R> ZT_HD V3<R> vdiv(V3<R> a, R s) { return scale(a, R(1) / s); }
template <typename R> ZT_HD V3<R> neg(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> ZT_HD R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> ZT_HD V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> ZT_HD R len2(V3<R> a) { return dot(a, a); }
template <typename R> ZT_HD R length(V3<R> a) { return sqrtT(len2(a)); }
template <typename R> ZT_HD V3<R> normalize(V3<R> a) { return vdiv(a, length(a)); }
template <typename R> ZT_HD V3<R> lerp(V3<R> a, V3<R> b, R t) { return add(scale(a, R(1) - t), scale(b, t)); }
template <typename R> ZT_HD R maxComponent(V3<R> a) { return fmaxT(a.x, fmaxT(a.y, a.z)); }
template <typename R> ZT_D bool nearZero(V3<R> a) {
  const R eps = R(1e-8);
  return fabs(a.x) < eps && fabs(a.y) < eps && fabs(a.z) < eps;
}
template <typename R> ZT_D bool isFinite(V3<R> a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
template <typename R> ZT_HD V3<R> reflect(V3<R> v, V3<R> n) { return sub(v, scale(n, R(2) * dot(v, n))); }
template <typename R> ZT_HD V3<R> refract(V3<R> uv, V3<R> n, R eta) {  // vec.zig:119-125
  const R cos_theta = fminT(dot(neg(uv), n), R(1));
  const V3<R> r_perp = scale(add(uv, scale(n, cos_theta)), eta);
  const V3<R> r_par = scale(n, -sqrtT(fabsT(R(1) - len2(r_perp))));
  return add(r_perp, r_par);
}
template <typename R> ZT_D R getc(V3<R> a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }
template <typename T, typename R> ZT_HD V3<T> cvt(V3<R> a) { return V3<T>{(T)a.x, (T)a.y, (T)a.z}; }

template <typename R> struct K;
template <> struct K<double> {
  static constexpr double pi = 3.14159265358979311600;      // f64(pi)
  static constexpr double two_pi = 6.28318530717958623200;  // f64(2*pi) == 2*f64(pi)
  static constexpr bool exact = true;
};
template <> struct K<float> {
  static constexpr float pi = 3.14159265358979311600f;
  static constexpr float two_pi = 6.28318530717958623200f;
  static constexpr bool exact = false;
};

// ---------------------------------------------------------------- rng.zig:10-88
ZT_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
ZT_HD uint64_t splitmix64(uint64_t* state) {
  *state += 0x9E3779B97F4A7C15ull;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s0, s1, s2, s3;
  ZT_HD uint64_t next() {  // xoshiro256++, rng.zig:27-38
    const uint64_t result = rotl64(s0 + s3, 23) + s0;
    const uint64_t t = s1 << 17;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl64(s3, 45);
    return result;
  }
  // rng.zig:41-43, exact f64 uniform in [0, 1).
  ZT_HD double u64f() { return (double)(next() >> 11) * 0x1.0p-53; }
  // The fp32 path takes the top 24 bits of the same draw: the f64 uniform
  // rounded toward zero to 24 bits, so both paths consume identical streams.
  ZT_HD float u32f() { return (float)(uint32_t)(next() >> 40) * 0x1.0p-24f; }
};
ZT_HD Rng rngInit(uint64_t seed) {  // rng.zig:13-18
  uint64_t sm = seed;
  Rng r;
  r.s0 = splitmix64(&sm);
  r.s1 = splitmix64(&sm);
  r.s2 = splitmix64(&sm);
  r.s3 = splitmix64(&sm);
  return r;
}
ZT_HD Rng rngForStream(uint64_t seed, uint64_t stream) {  // rng.zig:22-25
  uint64_t sm = seed ^ (stream * 0x9E3779B97F4A7C15ull);
  return rngInit(splitmix64(&sm));
}
template <typename R> ZT_D R uniform(Rng& r);
template <> ZT_D double uniform<double>(Rng& r) { return r.u64f(); }
template <> ZT_D float uniform<float>(Rng& r) { return r.u32f(); }

ZT_D void sincosT(float x, float* s, float* c) { sincosf(x, s, c); }
ZT_D void sincosT(double x, double* s, double* c) { sincos(x, s, c); }

template <typename R> ZT_D V3<R> unitVector(Rng& rng) {  // rng.zig:58-63
  const R z = R(-1) + R(2) * uniform<R>(rng);
  const R phi = K<R>::two_pi * uniform<R>(rng);
  const R r = sqrt(fmax(R(0), R(1) - z * z));
  R s, c;
  sincosT(phi, &s, &c);
  return mk(r * c, r * s, z);
}

// ---------------------------------------------------------------- device scene layout
// BVH node: the boxes of both children of one interior node of the host BVH
// (conservatively rounded to fp32), so one 64-byte fetch serves one pop.
//   a = (l.min.x, l.min.y, l.min.z, l.max.x)
//   b = (l.max.y, l.max.z, r.min.x, r.min.y)
//   c = (r.min.z, r.max.x, r.max.y, r.max.z)
//   d = (left code, right code, -, -)
// code >= 0: interior device node; code == kNoChild: empty slot;
// code < 0 otherwise: leaf, -1 - code == (first << 4) | count, count in 1..15.
struct __align__(16) DNode {
  float4 a, b, c;
  int4 d;
};
static constexpr int kNoChild = INT32_MIN;

// Four-wide node of the fp32 product path: one 128-byte line holding four
// child boxes (the same conservatively rounded fp32 boxes as the two-child
// nodes, laid out per coordinate so a node is eight 16-byte loads) and their
// codes (as in DNode; code >= 0 indexes the four-wide array). Empty slots
// have code kNoChild and an inverted box.
struct __align__(128) DNode4 {
  float4 lox, loy, loz, hix, hiy, hiz;
  int4 child;
  int4 _pad;
};
static_assert(sizeof(DNode4) == 128, "one four-wide node per 128-byte line");
// Most four-wide nodes a scene may have: node byte offsets (index * 128)
// are 32-bit, so traversal addresses a node as base + one 32-bit offset.
static constexpr uint32_t kMaxNodes4 = 1u << 25;
static_assert((uint64_t)kMaxNodes4 * sizeof(DNode4) <= (1ull << 32), "32-bit node offsets");

// The same four-wide node in 64 bytes: child boxes as 8-bit offsets from a
// per-node origin in steps of a power of two per axis. A bound decodes as
// fma(q, 2^e, origin) -- the product is exact, so one rounding -- and the
// host picks every q so that the decoded box contains the fp32 box it
// replaces. qlo / qhi hold byte k for child k; an empty slot has code
// kNoChild and an inverted box.
struct __align__(64) DNode4Q {
  float ox, oy, oz;
  uint32_t exps;  // bytes 0..2: e_x + 127, e_y + 127, e_z + 127
  int4 child;
  uint32_t qlox, qhix, qloy, qhiy, qloz, qhiz;
  uint32_t _pad[2];
};
static_assert(sizeof(DNode4Q) == 64, "one compressed four-wide node per 64 bytes");
static constexpr uint32_t kMaxLeafCount = 15;
static constexpr uint32_t kMaxLeafFirst = (1u << 27) - 1;
ZT_HD int leafCode(uint32_t first, uint32_t count) { return -1 - (int)((first << 4) | count); }

// Primitive records in leaf order. kind in bits 30-31 of meta, material in 0-29.
//   fp32: p0 = (v0 or centre, meta bits), p1 = (e1 or (radius,0,0)),
//         p2 = (e2 or (0,0,0), meta bits again)
//   fp64: a[0..2] v0/centre, a[3..5] e1/(radius,..), a[6..8] e2
struct __align__(16) PrimF {
  float4 p0, p1, p2;
};
struct __align__(16) PrimD {
  double a[9];
  uint32_t meta;
  uint32_t _pad;
};
static constexpr uint32_t kKindSphere = 0, kKindTriangle = 1;
ZT_HD uint32_t metaKind(uint32_t m) { return m >> 30; }
ZT_HD uint32_t metaMaterial(uint32_t m) { return m & 0x3FFFFFFFu; }

template <typename R>
struct Plane {
  V3<R> point;
  V3<R> normal;
  int material;
  int _pad;
};

template <typename R>
struct Mat {
  V3<R> color;  // albedo or emitted colour
  R param;      // fuzz or ior
  int kind;     // ZT_LAMBERTIAN ...
  int _pad;
};

// One sampled emitter (light.zig:31-36), with everything `sample` and
// `pdfHit` need precomputed in f64 on the host.
template <typename R>
struct Light {
  V3<R> p0, p1, p2;  // triangle v0 v1 v2; sphere centre in p0
  V3<R> radiance;
  V3<R> normal;  // triangle: normalize(cross(v1-v0, v2-v0))
  R radius;      // sphere
  R area;        // triangle: 0.5*|cross(v1-v0, v2-v0)|
  R select;      // weight / total
  int kind;      // kKindSphere / kKindTriangle
  int prim;      // device primitive slot of the emitter
  int material;  // emissive material of the emitter
  int _pad;
};

// Camera after Camera.init (camera.zig:29-55), always f64: ray generation is
// done in f64 by both paths so primary rays match the reference to rounding.
struct CamD {
  double origin[3], upper_left[3], horizontal[3], vertical[3], u[3], v[3];
  double lens_radius;
};

template <typename R>
struct SceneView {
  const DNode* nodes;
  const DNode4* nodes4;    // fp32 only: four-wide BVH (null: two-child traversal)
  const DNode4Q* nodes4q;  // fp32 only: the four-wide BVH in 64-byte nodes (null: use nodes4)
  const void* prims;  // PrimF for float, PrimD for double
  const Plane<R>* planes;
  const Mat<R>* mats;
  const Light<R>* lights;
  const double* cdf;
  const int* prim_light;  // primitive slot -> light index or -1
  int num_nodes;
  int num_prims;
  int num_planes;
  int num_lights;
  double total;
  CamD cam;
  int bg_kind;
  int _pad;
  V3<R> bg_a, bg_b;
};

// 16-byte loads from device memory through the global address space. The
// batched kernels read their scene's arrays through pointers that are
// themselves loaded from the scene table, which the compiler cannot prove
// global; it would emit flat loads, which also count against (and wait
// with) the wave's LDS traffic. Every pointer passed here addresses device
// memory.
typedef float GF4 __attribute__((ext_vector_type(4)));
typedef int GI4 __attribute__((ext_vector_type(4)));
ZT_D GF4 loadGF4(const void* base, uint64_t byte_off) {
  return *(const __attribute__((address_space(1))) GF4*)((const __attribute__((address_space(1))) char*)base +
                                                          byte_off);
}
ZT_D V3<float> loadGF3(const void* base, uint64_t byte_off) {
  const __attribute__((address_space(1))) float* p =
      (const __attribute__((address_space(1))) float*)((const __attribute__((address_space(1))) char*)base + byte_off);
  return V3<float>{p[0], p[1], p[2]};
}
ZT_D GI4 loadGI4(const void* base, uint64_t byte_off) {
  return *(const __attribute__((address_space(1))) GI4*)((const __attribute__((address_space(1))) char*)base +
                                                          byte_off);
}

// Decoded primitive for the intersection routines.
template <typename R>
struct Prim {
  V3<R> a, b, c;  // triangle v0 e1 e2; sphere centre (radius in b.x)
  uint32_t meta;
};
template <typename R> ZT_D Prim<R> loadPrim(const SceneView<R>& sc, int k);
template <> ZT_D Prim<float> loadPrim<float>(const SceneView<float>& sc, int k) {
  // The kind is read from the third vector, which every caller then needs
  // before it branches on the kind: the three loads go out together (with
  // the kind taken from the first, a triangle's e2 waited for it to return).
  const uint64_t at = (uint64_t)k * sizeof(PrimF);
  const GF4 p2 = loadGF4(sc.prims, at + 32);
  Prim<float> r;
  r.a = loadGF3(sc.prims, at);
  r.b = loadGF3(sc.prims, at + 16);
  r.c = mk(p2.x, p2.y, p2.z);
  r.meta = __float_as_uint(p2.w);
  return r;
}
template <> ZT_D Prim<double> loadPrim<double>(const SceneView<double>& sc, int k) {
  const PrimD* p = (const PrimD*)sc.prims + k;
  Prim<double> r;
  r.a = mk(p->a[0], p->a[1], p->a[2]);
  r.b = mk(p->a[3], p->a[4], p->a[5]);
  r.c = mk(p->a[6], p->a[7], p->a[8]);
  r.meta = p->meta;
  return r;
}

}  // namespace ztd
