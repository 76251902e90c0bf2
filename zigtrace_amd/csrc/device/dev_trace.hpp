// Ray/primitive tests and the nearest-child-first BVH traversal.
//
// Semantics follow geometry.zig:37-99 (open interval (t_min, t_max), near
// root before far root, Moller-Trumbore with |det| < 1e-12 rejection) and
// bvh.zig:80-126 (prune popped entries whose entry distance exceeds the
// closest hit; push the farther child first; ties go left).
//
// Two arithmetic regimes:
//   * double: the reference's f64 formulas operation for operation. BVH
//     boxes are the fp32 boxes rounded outward, so traversal visits a
//     superset of the reference's nodes and the closest hit is unchanged.
//   * float: the product path. Self-intersection is avoided by identity, not
//     by an offset: a ray leaving a flat primitive never tests it again, and
//     a ray leaving a sphere only finds that sphere's far side (t = 2h/a).
//     The reference's t_min = 1e-4 still applies to every other primitive.
#pragma once

#include "dev_common.hpp"

namespace ztd {

// Closest hit (or any hit) result: t and the primitive id, where ids
// [0, num_prims) are leaf-ordered BVH primitives and num_prims + j is plane j.
template <typename R>
struct HitRec {
  R t;
  int prim;  // -1: no hit
};

template <typename R>
ZT_D bool sphereHit(const Prim<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout) {
  const R radius = p.b.x;
  if constexpr (K<R>::exact) {  // geometry.zig:37-49
    const V3<R> oc = sub(p.a, o);
    const R a = len2(d);
    const R h = dot(d, oc);
    const R c = len2(oc) - radius * radius;
    const R disc = h * h - a * c;
    if (disc < 0) return false;
    const R sq = sqrt(disc);
    R t = (h - sq) / a;
    if (t <= tmin || t >= tmax) {
      t = (h + sq) / a;
      if (t <= tmin || t >= tmax) return false;
    }
    *tout = t;
    return true;
  } else {
    // Same roots, computed with the perpendicular-distance discriminant and
    // the stable quadratic form (Haines et al., Ray Tracing Gems ch. 7).
    const V3<R> f = sub(o, p.a);
    const R a = len2(d);
    const R h = -dot(f, d);
    const V3<R> l = add(f, scale(d, h / a));
    const R disc = a * (radius * radius - len2(l));
    if (disc < 0) return false;
    const R c = len2(f) - radius * radius;
    const R q = h + copysign(sqrt(disc), h);
    const R t0 = c / q, t1 = q / a;
    const R tn = fmin(t0, t1), tf = fmax(t0, t1);
    R t = tn;
    if (!(t > tmin && t < tmax)) {
      t = tf;
      if (!(t > tmin && t < tmax)) return false;
    }
    *tout = t;
    return true;
  }
}

// A ray that starts on sphere `p`: its only possible hit on that sphere is
// the far end of the chord, t = 2h/a (the near root is the origin itself).
template <typename R>
ZT_D bool selfSphereHit(const Prim<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout) {
  const R h = dot(d, sub(p.a, o));
  const R t = R(2) * h / len2(d);
  if (!(t > tmin && t < tmax)) return false;
  *tout = t;
  return true;
}

template <typename R>
ZT_D bool triangleHit(const Prim<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout) {  // geometry.zig:83-97
  const V3<R> e1 = p.b, e2 = p.c;
  const V3<R> pv = cross(d, e2);
  const R det = dot(e1, pv);
  if (fabs(det) < R(1e-12)) return false;
  const R inv_det = R(1) / det;
  const V3<R> s = sub(o, p.a);
  const R u = dot(s, pv) * inv_det;
  if (u < 0 || u > 1) return false;
  const V3<R> q = cross(s, e1);
  const R v = dot(d, q) * inv_det;
  if (v < 0 || u + v > 1) return false;
  const R t = dot(e2, q) * inv_det;
  if (t <= tmin || t >= tmax) return false;
  *tout = t;
  return true;
}

template <typename R>
ZT_D bool planeHit(const Plane<R>& pl, V3<R> o, V3<R> d, R tmin, R tmax, R* tout) {  // geometry.zig:67-73
  const R denom = dot(pl.normal, d);
  if (fabs(denom) < R(1e-12)) return false;
  const R t = dot(sub(pl.point, o), pl.normal) / denom;
  if (t <= tmin || t >= tmax) return false;
  *tout = t;
  return true;
}

// Slab test of one child box (aabb.zig:68-84): lo starts at t_min, hi at the
// current closest hit; a NaN slab (0*inf) leaves the interval unchanged,
// which fmax/fmin reproduce. The fp32 path widens the far distance by
// 1 + 2*gamma(3) so rounding never loses a box the exact test would hit.
// slabNear: the same test given each axis's near plane (n*) and far plane
// (f*) along the ray.
template <typename R>
ZT_D bool slabNear(float nx, float ny, float nz, float fx, float fy, float fz, V3<R> o, V3<R> inv, R tmin, R tmax,
                   R* tout) {
  const R tx0 = (R(nx) - o.x) * inv.x;
  const R tx1 = (R(fx) - o.x) * inv.x;
  const R ty0 = (R(ny) - o.y) * inv.y;
  const R ty1 = (R(fy) - o.y) * inv.y;
  const R tz0 = (R(nz) - o.z) * inv.z;
  const R tz1 = (R(fz) - o.z) * inv.z;
  const R lo = fmax(fmax(tmin, tx0), fmax(ty0, tz0));
  R far = fmin(tx1, fmin(ty1, tz1));
  if constexpr (!K<R>::exact) far = far * (1.0f + 3.0f * 0x1.0p-23f);
  const R hi = fmin(tmax, far);
  *tout = lo;
  return !(hi < lo);
}
// The fp32 four-wide walker's form of slabNear: each plane distance is one
// fma, p * inv - oi, with the ray's origin-over-direction oi = o * inv taken
// from per-ray offsets shifted to cover its rounding (onear >= o * inv for
// near planes, ofar <= o * inv for far planes), so a near distance is never
// above the exact (p - o) * inv and a far distance never below it, up to the
// fma's own rounding, which the far distance's widening covers as in
// slabNear. Boxes the exact test hits are never lost.
ZT_D bool slabFma(float nx, float ny, float nz, float fx, float fy, float fz, V3<float> inv, V3<float> onear,
                  V3<float> ofar, float tmin, float tmax, float* tout) {
  const float tx0 = __builtin_fmaf(nx, inv.x, -onear.x);
  const float tx1 = __builtin_fmaf(fx, inv.x, -ofar.x);
  const float ty0 = __builtin_fmaf(ny, inv.y, -onear.y);
  const float ty1 = __builtin_fmaf(fy, inv.y, -ofar.y);
  const float tz0 = __builtin_fmaf(nz, inv.z, -onear.z);
  const float tz1 = __builtin_fmaf(fz, inv.z, -ofar.z);
  const float lo = fmaxf(fmaxf(tmin, tx0), fmaxf(ty0, tz0));
  const float far = fminf(tx1, fminf(ty1, tz1)) * (1.0f + 3.0f * 0x1.0p-23f);
  const float hi = fminf(tmax, far);
  *tout = lo;
  return !(hi < lo);
}
template <typename R>
ZT_D bool slab(float bx0, float by0, float bz0, float bx1, float by1, float bz1, V3<R> o, V3<R> inv, bool nx,
               bool ny, bool nz, R tmin, R tmax, R* tout) {
  return slabNear<R>(nx ? bx1 : bx0, ny ? by1 : by0, nz ? bz1 : bz0, nx ? bx0 : bx1, ny ? by0 : by1, nz ? bz0 : bz1,
                     o, inv, tmin, tmax, tout);
}

// Per-lane traversal stack: the top `cap` entries live in LDS (laid out
// [entry][lane], 4-byte words, so a wave's accesses are bank-conflict free
// whatever its lanes' depths), the rest spill in halves to a global buffer
// the launch provides, kSpill entries per launched thread laid out
// [entry][thread] (coalesced across a wave). Nothing goes to private scratch,
// whose size would otherwise cap how many waves run at once.
static constexpr int kLds = 16;  // default LDS entries per lane
static constexpr int kSpill = 80;
// Deepest device tree the stack can traverse for the smallest LDS capacity
// used (8 entries), checked at upload: a two-child BVH needs at most
// depth + 1 entries with nearest-first pushes, a four-wide one 3 * depth + 1.
static constexpr int kMaxTreeDepth = 8 + kSpill - 1;

struct LaneStack {
  int* code;    // LDS: code[i * stride]
  float* tval;  // LDS: tval[i * stride]
  int stride;
  int cap;           // LDS entries per lane (even, >= 8)
  int* scode;        // global spill: scode[i * sstride]
  float* stval;      // global spill: stval[i * sstride]
  uint32_t sstride;  // threads in the launch
};

// LDS bytes of a stack of `cap` entries per lane for `threads` lanes.
ZT_HD size_t stackBytes(int cap, int threads) { return (size_t)2 * cap * threads * sizeof(int); }
// Global spill bytes for a launch of `threads` threads.
ZT_HD size_t spillBytes(uint64_t threads) { return (size_t)2 * kSpill * threads * sizeof(int); }
// kLanes: the workgroup size when the kernel fixes it (then the entry
// addresses take a shift instead of a 32-bit multiply), else 0.
template <int kLanes = 0>
ZT_D LaneStack laneStack(int* lds, int cap, int* spill) {
  const uint32_t lanes = kLanes ? (uint32_t)kLanes : blockDim.x;
  const uint32_t threads = gridDim.x * lanes;
  const uint32_t gtid = blockIdx.x * lanes + threadIdx.x;
  return LaneStack{lds + threadIdx.x, (float*)(lds + cap * lanes) + threadIdx.x, (int)lanes, cap,
                   spill + gtid, (float*)(spill + (size_t)kSpill * threads) + gtid, threads};
}

template <typename R>
ZT_D float entryT(R t) {
  if constexpr (K<R>::exact) return __double2float_rd(t);  // rounding down keeps pruning conservative
  else return t;
}

// Tests the primitives of leaf `code` in slot order (scene.zig:100-104 via
// bvh.zig:108-117), shrinking *closest. Returns true when an any-hit query
// found its hit.
// Primitive k against a ray whose interval is (tmin, tmax); the fp32 path
// skips the shadow ray's target and treats the ray's own primitive by
// identity (a sphere then yields only the far end of its chord).
template <typename R>
ZT_D bool primHit(const SceneView<R>& sc, int k, V3<R> o, V3<R> d, R tmin, R tmax, int self, int excl, R* t) {
  constexpr bool kIdentity = !K<R>::exact;
  if (kIdentity && k == excl) return false;
  const Prim<R> p = loadPrim<R>(sc, k);
  if (metaKind(p.meta) == kKindTriangle) return (kIdentity && k == self) ? false : triangleHit<R>(p, o, d, tmin, tmax, t);
  return (kIdentity && k == self) ? selfSphereHit<R>(p, o, d, tmin, tmax, t) : sphereHit<R>(p, o, d, tmin, tmax, t);
}

template <typename R>
ZT_D bool leafHits(const SceneView<R>& sc, int code, V3<R> o, V3<R> d, R tmin, R* closest, int* best, int self,
                   int excl, bool anyhit) {
  const int v = -1 - code;
  const int first = v >> 4, count = v & 15;
  for (int k = first; k < first + count; ++k) {
    R t;
    if (primHit<R>(sc, k, o, d, tmin, *closest, self, excl, &t)) {
      *closest = t;
      *best = k;
      if (anyhit) return true;
    }
  }
  return false;
}

// Planes after the BVH (scene.zig:113-114).
template <typename R>
ZT_D void planeHits(const SceneView<R>& sc, V3<R> o, V3<R> d, R tmin, R* closest, int* best, int self, int excl,
                    bool anyhit) {
  constexpr bool kIdentity = !K<R>::exact;
  for (int j = 0; j < sc.num_planes; ++j) {
    const int id = sc.num_prims + j;
    if (kIdentity && (id == self || id == excl)) continue;
    R t;
    if (planeHit<R>(sc.planes[j], o, d, tmin, *closest, &t)) {
      *closest = t;
      *best = id;
      if (anyhit) return;
    }
  }
}

// Nearest-hit (anyhit == false) or any-hit query over BVH + planes.
// self/excl: primitive ids to treat specially (fp32 only, see above); -1 none.
// Sets *overflow if the stack would exceed its capacity (cannot happen for a
// host BVH of depth <= 66, kept as a checked invariant).
template <typename R>
ZT_D HitRec<R> traverse(const SceneView<R>& sc, const LaneStack& stk, V3<R> o, V3<R> d, R tmin, R tmax, int self,
                        int excl, bool anyhit, int* overflow) {
  constexpr bool kIdentity = !K<R>::exact;
  R closest = tmax;
  int best = -1;
  bool done = false;
  if (sc.num_nodes > 0) {
    const V3<R> inv = mk(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    const bool nx = inv.x < 0, ny = inv.y < 0, nz = inv.z < 0;
    int sp = 0, nsp = 0;
    stk.code[0] = 0;
    stk.tval[0] = entryT<R>(tmin);
    sp = 1;
    const int half = stk.cap >> 1;
    auto push = [&](int c, R t) {
      if (sp == stk.cap) {
        if (nsp + half > kSpill) {
          *overflow = 1;
          return;
        }
        for (int i = 0; i < half; ++i) {
          stk.scode[(size_t)(nsp + i) * stk.sstride] = stk.code[i * stk.stride];
          stk.stval[(size_t)(nsp + i) * stk.sstride] = stk.tval[i * stk.stride];
          stk.code[i * stk.stride] = stk.code[(i + half) * stk.stride];
          stk.tval[i * stk.stride] = stk.tval[(i + half) * stk.stride];
        }
        nsp += half;
        sp = half;
      }
      stk.code[sp * stk.stride] = c;
      stk.tval[sp * stk.stride] = entryT<R>(t);
      ++sp;
    };
    for (;;) {
      if (sp == 0) {
        if (nsp == 0) break;
        nsp -= half;
        for (int i = 0; i < half; ++i) {
          stk.code[i * stk.stride] = stk.scode[(size_t)(nsp + i) * stk.sstride];
          stk.tval[i * stk.stride] = stk.stval[(size_t)(nsp + i) * stk.sstride];
        }
        sp = half;
      }
      --sp;
      const int code = stk.code[sp * stk.stride];
      const float et = stk.tval[sp * stk.stride];
      if (R(et) > closest) continue;
      if (code >= 0) {
        const DNode n = sc.nodes[code];
        R tl, tr;
        const bool hl = slab<R>(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, o, inv, nx, ny, nz, tmin, closest, &tl) &&
                        n.d.x != kNoChild;
        const bool hr = slab<R>(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, o, inv, nx, ny, nz, tmin, closest, &tr) &&
                        n.d.y != kNoChild;
        if (hl && hr) {
          const bool near_left = tl <= tr;
          push(near_left ? n.d.y : n.d.x, near_left ? tr : tl);
          push(near_left ? n.d.x : n.d.y, near_left ? tl : tr);
        } else if (hl) {
          push(n.d.x, tl);
        } else if (hr) {
          push(n.d.y, tr);
        }
      } else if (leafHits<R>(sc, code, o, d, tmin, &closest, &best, self, excl, anyhit)) {
        done = true;
        break;
      }
    }
  }
  if (!done) planeHits<R>(sc, o, d, tmin, &closest, &best, self, excl, anyhit);
  return HitRec<R>{closest, best};
}

// The four bounds packed in `q` (byte k = child k) decoded as
// fma(q_k, s, origin): q_k * s is exact (s is a power of two), so the result
// is the one rounding of the exact bound, as the host assumed.
ZT_D float4 decodeBounds(uint32_t q, float s, float origin) {
  return make_float4(__builtin_fmaf((float)(q & 0xFFu), s, origin), __builtin_fmaf((float)((q >> 8) & 0xFFu), s, origin),
                     __builtin_fmaf((float)((q >> 16) & 0xFFu), s, origin), __builtin_fmaf((float)(q >> 24), s, origin));
}

// Compare-exchange of (entry distance, child code) pairs: after it, a holds
// the nearer entry.
ZT_D void sortChild(float& ka, int& ca, float& kb, int& cb) {
  const bool sw = ka > kb;
  const float k = sw ? kb : ka;
  const int c = sw ? cb : ca;
  kb = sw ? ka : kb;
  cb = sw ? ca : cb;
  ka = k;
  ca = c;
}

// Brute force over every primitive in slot order, then planes
// (geometry.zig:129-139 and scene.zig:109-117): the reference the BVH
// traversal is checked against on the device, and the path for scenes
// handed over without a BVH.
template <typename R>
ZT_D HitRec<R> bruteForce(const SceneView<R>& sc, V3<R> o, V3<R> d, R tmin, R tmax, int self, int excl,
                          bool anyhit) {
  constexpr bool kIdentity = !K<R>::exact;
  R closest = tmax;
  int best = -1;
  for (int k = 0; k < sc.num_prims; ++k) {
    R t;
    if (primHit<R>(sc, k, o, d, tmin, closest, self, excl, &t)) {
      closest = t;
      best = k;
      if (anyhit) return HitRec<R>{closest, best};
    }
  }
  for (int j = 0; j < sc.num_planes; ++j) {
    const int id = sc.num_prims + j;
    if (kIdentity && (id == self || id == excl)) continue;
    R t;
    if (planeHit<R>(sc.planes[j], o, d, tmin, closest, &t)) {
      closest = t;
      best = id;
      if (anyhit) break;
    }
  }
  return HitRec<R>{closest, best};
}

// The four-wide walk as a resumable state machine whose state lives in
// registers: `start` sets up a query; `inner` does one pop or one interior
// node (the four slab tests, a 5-comparator sorting network on (entry
// distance, code), the farther hits pushed farthest first, the nearest held
// in `cur`); `leaf` tests the leaf held in `cur`; `finish` adds the planes
// and returns the hit. Popped entries beyond the closest hit are dropped.
// The closest hit is the primitive the two-child traversal finds, except
// that two primitives at exactly the same t may be resolved the other way
// (the visiting order differs). `traverse4` drives one query to the end;
// the wave-scheduled megakernel interleaves many.
// The query a walk answers; kept by the caller (the wave-scheduled kernel
// rebuilds it from the path state), so the walk itself holds only what
// changes.
// A ray query of the resumable walkers (Trav4, Trav2).
template <typename R>
struct QueryR {
  V3<R> o, d;
  R tmin;
  int self, excl;
  bool anyhit;
};
using Query4 = QueryR<float>;

struct Trav4 {
  V3<float> inv;
  V3<float> onear, ofar;  // o * inv rounded, raised / lowered by 2^-22 of itself (slabFma)
  float closest;
  int best, cur, sp, nsp;
  bool found;

  ZT_D void start(const Query4& q, float tmax) {
    // A zero direction component (a draw of exactly 0.5 makes one, e.g. off
    // an axis-aligned plane) has an infinite reciprocal, and o * inf minus
    // itself would be NaN: that axis would stop culling and the ray walk most
    // of the tree. Reciprocals are held to +-2^100 instead, which still culls
    // every box the exact test culls but those whose face lies within
    // 2^-100 * tmin of the origin on that axis.
    inv = mk(fminf(fmaxf(1.0f / q.d.x, -0x1.0p100f), 0x1.0p100f),
             fminf(fmaxf(1.0f / q.d.y, -0x1.0p100f), 0x1.0p100f),
             fminf(fmaxf(1.0f / q.d.z, -0x1.0p100f), 0x1.0p100f));
    // |fl(o * inv) - o * inv| <= 2^-24 |fl(o * inv)|; the shift (an exact
    // power-of-two scaling) covers that and the rounding of the sum.
    const V3<float> oi = mk(q.o.x * inv.x, q.o.y * inv.y, q.o.z * inv.z);
    const V3<float> e = mk(fabsf(oi.x) * 0x1.0p-22f, fabsf(oi.y) * 0x1.0p-22f, fabsf(oi.z) * 0x1.0p-22f);
    onear = mk(oi.x + e.x, oi.y + e.y, oi.z + e.z);
    ofar = mk(oi.x - e.x, oi.y - e.y, oi.z - e.z);
    closest = tmax;
    best = -1;
    found = false;
    cur = 0;  // root
    sp = 0;
    nsp = 0;
  }
  ZT_D bool atLeaf() const { return cur < 0 && cur != kNoChild; }
  // Pushes the m (0..3) farther hits, farthest first, so the nearest of them
  // is popped first; one overflow check for the group.
  ZT_D void push3(const LaneStack& stk, int m, int c1, float k1, int c2, float k2, int c3, float k3,
                  int* overflow) {
    if (m == 0) return;
    if (sp + m > stk.cap) {
      const int half = stk.cap >> 1;
      if (nsp + half > kSpill) {
        *overflow = 1;
        return;
      }
      for (int i = 0; i < half; ++i) {
        stk.scode[(size_t)(nsp + i) * stk.sstride] = stk.code[i * stk.stride];
        stk.stval[(size_t)(nsp + i) * stk.sstride] = stk.tval[i * stk.stride];
      }
      for (int i = half; i < sp; ++i) {
        stk.code[(i - half) * stk.stride] = stk.code[i * stk.stride];
        stk.tval[(i - half) * stk.stride] = stk.tval[i * stk.stride];
      }
      nsp += half;
      sp -= half;
    }
    // c1 on top, then c2, then c3: each written to its slot, no selects.
    const int top = sp + m - 1;
    stk.code[top * stk.stride] = c1;
    stk.tval[top * stk.stride] = k1;
    if (m >= 2) {
      stk.code[(top - 1) * stk.stride] = c2;
      stk.tval[(top - 1) * stk.stride] = k2;
    }
    if (m == 3) {
      stk.code[(top - 2) * stk.stride] = c3;
      stk.tval[(top - 2) * stk.stride] = k3;
    }
    sp += m;
  }
  // One interior node, after popping until an entry survives pruning when
  // none is held (so a popped node costs no separate step). Returns true
  // when the BVH is exhausted; a popped leaf is held for the leaf step.
  // Q: the nodes are the 64-byte DNode4Q (else the 128-byte DNode4).
  template <bool Q>
  ZT_D bool inner(const SceneView<float>& sc, const LaneStack& stk, const Query4& q, int* overflow) {
    while (cur == kNoChild) {
      if (sp == 0) {
        if (nsp == 0) return true;
        const int half = stk.cap >> 1;
        nsp -= half;
        for (int i = 0; i < half; ++i) {
          stk.code[i * stk.stride] = stk.scode[(size_t)(nsp + i) * stk.sstride];
          stk.tval[i * stk.stride] = stk.stval[(size_t)(nsp + i) * stk.sstride];
        }
        sp = half;
      }
      --sp;
      // Both words read together (one LDS round trip per pop, not two).
      const float tv = stk.tval[sp * stk.stride];
      const int code = stk.code[sp * stk.stride];
      cur = tv <= closest ? code : kNoChild;
    }
    if (cur < 0) return false;
    float k0, k1, k2, k3;
    bool h0, h1, h2, h3;
    int c0, c1, c2, c3;
    const bool nx = inv.x < 0, ny = inv.y < 0, nz = inv.z < 0;
    if constexpr (Q) {
      float4 lx, ly, lz, hx, hy, hz;
      const uint4* p = (const uint4*)(sc.nodes4q + cur);
      const uint4 a = p[0], b = p[1], c = p[2];
      const float ox = __uint_as_float(a.x), oy = __uint_as_float(a.y), oz = __uint_as_float(a.z);
      const float sx = __uint_as_float((a.w & 0xFFu) << 23);
      const float sy = __uint_as_float(((a.w >> 8) & 0xFFu) << 23);
      const float sz = __uint_as_float(((a.w >> 16) & 0xFFu) << 23);
      c0 = (int)b.x;
      c1 = (int)b.y;
      c2 = (int)b.z;
      c3 = (int)b.w;
      lx = decodeBounds(c.x, sx, ox);
      hx = decodeBounds(c.y, sx, ox);
      ly = decodeBounds(c.z, sy, oy);
      hy = decodeBounds(c.w, sy, oy);
      const uint2 d = *(const uint2*)(p + 3);
      lz = decodeBounds(d.x, sz, oz);
      hz = decodeBounds(d.y, sz, oz);
      h0 = slab<float>(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, q.o, inv, nx, ny, nz, q.tmin, closest, &k0);
      h1 = slab<float>(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, q.o, inv, nx, ny, nz, q.tmin, closest, &k1);
      h2 = slab<float>(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, q.o, inv, nx, ny, nz, q.tmin, closest, &k2);
      h3 = slab<float>(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, q.o, inv, nx, ny, nz, q.tmin, closest, &k3);
    } else {
      // Each axis's near and far rows (lo, or hi when the direction is
      // negative) picked by address -- byte offsets lo 0/16/32, hi 48/64/80
      // -- instead of by a select per child and bound.
      const void* base = sc.nodes4;
      const uint32_t at = (uint32_t)cur << 7;
      const uint32_t ox = nx ? 48u : 0u, oy = ny ? 64u : 16u, oz = nz ? 80u : 32u;
      const GF4 ax = loadGF4(base, at | ox), fx = loadGF4(base, at | (ox ^ 48u));
      const GF4 ay = loadGF4(base, at | oy), fy = loadGF4(base, at | (oy ^ 80u));
      const GF4 az = loadGF4(base, at | oz), fz = loadGF4(base, at | (oz ^ 112u));
      const GI4 ch = loadGI4(base, at | 96u);
      c0 = ch.x;
      c1 = ch.y;
      c2 = ch.z;
      c3 = ch.w;
      h0 = slabFma(ax.x, ay.x, az.x, fx.x, fy.x, fz.x, inv, onear, ofar, q.tmin, closest, &k0);
      h1 = slabFma(ax.y, ay.y, az.y, fx.y, fy.y, fz.y, inv, onear, ofar, q.tmin, closest, &k1);
      h2 = slabFma(ax.z, ay.z, az.z, fx.z, fy.z, fz.z, inv, onear, ofar, q.tmin, closest, &k2);
      h3 = slabFma(ax.w, ay.w, az.w, fx.w, fy.w, fz.w, inv, onear, ofar, q.tmin, closest, &k3);
    }
    // Misses (and empty slots, whose boxes are inverted) sort last.
    const bool v0 = h0 && c0 != kNoChild, v1 = h1 && c1 != kNoChild, v2 = h2 && c2 != kNoChild,
               v3 = h3 && c3 != kNoChild;
    if (!v0) k0 = INFINITY, c0 = kNoChild;
    if (!v1) k1 = INFINITY, c1 = kNoChild;
    if (!v2) k2 = INFINITY, c2 = kNoChild;
    if (!v3) k3 = INFINITY, c3 = kNoChild;
    sortChild(k0, c0, k1, c1);
    sortChild(k2, c2, k3, c3);
    sortChild(k0, c0, k2, c2);
    sortChild(k1, c1, k3, c3);
    sortChild(k1, c1, k2, c2);
    const int hits = (int)v0 + (int)v1 + (int)v2 + (int)v3;
    push3(stk, hits > 0 ? hits - 1 : 0, c1, k1, c2, k2, c3, k3, overflow);
    cur = c0;  // kNoChild when nothing was hit: pop next
    return false;
  }
  // Tests the held leaf. Returns true when an any-hit query found its hit.
  ZT_D bool leaf(const SceneView<float>& sc, const Query4& q) {
    const int code = cur;
    cur = kNoChild;
    if (leafHits<float>(sc, code, q.o, q.d, q.tmin, &closest, &best, q.self, q.excl, q.anyhit)) {
      found = true;
      return true;
    }
    return false;
  }
  ZT_D HitRec<float> finish(const SceneView<float>& sc, const Query4& q) {
    if (!found) planeHits<float>(sc, q.o, q.d, q.tmin, &closest, &best, q.self, q.excl, q.anyhit);
    return HitRec<float>{closest, best};
  }
};

// One four-wide query to the end, while-while: a lane pops and visits
// interior nodes until it holds a leaf (or has nothing left), then tests that
// leaf, so a wave runs node work and leaf work in separate passes.
template <bool Q>
ZT_D HitRec<float> traverse4Impl(const SceneView<float>& sc, const LaneStack& stk, const Query4& q, float tmax,
                                 int* overflow) {
  Trav4 t;
  t.start(q, tmax);
  for (;;) {
    bool empty = false;
    while (!t.atLeaf()) {
      if (t.inner<Q>(sc, stk, q, overflow)) {
        empty = true;
        break;
      }
    }
    if (empty || t.leaf(sc, q)) break;
  }
  return t.finish(sc, q);
}
ZT_D HitRec<float> traverse4(const SceneView<float>& sc, const LaneStack& stk, V3<float> o, V3<float> d, float tmin,
                             float tmax, int self, int excl, bool anyhit, int* overflow) {
  const Query4 q{o, d, tmin, self, excl, anyhit};
  return sc.nodes4q ? traverse4Impl<true>(sc, stk, q, tmax, overflow) : traverse4Impl<false>(sc, stk, q, tmax, overflow);
}

// The two-child walk of `traverse` (bvh.zig:80-126), resumable one step at a
// time for the wave-scheduled kernels: the same entries are popped, pruned
// and tested in the same order with the same closest distance, so the hit,
// ties included, is the one `traverse` returns. A visited node's nearer hit
// child is held rather than pushed and popped straight away; pruning could
// not drop it (its entry distance is at most the closest hit). The tree must
// have a root (num_nodes > 0).
template <typename R>
struct Trav2 {
  V3<R> inv;
  R closest;
  int best, cur, sp, nsp;
  bool found;

  ZT_D void start(const QueryR<R>& q, R tmax) {
    inv = mk(R(1) / q.d.x, R(1) / q.d.y, R(1) / q.d.z);
    closest = tmax;
    best = -1;
    found = false;
    sp = 0;
    nsp = 0;
    cur = R(entryT<R>(q.tmin)) > tmax ? kNoChild : 0;  // the root, pruned as `traverse` prunes it
  }
  ZT_D bool atLeaf() const { return cur < 0 && cur != kNoChild; }
  ZT_D void push(const LaneStack& stk, int c, R t, int* overflow) {
    if (sp == stk.cap) {
      const int half = stk.cap >> 1;
      if (nsp + half > kSpill) {
        *overflow = 1;
        return;
      }
      for (int i = 0; i < half; ++i) {
        stk.scode[(size_t)(nsp + i) * stk.sstride] = stk.code[i * stk.stride];
        stk.stval[(size_t)(nsp + i) * stk.sstride] = stk.tval[i * stk.stride];
        stk.code[i * stk.stride] = stk.code[(i + half) * stk.stride];
        stk.tval[i * stk.stride] = stk.tval[(i + half) * stk.stride];
      }
      nsp += half;
      sp = half;
    }
    stk.code[sp * stk.stride] = c;
    stk.tval[sp * stk.stride] = entryT<R>(t);
    ++sp;
  }
  // Pops until an entry survives pruning (when none is held), then visits it
  // if it is a node. Returns true when the tree is exhausted; a leaf is held
  // for the leaf step.
  ZT_D bool inner(const SceneView<R>& sc, const LaneStack& stk, const QueryR<R>& q, int* overflow) {
    while (cur == kNoChild) {
      if (sp == 0) {
        if (nsp == 0) return true;
        const int half = stk.cap >> 1;
        nsp -= half;
        for (int i = 0; i < half; ++i) {
          stk.code[i * stk.stride] = stk.scode[(size_t)(nsp + i) * stk.sstride];
          stk.tval[i * stk.stride] = stk.stval[(size_t)(nsp + i) * stk.sstride];
        }
        sp = half;
      }
      --sp;
      const float tv = stk.tval[sp * stk.stride];
      const int code = stk.code[sp * stk.stride];
      cur = R(tv) > closest ? kNoChild : code;
    }
    if (cur < 0) return false;
    const DNode n = sc.nodes[cur];
    const bool nx = inv.x < 0, ny = inv.y < 0, nz = inv.z < 0;
    R tl, tr;
    const bool hl = slab<R>(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, q.o, inv, nx, ny, nz, q.tmin, closest, &tl) &&
                    n.d.x != kNoChild;
    const bool hr = slab<R>(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, q.o, inv, nx, ny, nz, q.tmin, closest, &tr) &&
                    n.d.y != kNoChild;
    if (hl && hr) {
      const bool near_left = tl <= tr;
      push(stk, near_left ? n.d.y : n.d.x, near_left ? tr : tl, overflow);
      cur = near_left ? n.d.x : n.d.y;
    } else {
      cur = hl ? n.d.x : (hr ? n.d.y : kNoChild);
    }
    return false;
  }
  // Tests the held leaf. Returns true when an any-hit query found its hit.
  ZT_D bool leaf(const SceneView<R>& sc, const QueryR<R>& q) {
    const int code = cur;
    cur = kNoChild;
    if (leafHits<R>(sc, code, q.o, q.d, q.tmin, &closest, &best, q.self, q.excl, q.anyhit)) {
      found = true;
      return true;
    }
    return false;
  }
  ZT_D HitRec<R> finish(const SceneView<R>& sc, const QueryR<R>& q) {
    if (!found) planeHits<R>(sc, q.o, q.d, q.tmin, &closest, &best, q.self, q.excl, q.anyhit);
    return HitRec<R>{closest, best};
  }
};

// One ray's closest-hit (or any-hit) query (scene.zig:108-116): brute force
// when the scene has no BVH, the four-wide BVH on the fp32 path when the view
// carries one, else the two-child traversal.
template <typename R>
ZT_D HitRec<R> query(const SceneView<R>& sc, const LaneStack& stk, V3<R> o, V3<R> d, R tmin, R tmax, int self,
                     int excl, bool anyhit, int* overflow) {
  if (sc.num_nodes == 0 && sc.num_prims > 0) return bruteForce<R>(sc, o, d, tmin, tmax, self, excl, anyhit);
  if constexpr (!K<R>::exact) {
    if (sc.nodes4) return traverse4(sc, stk, o, d, tmin, tmax, self, excl, anyhit, overflow);
  }
  return traverse<R>(sc, stk, o, d, tmin, tmax, self, excl, anyhit, overflow);
}

// Surface data of a hit (geometry.zig:20-29): point, ray-facing normal,
// front-face flag and material.
template <typename R>
struct Surface {
  V3<R> p, n;
  bool front;
  int material;
};

template <typename R>
ZT_D Surface<R> surfaceAt(const SceneView<R>& sc, V3<R> o, V3<R> d, HitRec<R> h) {
  Surface<R> s;
  s.p = add(o, scale(d, h.t));
  V3<R> outward;
  if (h.prim >= sc.num_prims) {
    const Plane<R>& pl = sc.planes[h.prim - sc.num_prims];
    outward = pl.normal;
    s.material = pl.material;
  } else {
    const Prim<R> p = loadPrim<R>(sc, h.prim);
    s.material = (int)metaMaterial(p.meta);
    if (metaKind(p.meta) == kKindTriangle) {
      outward = normalize(cross(p.b, p.c));
    } else {
      outward = vdiv(sub(s.p, p.a), p.b.x);
    }
  }
  s.front = dot(d, outward) < 0;
  s.n = s.front ? outward : neg(outward);
  return s;
}

}  // namespace ztd
