// Context lifetime and the host -> device scene re-layout.
//
// The host hands over the reference's data structures (f64 shapes, the
// 56-byte-node BVH with its index permutation, the light table). The device
// layout built here:
//   * primitives reordered into leaf order (slot k = shapes[indices[k]]), so a
//     leaf reads one contiguous run; triangles stored as v0, e1, e2;
//   * one 64-byte node per interior node of the host BVH holding both
//     children's boxes (rounded outward to fp32) and their codes;
//   * materials, planes and lights in fp32 (product path) and f64
//     (verification path), lights with their f64-precomputed normals, areas
//     and selection probabilities, and a primitive -> light map for MIS.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "../device/dev_path.hpp"
#include "ctx.hpp"

using namespace ztd;

namespace ztg {

void setDevice(zt_ctx* ctx) { ZT_HIP(hipSetDevice(ctx->device)); }

DevScene& sceneOf(zt_ctx* ctx, uint32_t id) {
  if (!ctx) throw Error(ZT_ERR_INVALID_ARGUMENT, "null context");
  if (id >= ctx->scenes.size() || !ctx->scenes[id]) throw Error(ZT_ERR_INVALID_ARGUMENT, "unknown scene id");
  return *ctx->scenes[id];
}

CamD cameraInit(const zt_camera_config& cfg, double aspect) {  // camera.zig:29-55
  const V3<double> from = mk(cfg.look_from[0], cfg.look_from[1], cfg.look_from[2]);
  const V3<double> at = mk(cfg.look_at[0], cfg.look_at[1], cfg.look_at[2]);
  const V3<double> up = mk(cfg.up[0], cfg.up[1], cfg.up[2]);
  const double theta = cfg.vfov * 0.017453292519943295;  // degreesToRadians: vfov * f64(pi/180)
  const double half_h = std::tan(theta / 2);
  const double viewport_h = 2.0 * half_h;
  const double viewport_w = aspect * viewport_h;
  const V3<double> w = normalize(sub(from, at));
  const V3<double> u = normalize(cross(up, w));
  const V3<double> v = cross(w, u);
  const double focus = cfg.has_focus_dist ? cfg.focus_dist : length(sub(from, at));
  const V3<double> horizontal = scale(u, viewport_w * focus);
  const V3<double> vertical = scale(v, -viewport_h * focus);
  const V3<double> ul = sub(sub(sub(from, scale(w, focus)), scale(horizontal, 0.5)), scale(vertical, 0.5));
  CamD c;
  const V3<double>* src[6] = {&from, &ul, &horizontal, &vertical, &u, &v};
  double* dst[6] = {c.origin, c.upper_left, c.horizontal, c.vertical, c.u, c.v};
  for (int i = 0; i < 6; ++i) {
    dst[i][0] = src[i]->x;
    dst[i][1] = src[i]->y;
    dst[i][2] = src[i]->z;
  }
  c.lens_radius = cfg.aperture / 2;
  return c;
}

template <typename R>
SceneView<R> viewOf(const DevScene& s) {
  SceneView<R> v{};
  v.nodes = s.brute_force ? nullptr : s.nodes.as<DNode>();
  v.num_nodes = s.brute_force ? 0 : s.num_nodes;
  v.nodes4 = (std::is_same<R, float>::value && !s.brute_force && s.num_nodes4) ? s.nodes4.as<DNode4>() : nullptr;
  v.nodes4q = nullptr;  // 64-byte nodes only when a tuning asks for them
  if constexpr (std::is_same<R, float>::value) {
    v.prims = s.primsF.p;
    v.planes = s.planesF.as<Plane<float>>();
    v.mats = s.matsF.as<Mat<float>>();
    v.lights = s.lightsF.as<Light<float>>();
  } else {
    v.prims = s.primsD.p;
    v.planes = s.planesD.as<Plane<double>>();
    v.mats = s.matsD.as<Mat<double>>();
    v.lights = s.lightsD.as<Light<double>>();
  }
  v.cdf = s.cdf.as<double>();
  v.prim_light = s.prim_light.as<int>();
  v.num_prims = s.num_prims;
  v.num_planes = s.num_planes;
  v.num_lights = s.num_lights;
  v.total = s.light_total;
  v.bg_kind = (int)s.background.kind;
  v.bg_a = mk((R)s.background.a[0], (R)s.background.a[1], (R)s.background.a[2]);
  v.bg_b = mk((R)s.background.b[0], (R)s.background.b[1], (R)s.background.b[2]);
  return v;
}
template SceneView<float> viewOf<float>(const DevScene&);
template SceneView<double> viewOf<double>(const DevScene&);

namespace {

float roundDown(double v) {
  float f = (float)v;
  if ((double)f > v) f = std::nextafter(f, -INFINITY);
  return f;
}
float roundUp(double v) {
  float f = (float)v;
  if ((double)f < v) f = std::nextafter(f, INFINITY);
  return f;
}

struct Box6 {
  float lo[3], hi[3];
};
Box6 boxOf(const zt_bvh_node& n) {
  Box6 b;
  for (int i = 0; i < 3; ++i) {
    b.lo[i] = roundDown(n.min[i]);
    b.hi[i] = roundUp(n.max[i]);
  }
  return b;
}
Box6 emptyBox() {
  Box6 b;
  for (int i = 0; i < 3; ++i) {
    b.lo[i] = 1e30f;
    b.hi[i] = -1e30f;
  }
  return b;
}

// Converts the host BVH into 64-byte two-child nodes in depth-first preorder.
struct TreeBuilder {
  const zt_bvh_node* ref;
  uint64_t num_ref;
  std::vector<DNode> out;
  std::vector<char> visited;
  int max_depth = 0;

  void setChild(size_t idx, int side, int code, const Box6& b) {
    DNode& n = out[idx];
    if (side == 0) {
      n.a = make_float4(b.lo[0], b.lo[1], b.lo[2], b.hi[0]);
      n.b.x = b.hi[1];
      n.b.y = b.hi[2];
      n.d.x = code;
    } else {
      n.b.z = b.lo[0];
      n.b.w = b.lo[1];
      n.c = make_float4(b.lo[2], b.hi[0], b.hi[1], b.hi[2]);
      n.d.y = code;
    }
  }
  // A leaf range [first, first+count). Ranges longer than a code can hold
  // (only possible for depth-capped leaves) become synthetic interior nodes
  // whose children share the leaf's box.
  int leaf(uint32_t first, uint32_t count, const Box6& b, int depth) {
    max_depth = std::max(max_depth, depth);
    if (first > kMaxLeafFirst) throw Error(ZT_ERR_BAD_SCENE, "too many primitives for the device layout");
    if (count <= kMaxLeafCount) return leafCode(first, count);
    const size_t idx = out.size();
    out.push_back(DNode{});
    const uint32_t half = count / 2;
    const int l = leaf(first, half, b, depth + 1);
    const int r = leaf(first + half, count - half, b, depth + 1);
    setChild(idx, 0, l, b);
    setChild(idx, 1, r, b);
    return (int)idx;
  }
  int node(uint64_t i, int depth) {
    if (visited[i]) throw Error(ZT_ERR_BAD_SCENE, "bvh node referenced twice");
    visited[i] = 1;
    if (depth > 1000) throw Error(ZT_ERR_BAD_SCENE, "bvh too deep");
    const zt_bvh_node& n = ref[i];
    if (n.count > 0) return leaf(n.first, n.count, boxOf(n), depth);
    max_depth = std::max(max_depth, depth);
    const size_t idx = out.size();
    out.push_back(DNode{});
    const uint64_t l = n.first;
    const int lc = node(l, depth + 1);
    const int rc = node(l + 1, depth + 1);
    setChild(idx, 0, lc, boxOf(ref[l]));
    setChild(idx, 1, rc, boxOf(ref[l + 1]));
    return (int)idx;
  }
  void build() {
    visited.assign(num_ref, 0);
    if (ref[0].count > 0) {  // the whole scene is one leaf
      out.push_back(DNode{});
      const int lc = leaf(ref[0].first, ref[0].count, boxOf(ref[0]), 1);
      setChild(0, 0, lc, boxOf(ref[0]));
      setChild(0, 1, kNoChild, emptyBox());
    } else {
      node(0, 0);
    }
  }
};

// Collapses the two-child device tree into four-wide nodes (DNode4): each
// four-wide node takes a two-child node's children and, while it has fewer
// than four, replaces its interior child of largest surface area by that
// child's two children. Nodes are emitted in depth-first preorder; boxes are
// copied bit for bit from the two-child nodes.
struct Collapse4 {
  const std::vector<DNode>& in;
  std::vector<DNode4> out;
  int max_depth = 0;

  struct Child {
    int code;
    Box6 box;
  };
  static Box6 leftBox(const DNode& n) {
    return Box6{{n.a.x, n.a.y, n.a.z}, {n.a.w, n.b.x, n.b.y}};
  }
  static Box6 rightBox(const DNode& n) {
    return Box6{{n.b.z, n.b.w, n.c.x}, {n.c.y, n.c.z, n.c.w}};
  }
  static double area(const Box6& b) {
    const double dx = (double)b.hi[0] - b.lo[0], dy = (double)b.hi[1] - b.lo[1], dz = (double)b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
  }
  int node(int i, int depth) {
    max_depth = std::max(max_depth, depth);
    std::vector<Child> ch;
    const DNode& n = in[i];
    if (n.d.x != kNoChild) ch.push_back({n.d.x, leftBox(n)});
    if (n.d.y != kNoChild) ch.push_back({n.d.y, rightBox(n)});
    while (ch.size() < 4) {
      int pick = -1;
      double best = -1;
      for (size_t k = 0; k < ch.size(); ++k)
        if (ch[k].code >= 0 && area(ch[k].box) > best) {
          best = area(ch[k].box);
          pick = (int)k;
        }
      if (pick < 0) break;
      const DNode& c = in[ch[pick].code];
      std::vector<Child> repl;
      if (c.d.x != kNoChild) repl.push_back({c.d.x, leftBox(c)});
      if (c.d.y != kNoChild) repl.push_back({c.d.y, rightBox(c)});
      ch.erase(ch.begin() + pick);
      ch.insert(ch.begin() + pick, repl.begin(), repl.end());
    }
    const int idx = (int)out.size();
    out.push_back(DNode4{});
    float lo[3][4], hi[3][4];
    int code[4];
    for (int k = 0; k < 4; ++k) {
      const Box6 b = k < (int)ch.size() ? ch[k].box : emptyBox();
      for (int a = 0; a < 3; ++a) {
        lo[a][k] = b.lo[a];
        hi[a][k] = b.hi[a];
      }
      code[k] = k < (int)ch.size() ? ch[k].code : kNoChild;
    }
    for (int k = 0; k < 4; ++k)
      if (code[k] >= 0) code[k] = node(code[k], depth + 1);
      else if (code[k] != kNoChild) max_depth = std::max(max_depth, depth + 1);
    DNode4& o = out[idx];
    o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    o.child = make_int4(code[0], code[1], code[2], code[3]);
    o._pad = make_int4(0, 0, 0, 0);
    return idx;
  }
};

// Compresses four-wide nodes to 64 bytes (DNode4Q). Per node and axis the
// origin is the smallest child lower bound and the step s = 2^e the smallest
// power of two that lets every child bound be reached within 255 steps;
// each lower bound q is the largest with fma(q, s, origin) <= lo and each
// upper bound the smallest with fma(q, s, origin) >= hi, evaluated with the
// same single-rounding fma the device decodes with, so a decoded box always
// contains the fp32 box (and the traversal's hits are those of the fp32
// boxes, up to exact ties in t).
struct Quantize4 {
  static bool axis(const float* lo, const float* hi, const bool* valid, float* origin, int* e_out, uint8_t* qlo,
                   uint8_t* qhi) {
    float o = INFINITY, m = -INFINITY;
    for (int k = 0; k < 4; ++k)
      if (valid[k]) {
        o = std::min(o, lo[k]);
        m = std::max(m, hi[k]);
      }
    if (!(o <= m)) {  // no valid child: any inverted box
      *origin = 0;
      *e_out = 0;
      for (int k = 0; k < 4; ++k) qlo[k] = 255, qhi[k] = 0;
      return true;
    }
    const double ext = (double)m - (double)o;
    int e = ext > 0 ? (int)std::ceil(std::log2(ext / 255.0)) - 1 : -126;
    e = std::max(-126, std::min(127, e));
    for (; e <= 127; ++e) {
      const float st = std::ldexp(1.0f, e);
      bool ok = true;
      for (int k = 0; k < 4 && ok; ++k) {
        if (!valid[k]) {
          qlo[k] = 255;
          qhi[k] = 0;
          continue;
        }
        double fl = std::floor(((double)lo[k] - (double)o) / st);
        int ql = (int)std::max(0.0, std::min(255.0, fl));
        while (ql > 0 && std::fmaf((float)ql, st, o) > lo[k]) --ql;
        while (ql < 255 && std::fmaf((float)(ql + 1), st, o) <= lo[k]) ++ql;
        if (std::fmaf((float)ql, st, o) > lo[k]) ok = false;
        double ch = std::ceil(((double)hi[k] - (double)o) / st);
        int qh = (int)std::max(0.0, std::min(256.0, ch));
        while (qh <= 255 && std::fmaf((float)qh, st, o) < hi[k]) ++qh;
        while (qh > 0 && qh <= 255 && std::fmaf((float)(qh - 1), st, o) >= hi[k]) --qh;
        if (qh > 255) ok = false;
        qlo[k] = (uint8_t)ql;
        qhi[k] = (uint8_t)std::min(qh, 255);
      }
      if (ok) {
        *origin = o;
        *e_out = e;
        return true;
      }
    }
    return false;
  }
  static bool node(const DNode4& n, DNode4Q* out) {
    const int code[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
    bool valid[4];
    for (int k = 0; k < 4; ++k) valid[k] = code[k] != kNoChild;
    const float lo[3][4] = {{n.lox.x, n.lox.y, n.lox.z, n.lox.w}, {n.loy.x, n.loy.y, n.loy.z, n.loy.w},
                            {n.loz.x, n.loz.y, n.loz.z, n.loz.w}};
    const float hi[3][4] = {{n.hix.x, n.hix.y, n.hix.z, n.hix.w}, {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w},
                            {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w}};
    float origin[3];
    int e[3];
    uint8_t ql[3][4], qh[3][4];
    for (int a = 0; a < 3; ++a)
      if (!axis(lo[a], hi[a], valid, &origin[a], &e[a], ql[a], qh[a])) return false;
    auto pack = [](const uint8_t* q) {
      return (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
    };
    *out = DNode4Q{};
    out->ox = origin[0];
    out->oy = origin[1];
    out->oz = origin[2];
    out->exps = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16);
    out->child = n.child;
    out->qlox = pack(ql[0]);
    out->qhix = pack(qh[0]);
    out->qloy = pack(ql[1]);
    out->qhiy = pack(qh[1]);
    out->qloz = pack(ql[2]);
    out->qhiz = pack(qh[2]);
    return true;
  }
};

void validate(const zt_scene_view* v) {
  auto bad = [](const std::string& m) { throw Error(ZT_ERR_BAD_SCENE, m); };
  if (!v) throw Error(ZT_ERR_INVALID_ARGUMENT, "null scene view");
  if (v->num_materials && !v->materials) bad("materials pointer is null");
  if (v->num_shapes && !v->shapes) bad("shapes pointer is null");
  if (v->num_planes && !v->planes) bad("planes pointer is null");
  if (v->num_shapes > (uint64_t)kMaxLeafFirst || v->num_planes > (1u << 24) || v->num_materials > 0x3FFFFFFFull)
    bad("scene too large");
  for (uint64_t i = 0; i < v->num_materials; ++i)
    if (v->materials[i].kind > ZT_EMISSIVE) bad("material kind out of range");
  for (uint64_t i = 0; i < v->num_shapes; ++i) {
    if (v->shapes[i].kind != ZT_SPHERE && v->shapes[i].kind != ZT_TRIANGLE) bad("bounded shape list holds a plane");
    if (v->shapes[i].material >= v->num_materials) bad("shape material out of range");
  }
  for (uint64_t i = 0; i < v->num_planes; ++i) {
    if (v->planes[i].kind != ZT_PLANE) bad("plane list holds a bounded shape");
    if (v->planes[i].material >= v->num_materials) bad("plane material out of range");
  }
  if (v->num_nodes) {
    if (!v->nodes || !v->indices) bad("bvh pointers are null");
    if (v->num_shapes == 0 || v->num_nodes > 2 * v->num_shapes - 1) bad("bvh node count inconsistent");
    for (uint64_t i = 0; i < v->num_nodes; ++i) {
      const zt_bvh_node& n = v->nodes[i];
      if (n.count > 0) {
        if ((uint64_t)n.first + n.count > v->num_shapes) bad("bvh leaf range out of bounds");
      } else if ((uint64_t)n.first <= i || (uint64_t)n.first + 1 >= v->num_nodes) {
        bad("bvh interior node children out of order or out of bounds");
      }
    }
    std::vector<char> seen(v->num_shapes, 0);
    for (uint64_t i = 0; i < v->num_shapes; ++i) {
      const uint32_t k = v->indices[i];
      if (k >= v->num_shapes || seen[k]) bad("bvh indices are not a permutation");
      seen[k] = 1;
    }
  }
  if (v->num_lights) {
    if (!v->lights || !v->light_cdf) bad("light pointers are null");
    for (uint64_t i = 0; i < v->num_lights; ++i)
      if (v->lights[i].shape >= v->num_shapes) bad("light shape out of range");
  }
  if (v->background.kind > ZT_BG_SKY) bad("background kind out of range");
}

V3<double> vec(const double* a) { return mk(a[0], a[1], a[2]); }

template <typename R>
Light<R> lightCast(const Light<double>& L) {
  Light<R> o;
  o.p0 = cvt<R>(L.p0);
  o.p1 = cvt<R>(L.p1);
  o.p2 = cvt<R>(L.p2);
  o.radiance = cvt<R>(L.radiance);
  o.normal = cvt<R>(L.normal);
  o.radius = (R)L.radius;
  o.area = (R)L.area;
  o.select = (R)L.select;
  o.kind = L.kind;
  o.prim = L.prim;
  o.material = L.material;
  o._pad = 0;
  return o;
}

}  // namespace

// Rebuilds the fp32/f64 material and light tables from host copies (used at
// upload and whenever parameters override material colours).
void uploadMaterialsAndLights(DevScene& s, hipStream_t stream) {
  std::vector<Mat<float>> mf(s.materials.size());
  std::vector<Mat<double>> md(s.materials.size());
  for (size_t i = 0; i < s.materials.size(); ++i) {
    const zt_material& m = s.materials[i];
    md[i] = Mat<double>{vec(m.color), m.param, (int)m.kind, 0};
    mf[i] = Mat<float>{cvt<float>(vec(m.color)), (float)m.param, (int)m.kind, 0};
  }
  s.matsF.upload(mf, stream);
  s.matsD.upload(md, stream);
  std::vector<Light<float>> lf;
  for (const auto& L : s.lights_host) lf.push_back(lightCast<float>(L));
  s.lightsF.upload(lf, stream);
  s.lightsD.upload(s.lights_host, stream);
}

}  // namespace ztg

using namespace ztg;

extern "C" {

int zt_gpu_create(int device, zt_ctx** out) {
  if (!out) return ZT_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto ctx = std::make_unique<zt_ctx>();
  ctx->device = device;
  const int rc = guarded(nullptr, [&] {
    int n = 0;
    ZT_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(ZT_ERR_INVALID_ARGUMENT, "no such device");
    ZT_HIP(hipSetDevice(device));
    ZT_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ZT_HIP(hipEventCreate(&ctx->ev0));
    ZT_HIP(hipEventCreate(&ctx->ev1));
    ctx->errflag.alloc(ztd::kStatusBytes);
    ZT_HIP(hipHostMalloc((void**)&ctx->status_host, ztd::kStatusBytes, hipHostMallocDefault));
    initNoAllocScratch(device);
  });
  if (rc != ZT_OK) {
    if (ctx->status_host) (void)hipHostFree(ctx->status_host);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    return rc;
  }
  *out = ctx.release();
  return ZT_OK;
}

void zt_gpu_destroy(zt_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->scenes.clear();
  ctx->out.reset();
  ctx->errflag.reset();
  ctx->probe_out.reset();
  if (ctx->status_host) (void)hipHostFree(ctx->status_host);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* zt_gpu_last_error(const zt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* zt_gpu_stream(zt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int zt_gpu_get_stats(const zt_ctx* ctx, zt_gpu_stats* out) {
  if (!ctx || !out) return ZT_ERR_INVALID_ARGUMENT;
  *out = ctx->stats;
  return ZT_OK;
}

int zt_gpu_device_alloc(zt_ctx* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    setDevice(ctx);
    *out = nullptr;
    const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
      *out = nullptr;
      throw Error(e == hipErrorOutOfMemory ? ZT_ERR_OUT_OF_MEMORY : ZT_ERR_HIP, hipGetErrorString(e));
    }
  });
}

int zt_gpu_device_free(zt_ctx* ctx, void* ptr) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    setDevice(ctx);
    ZT_HIP(hipStreamSynchronize(ctx->stream));
    if (ptr) ZT_HIP(hipFree(ptr));
  });
}

int zt_gpu_copy_to_host(zt_ctx* ctx, void* host, const void* device, uint64_t bytes) {
  if (!ctx || (bytes && (!host || !device))) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    setDevice(ctx);
    ZT_HIP(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, ctx->stream));
    ZT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int zt_gpu_copy_to_device(zt_ctx* ctx, void* device, const void* host, uint64_t bytes) {
  if (!ctx || (bytes && (!host || !device))) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    setDevice(ctx);
    ZT_HIP(hipMemcpyAsync(device, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    ZT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int zt_gpu_scene_upload(zt_ctx* ctx, const zt_scene_view* v, uint32_t* scene_id) {
  if (!ctx || !scene_id) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    const auto t0 = std::chrono::steady_clock::now();
    validate(v);
    setDevice(ctx);
    auto s = std::make_unique<DevScene>();
    const uint64_t n = v->num_shapes;
    s->num_prims = (int)n;
    s->num_planes = (int)v->num_planes;
    s->num_mats = (int)v->num_materials;
    s->camera = v->camera;
    s->background = v->background;
    s->materials.assign(v->materials, v->materials + v->num_materials);

    // Slot order: leaf order when a BVH is given, shape order otherwise.
    std::vector<uint32_t> order(n), slot_of(n);
    for (uint64_t k = 0; k < n; ++k) order[k] = v->num_nodes ? v->indices[k] : (uint32_t)k;
    for (uint64_t k = 0; k < n; ++k) slot_of[order[k]] = (uint32_t)k;

    if (v->num_nodes) {
      TreeBuilder tb{v->nodes, v->num_nodes, {}, {}, 0};
      tb.build();
      s->tree_depth = tb.max_depth;
      if (tb.max_depth + 1 > kMaxTreeDepth) {
        s->brute_force = true;  // deeper than the traversal stack: exact but slow fallback
      } else {
        s->num_nodes = (int)tb.out.size();
        s->nodes.upload(tb.out, ctx->stream);
        // The four-wide traversal pushes up to three entries per level.
        Collapse4 c4{tb.out, {}, 0};
        c4.node(0, 0);
        s->tree_depth4 = c4.max_depth;
        if (3 * c4.max_depth + 1 <= 8 + kSpill && c4.out.size() <= kMaxNodes4) {
          s->num_nodes4 = (int)c4.out.size();
          s->nodes4.upload(c4.out, ctx->stream);
          std::vector<DNode4Q> q(c4.out.size());
          bool ok = true;
          for (size_t i = 0; i < q.size() && ok; ++i) ok = Quantize4::node(c4.out[i], &q[i]);
          if (ok) s->nodes4q.upload(q, ctx->stream);  // else only the 128-byte form
        }
      }
    } else {
      s->brute_force = n > 0;
    }

    std::vector<PrimF> pf(n);
    std::vector<PrimD> pd(n);
    for (uint64_t k = 0; k < n; ++k) {
      const zt_shape& sh = v->shapes[order[k]];
      const uint32_t meta = ((sh.kind == ZT_TRIANGLE ? kKindTriangle : kKindSphere) << 30) | sh.material;
      PrimD d{};
      d.meta = meta;
      if (sh.kind == ZT_TRIANGLE) {
        const V3<double> v0 = vec(sh.v), v1 = vec(sh.v + 3), v2 = vec(sh.v + 6);
        const V3<double> e1 = sub(v1, v0), e2 = sub(v2, v0);
        const double tmp[9] = {v0.x, v0.y, v0.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z};
        std::memcpy(d.a, tmp, sizeof tmp);
      } else {
        d.a[0] = sh.v[0];
        d.a[1] = sh.v[1];
        d.a[2] = sh.v[2];
        d.a[3] = sh.v[3];
      }
      pd[k] = d;
      PrimF f{};
      float meta_f;
      std::memcpy(&meta_f, &meta, 4);
      f.p0 = make_float4((float)d.a[0], (float)d.a[1], (float)d.a[2], meta_f);
      f.p1 = make_float4((float)d.a[3], (float)d.a[4], (float)d.a[5], 0.f);
      f.p2 = make_float4((float)d.a[6], (float)d.a[7], (float)d.a[8], meta_f);
      pf[k] = f;
    }
    s->primsF.upload(pf, ctx->stream);
    s->primsD.upload(pd, ctx->stream);

    std::vector<Plane<float>> plf(v->num_planes);
    std::vector<Plane<double>> pld(v->num_planes);
    for (uint64_t j = 0; j < v->num_planes; ++j) {
      const zt_shape& sh = v->planes[j];
      pld[j] = Plane<double>{vec(sh.v), vec(sh.v + 3), (int)sh.material, 0};
      plf[j] = Plane<float>{cvt<float>(pld[j].point), cvt<float>(pld[j].normal), (int)sh.material, 0};
    }
    s->planesF.upload(plf, ctx->stream);
    s->planesD.upload(pld, ctx->stream);

    // Lights (light.zig:68-94 built on the host): precompute per light what
    // sample/pdfHit recompute in the reference, with the same f64 operations.
    std::vector<int> prim_light(std::max<uint64_t>(n, 1), -1);
    s->num_lights = (int)v->num_lights;
    s->light_total = v->light_total;
    for (uint64_t i = 0; i < v->num_lights; ++i) {
      const zt_light& l = v->lights[i];
      const zt_shape& sh = v->shapes[l.shape];
      Light<double> L{};
      L.radiance = vec(l.radiance);
      L.select = l.weight / v->light_total;
      L.prim = (int)slot_of[l.shape];
      if (sh.kind == ZT_TRIANGLE) {
        L.kind = (int)kKindTriangle;
        L.p0 = vec(sh.v);
        L.p1 = vec(sh.v + 3);
        L.p2 = vec(sh.v + 6);
        const V3<double> c = cross(sub(L.p1, L.p0), sub(L.p2, L.p0));
        L.normal = normalize(c);
        L.area = 0.5 * length(c);
      } else {
        L.kind = (int)kKindSphere;
        L.p0 = vec(sh.v);
        L.radius = sh.v[3];
      }
      L.material = (int)sh.material;
      s->lights_host.push_back(L);
      prim_light[L.prim] = (int)i;
    }
    // Every bounded emissive shape, for rebuilding the light table when
    // material colours change (light.zig:68-94).
    for (uint64_t i = 0; i < n; ++i) {
      const zt_shape& sh = v->shapes[i];
      if (v->materials[sh.material].kind != ZT_EMISSIVE) continue;
      LightCandidate c{};
      c.prim = (int)slot_of[i];
      c.material = (int)sh.material;
      if (sh.kind == ZT_TRIANGLE) {
        c.kind = (int)kKindTriangle;
        c.p0 = vec(sh.v);
        c.p1 = vec(sh.v + 3);
        c.p2 = vec(sh.v + 6);
        const V3<double> x = cross(sub(c.p1, c.p0), sub(c.p2, c.p0));
        c.normal = normalize(x);
        c.area = 0.5 * length(x);
        c.weight_area = c.area;
      } else {
        c.kind = (int)kKindSphere;
        c.p0 = vec(sh.v);
        c.radius = sh.v[3];
        c.weight_area = c.radius > 0 ? 4 * M_PI * c.radius * c.radius : 0;
      }
      s->light_candidates.push_back(c);
    }
    s->param_offset.assign(v->num_materials, -1);
    for (uint64_t m = 0; m < v->num_materials; ++m)
      if (v->materials[m].kind != ZT_DIELECTRIC) {
        s->param_offset[m] = s->num_params;
        s->num_params += 3;
      }
    std::vector<int> po(s->param_offset);
    if (po.empty()) po.push_back(-1);
    s->param_offsets.upload(po, ctx->stream);
    std::vector<double> cdf(v->light_cdf, v->light_cdf + v->num_lights);
    s->cdf.upload(cdf, ctx->stream);
    s->prim_light.upload(prim_light, ctx->stream);
    uploadMaterialsAndLights(*s, ctx->stream);
    ZT_HIP(hipStreamSynchronize(ctx->stream));

    uint32_t id = 0;
    while (id < ctx->scenes.size() && ctx->scenes[id]) ++id;
    if (id == ctx->scenes.size()) ctx->scenes.emplace_back();
    ctx->scenes[id] = std::move(s);
    *scene_id = id;
    ctx->stats.upload_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  });
}

int zt_gpu_scene_release(zt_ctx* ctx, uint32_t scene_id) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    sceneOf(ctx, scene_id);
    setDevice(ctx);
    ZT_HIP(hipStreamSynchronize(ctx->stream));
    ctx->scenes[scene_id].reset();
  });
}

}  // extern "C"
