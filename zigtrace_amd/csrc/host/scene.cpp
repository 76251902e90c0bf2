// Host scene API (scene.zig), Lights.build (light.zig:68-94) and the
// render.render entry (render.zig:152-180) that hands the built scene to the
// GPU through zt_gpu.h.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "host_scene.hpp"

namespace zth {

namespace {
double luminance(const double* c) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }  // light.zig:38-40
double triangleArea(const zt_shape& t) {                                                     // light.zig:42-44
  const double e1[3] = {t.v[3] - t.v[0], t.v[4] - t.v[1], t.v[5] - t.v[2]};
  const double e2[3] = {t.v[6] - t.v[0], t.v[7] - t.v[1], t.v[8] - t.v[2]};
  const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  return 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}
}  // namespace

void buildLights(const std::vector<zt_shape>& shapes, const std::vector<zt_material>& materials, HostLights* out) {
  out->items.clear();
  out->cdf.clear();
  for (size_t i = 0; i < shapes.size(); ++i) {
    const zt_shape& s = shapes[i];
    double a;
    if (s.kind == ZT_TRIANGLE) a = triangleArea(s);
    else if (s.kind == ZT_SPHERE) a = s.v[3] > 0 ? 4 * M_PI * s.v[3] * s.v[3] : 0;
    else continue;
    if (s.material >= materials.size()) continue;
    const zt_material& m = materials[s.material];
    if (m.kind != ZT_EMISSIVE) continue;
    const double w = luminance(m.color) * a;
    if (!(w > 0) || !std::isfinite(w)) continue;
    zt_light l{};
    l.shape = (uint32_t)i;
    std::memcpy(l.radiance, m.color, sizeof l.radiance);
    l.weight = w;
    out->items.push_back(l);
  }
  double sum = 0;
  for (const zt_light& l : out->items) {
    sum += l.weight;
    out->cdf.push_back(sum);
  }
  out->total = sum;
}

}  // namespace zth

namespace {

// Per-thread default GPU context (zt_render) and the device copies of scenes
// in it, keyed by the scene's process-unique id and version. Every thread's
// record is listed in a registry, so destroying a scene releases its copies
// in every context; a record's mutex is held while its thread renders and
// while another thread drops copies from it.
std::atomic<uint64_t> g_next_uid{1};

struct DeviceCopy {
  uint64_t uid, version;
  uint32_t id;
};
struct ThreadGpu;
struct Registry {
  std::mutex mu;
  std::vector<ThreadGpu*> all;
};
// Never destroyed: threads may exit while the process tears down statics.
Registry& registry() {
  static Registry* r = new Registry;
  return *r;
}
// The thread that loaded the library: its context is left to the end of the
// process (destroying it from a thread_local destructor at exit could run
// after the HIP runtime's own teardown). Other threads free theirs on exit.
const std::thread::id g_loader_thread = std::this_thread::get_id();

struct ThreadGpu {
  std::mutex mu;
  zt_ctx* ctx = nullptr;
  std::vector<DeviceCopy> copies;
  ThreadGpu() {
    std::lock_guard<std::mutex> lock(registry().mu);
    registry().all.push_back(this);
  }
  ~ThreadGpu() {
    {
      std::lock_guard<std::mutex> lock(registry().mu);
      auto& v = registry().all;
      v.erase(std::remove(v.begin(), v.end(), this), v.end());
    }
    if (ctx && std::this_thread::get_id() != g_loader_thread) zt_gpu_destroy(ctx);
  }
  // Drops the copies of scene `uid` (all versions). Caller holds `mu`.
  void drop(uint64_t uid) {
    for (size_t i = 0; i < copies.size();) {
      if (copies[i].uid == uid) {
        zt_gpu_scene_release(ctx, copies[i].id);
        copies.erase(copies.begin() + (long)i);
      } else {
        ++i;
      }
    }
  }
};
thread_local ThreadGpu g_gpu;

}  // namespace

extern "C" {

zt_scene* zt_scene_create(void) {
  zt_scene* s = new (std::nothrow) zt_scene();
  if (s) s->uid = g_next_uid.fetch_add(1);
  return s;
}

void zt_scene_destroy(zt_scene* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(registry().mu);
    for (ThreadGpu* t : registry().all) {
      std::lock_guard<std::mutex> tl(t->mu);
      t->drop(s->uid);
    }
  }
  delete s;
}

int zt_scene_add_material(zt_scene* s, const zt_material* m, uint32_t* id_out) {
  if (!s || !m) return ZT_ERR_INVALID_ARGUMENT;
  if (m->kind > ZT_EMISSIVE) return ZT_ERR_INVALID_ARGUMENT;
  s->materials.push_back(*m);
  ++s->version;
  if (id_out) *id_out = (uint32_t)(s->materials.size() - 1);
  return ZT_OK;
}

int zt_scene_add(zt_scene* s, const zt_shape* sh) {
  if (!s || !sh || sh->kind > ZT_PLANE) return ZT_ERR_INVALID_ARGUMENT;
  s->invalidate();
  if (sh->kind == ZT_PLANE) s->planes.push_back(*sh);
  else s->shapes.push_back(*sh);
  return ZT_OK;
}

int zt_scene_add_many(zt_scene* s, const zt_shape* sh, uint64_t n) {
  if (!s || (n && !sh)) return ZT_ERR_INVALID_ARGUMENT;
  for (uint64_t i = 0; i < n; ++i)
    if (sh[i].kind > ZT_PLANE) return ZT_ERR_INVALID_ARGUMENT;
  s->invalidate();
  for (uint64_t i = 0; i < n; ++i) {
    if (sh[i].kind == ZT_PLANE) s->planes.push_back(sh[i]);
    else s->shapes.push_back(sh[i]);
  }
  return ZT_OK;
}

int zt_scene_add_quad(zt_scene* s, const double c[3], const double u[3], const double v[3], uint32_t material) {
  if (!s || !c || !u || !v) return ZT_ERR_INVALID_ARGUMENT;
  zt_shape a{}, b{};
  a.kind = b.kind = ZT_TRIANGLE;
  a.material = b.material = material;
  for (int i = 0; i < 3; ++i) {
    const double p1 = c[i] + u[i];
    const double p2 = (c[i] + u[i]) + v[i];
    const double p3 = c[i] + v[i];
    a.v[i] = c[i];
    a.v[3 + i] = p1;
    a.v[6 + i] = p2;
    b.v[i] = c[i];
    b.v[3 + i] = p2;
    b.v[6 + i] = p3;
  }
  zt_scene_add(s, &a);
  return zt_scene_add(s, &b);
}

int zt_scene_set_camera(zt_scene* s, const zt_camera_config* c) {
  if (!s || !c) return ZT_ERR_INVALID_ARGUMENT;
  s->camera = *c;
  ++s->version;
  return ZT_OK;
}

int zt_scene_set_background(zt_scene* s, const zt_background* b) {
  if (!s || !b || b->kind > ZT_BG_SKY) return ZT_ERR_INVALID_ARGUMENT;
  s->background = *b;
  ++s->version;
  return ZT_OK;
}

int zt_scene_set_settings(zt_scene* s, uint64_t w, uint64_t h, uint64_t spp, uint64_t depth) {
  if (!s) return ZT_ERR_INVALID_ARGUMENT;
  s->width = w;
  s->height = h;
  s->spp = spp;
  s->max_depth = depth;
  return ZT_OK;
}

int zt_scene_build(zt_scene* s) {
  if (!s) return ZT_ERR_INVALID_ARGUMENT;
  s->invalidate();
  const int rc = zth::buildBvh(s->shapes.data(), s->shapes.size(), &s->nodes, &s->indices);
  if (rc != ZT_OK) return rc;
  zth::buildLights(s->shapes, s->materials, &s->lights);
  s->built = true;
  return ZT_OK;
}

int zt_scene_get_info(const zt_scene* s, zt_scene_info* o) {
  if (!s || !o) return ZT_ERR_INVALID_ARGUMENT;
  *o = zt_scene_info{};
  o->num_materials = s->materials.size();
  o->num_shapes = s->shapes.size();
  o->num_planes = s->planes.size();
  o->num_nodes = s->nodes.size();
  o->num_lights = s->lights.items.size();
  o->bvh_depth = zth::bvhDepth(s->nodes);
  o->width = s->width;
  o->height = s->height;
  o->spp = s->spp;
  o->max_depth = s->max_depth;
  o->built = s->built ? 1 : 0;
  return ZT_OK;
}

int zt_scene_get_view(const zt_scene* s, zt_scene_view* v) {
  if (!s || !v) return ZT_ERR_INVALID_ARGUMENT;
  *v = zt_scene_view{};
  v->materials = s->materials.data();
  v->num_materials = s->materials.size();
  v->shapes = s->shapes.data();
  v->num_shapes = s->shapes.size();
  v->planes = s->planes.data();
  v->num_planes = s->planes.size();
  // An unbuilt scene has no BVH and no lights (scene.zig:109-112, render.zig:62).
  v->nodes = s->built ? s->nodes.data() : nullptr;
  v->num_nodes = s->built ? s->nodes.size() : 0;
  v->indices = s->built ? s->indices.data() : nullptr;
  v->lights = s->lights.items.data();
  v->num_lights = s->lights.items.size();
  v->light_cdf = s->lights.cdf.data();
  v->light_total = s->lights.total;
  v->camera = s->camera;
  v->background = s->background;
  return ZT_OK;
}

int zt_bvh_build(const zt_shape* shapes, uint64_t n, zt_bvh_node* nodes_out, uint64_t* num_nodes,
                 uint32_t* indices_out) {
  return zt_bvh_build_threads(shapes, n, 0, nodes_out, num_nodes, indices_out);
}

int zt_bvh_build_threads(const zt_shape* shapes, uint64_t n, uint32_t threads, zt_bvh_node* nodes_out,
                         uint64_t* num_nodes, uint32_t* indices_out) {
  if ((n && !shapes) || !num_nodes) return ZT_ERR_INVALID_ARGUMENT;
  std::vector<zt_bvh_node> nodes;
  std::vector<uint32_t> idx;
  const int rc = zth::buildBvh(shapes, n, &nodes, &idx, threads);
  if (rc != ZT_OK) return rc;
  *num_nodes = nodes.size();
  if (nodes_out && !nodes.empty()) std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(zt_bvh_node));
  if (indices_out && !idx.empty()) std::memcpy(indices_out, idx.data(), idx.size() * 4);
  return ZT_OK;
}

int zt_default_context(zt_ctx** out) {
  if (!out) return ZT_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_gpu.mu);
  if (!g_gpu.ctx) {
    const int rc = zt_gpu_create(0, &g_gpu.ctx);
    if (rc != ZT_OK) {
      g_gpu.ctx = nullptr;
      return rc;
    }
  }
  *out = g_gpu.ctx;
  return ZT_OK;
}

}  // extern "C"

namespace {
// zt_scene_device_id for the calling thread's default context (caching the
// copy); caller holds g_gpu.mu.
int deviceIdLocked(zt_scene* s, zt_ctx* ctx, uint32_t* id_out) {
  if (ctx == g_gpu.ctx) {
    for (size_t i = 0; i < g_gpu.copies.size(); ++i) {
      DeviceCopy& c = g_gpu.copies[i];
      if (c.uid != s->uid) continue;
      if (c.version == s->version) {
        *id_out = c.id;
        return ZT_OK;
      }
      zt_gpu_scene_release(ctx, c.id);
      g_gpu.copies.erase(g_gpu.copies.begin() + (long)i);
      break;
    }
  }
  zt_scene_view v;
  zt_scene_get_view(s, &v);
  uint32_t id = 0;
  const int rc = zt_gpu_scene_upload(ctx, &v, &id);
  if (rc != ZT_OK) return rc;
  if (ctx == g_gpu.ctx) g_gpu.copies.push_back(DeviceCopy{s->uid, s->version, id});
  *id_out = id;
  return ZT_OK;
}
}  // namespace

extern "C" {

int zt_scene_device_id(zt_scene* s, zt_ctx* ctx, uint32_t* id_out) {
  if (!s || !ctx || !id_out) return ZT_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_gpu.mu);
  return deviceIdLocked(s, ctx, id_out);
}

int zt_render(zt_scene* s, const zt_render_options* o, const zt_gpu_tuning* tuning, double* pixels) {
  if (!o) return ZT_ERR_INVALID_ARGUMENT;
  // render.zig:153: option validation comes before anything else, even for
  // an unbuilt scene.
  if (o->width == 0 || o->height == 0 || o->spp == 0 || o->tile_size == 0) return ZT_ERR_INVALID_OPTIONS;
  if (!s || !pixels) return ZT_ERR_INVALID_ARGUMENT;
  zt_ctx* ctx = nullptr;
  int rc = zt_default_context(&ctx);
  if (rc != ZT_OK) return rc;
  // Held for the whole render: another thread destroying a scene waits
  // before it releases this context's copies.
  std::lock_guard<std::mutex> lock(g_gpu.mu);
  uint32_t id = 0;
  rc = deviceIdLocked(s, ctx, &id);
  if (rc != ZT_OK) return rc;
  const bool f64 = tuning && tuning->precision == ZT_PRECISION_FP64;
  const size_t n = (size_t)o->width * o->height * 3;
  if (f64) return zt_gpu_render(ctx, id, o, tuning, pixels, 0);
  std::vector<float> tmp;
  try {
    tmp.resize(n);
  } catch (const std::bad_alloc&) {
    return ZT_ERR_OUT_OF_MEMORY;
  }
  rc = zt_gpu_render(ctx, id, o, tuning, tmp.data(), 0);
  if (rc != ZT_OK) return rc;
  for (size_t i = 0; i < n; ++i) pixels[i] = (double)tmp[i];  // widen to the reference's []Vec3 (f64)
  return ZT_OK;
}

}  // extern "C"
