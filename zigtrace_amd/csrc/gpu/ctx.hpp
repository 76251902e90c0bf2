// Internal state behind the zt_gpu.h handles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <tuple>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../device/dev_common.hpp"
#include "zt_gpu.h"

// The ABI records keep the sizes the bindings are written against.
static_assert(sizeof(zt_gpu_tuning) == 52, "zt_gpu_tuning layout");
static_assert(sizeof(zt_gpu_stats) == 48, "zt_gpu_stats layout");
static_assert(sizeof(zt_render_options) == 64, "zt_render_options layout");
static_assert(sizeof(zt_ray) == 48 && sizeof(zt_hit) == 72, "zt_ray / zt_hit layout");

namespace ztg {

// Thrown inside the library, converted to a status code at the ABI edge.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define ZT_HIP(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw ::ztg::Error(ZT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

// A device allocation that frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    reset();
    if (n == 0) return;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(e == hipErrorOutOfMemory ? ZT_ERR_OUT_OF_MEMORY : ZT_ERR_HIP,
                  std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = n;
  }
  // Grow-only scratch.
  void ensure(size_t n) {
    if (n > bytes) alloc(n);
  }
  template <typename T>
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) ZT_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

// A bounded emissive shape that Lights.build would consider, with the area
// factor of its selection weight (weight = luminance(radiance) * weight_area).
struct LightCandidate : ztd::Light<double> {
  double weight_area;
};

// One uploaded scene: the device layout in both precisions plus what the
// host needs to rebuild per-render constants.
struct DevScene {
  DevBuf nodes, nodes4, nodes4q, primsF, primsD, planesF, planesD, matsF, matsD, lightsF, lightsD, cdf, prim_light;
  int num_nodes = 0, num_nodes4 = 0, num_prims = 0, num_planes = 0, num_lights = 0, num_mats = 0;
  int tree_depth = 0, tree_depth4 = 0;
  bool brute_force = false;  // no BVH handed over (or one too deep for the stack)
  double light_total = 0;
  zt_camera_config camera{};
  zt_background background{};
  // Host copies used to rebuild material/light tables when parameters change.
  std::vector<zt_material> materials;
  std::vector<ztd::Light<double>> lights_host;
  std::vector<LightCandidate> light_candidates;  // in shape order
  // Differentiable parameters: 3 per Lambertian / metal / emissive material.
  std::vector<int> param_offset;  // per material, -1 for dielectrics
  DevBuf param_offsets;
  int num_params = 0;
};

// Grow-only buffers of the wavefront renderer, reused across calls.
constexpr int kWfHostWords = 24;  // counter words copied to the host after each batch of bounces

struct WfBuffers {
  DevBuf scenes, slots, hits, hits_d, qext0, qext1, qsh, fext, fsh, ranges, counters, tally;
  // The bounce-loop graph and the bytes of every argument it was captured
  // with: reused while they match, captured again otherwise. Two
  // instantiations of it, one per batch in flight.
  std::vector<unsigned char> graph_key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec[2] = {};
  hipEvent_t done[2] = {};   // the batch in flight in slot i has finished
  uint32_t* host = nullptr;  // pinned: the counter words each batch ends with, [slot][kWfHostWords]
  void dropGraph() {
    for (auto& e : exec) {
      if (e) (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
    if (graph) (void)hipGraphDestroy(graph);
    graph = nullptr;
    graph_key.clear();
  }
  // The events and the pinned words, created on first use.
  void ensureHost() {
    if (host) return;
    for (auto& e : done) ZT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ZT_HIP(hipHostMalloc((void**)&host, 2 * kWfHostWords * sizeof(uint32_t), hipHostMallocDefault));
  }
  ~WfBuffers() {
    dropGraph();
    for (auto& e : done)
      if (e) (void)hipEventDestroy(e);
    if (host) (void)hipHostFree(host);
  }
};

// Autotuner record: the fastest configuration of one method for one class of
// render (image size and primitive count class, i.e. bit length of the count).
struct TuneKey {
  uint32_t method, precision, width, height, prim_class, brute;
  bool operator<(const TuneKey& o) const {
    return std::tie(method, precision, width, height, prim_class, brute) <
           std::tie(o.method, o.precision, o.width, o.height, o.prim_class, o.brute);
  }
};
struct TuneEntry {
  zt_gpu_tuning tu;
  double ms;
};

}  // namespace ztg

struct zt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::vector<std::unique_ptr<ztg::DevScene>> scenes;  // index = scene id; null = released
  zt_gpu_stats stats{};
  ztg::DevBuf out, errflag, probe_out;  // probe_out: the autotuner's trial frames
  ztg::DevBuf spill;  // traversal stack overflow, sized per launch (grow-only)
  ztg::DevBuf bands;  // per-XCD tile counters of the persistent megakernel
  ztg::DevBuf jobs;   // scene table of a many-scene wave-scheduled launch
  ztg::DevBuf tile_cost, tile_order;  // longest-first tile order of the wave-scheduled megakernel
  ztg::DevBuf adj_partial, adj_cols, adj_grad;  // backward: per-group partials, global lane columns, result
  ztg::WfBuffers wf;
  int* status_host = nullptr;  // pinned: where zt_gpu_backward_noalloc reads the status words
  std::map<ztg::TuneKey, ztg::TuneEntry> tuned;
};

namespace ztg {

DevScene& sceneOf(zt_ctx* ctx, uint32_t id);
// Resolves the device address of zt_gpu_backward_noalloc's scratch on
// `device` (the current device), once per device.
void initNoAllocScratch(int device);
void setDevice(zt_ctx* ctx);

// Camera.init (camera.zig:29-55) in f64 on the host.
ztd::CamD cameraInit(const zt_camera_config& cfg, double aspect);

template <typename R>
ztd::SceneView<R> viewOf(const DevScene& s);

// Runs `f`, converting exceptions to a status code and recording the message.
template <typename F>
int guarded(zt_ctx* ctx, F&& f) {
  try {
    f();
    if (ctx) ctx->err.clear();
    return ZT_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return ZT_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return ZT_ERR_INTERNAL;
  } catch (...) {
    if (ctx) ctx->err = "unknown exception";
    return ZT_ERR_INTERNAL;
  }
}

}  // namespace ztg
