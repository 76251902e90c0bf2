// Differentiable rendering: scene parameters, forward, and the path-replay
// adjoint.
//
// Parameters are the colours the estimator is linear in per channel: the
// albedo of every Lambertian and metal material and the radiance of every
// emissive material (3 floats each, in material order). Geometry, visibility
// and the light-selection probabilities are not differentiated.
//
// Backward re-traces every path from its pixel's RNG stream (nothing is kept
// from the forward call; memory is O(pixels) for dL/dimage only) and
// accumulates, per path, the contributions it adds to its radiance. With the
// Russian-roulette probability held fixed, every contribution X of a path is
// a product of the albedos of the vertices before it, so
//     d(delta . X)/d albedo_m[c] = n_m * delta[c] * X[c] / albedo_m[c],
// n_m being how many of those vertices use material m; emitter radiance
// enters X linearly. Holding the roulette probability fixed is what keeps the
// gradient unbiased: the estimator T/p * [u < p] has expectation T for any p,
// so its derivative is taken with p constant (differentiating through
// p = max(T) would give 0 for a single-channel path whose true derivative is 1).
//
// Each lane keeps a short list of the albedo materials its current path has
// visited (with counts) and of the emitters it has drawn from; at the end of
// a finite sample the list is divided through and added to the lane's own
// column of partial gradients (one value per parameter), with plain loads and
// stores: no atomics, so the bits do not depend on scheduling. Pixels are
// reduced in fixed 8x8 groups of the image: once a group's 64 lanes are done,
// lane l sums parameter l (l + 64, ...) over the group's columns in lane
// order, in double, into the group's row of partials; a second kernel sums
// the groups' rows in a fixed order. The gradient is therefore the same bits
// for every call and every launch shape (tile shape, grid size).

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../device/dev_path.hpp"
#include "ctx.hpp"
#include "launch.hpp"

using namespace ztd;

namespace ztg {

static constexpr double kRayEpsA = 1e-4;  // render.zig:26
static constexpr int kEmitterSlots = 4;
// Row stride of the per-lane columns (parameter p of lane l at p * kColStride
// + l): 65 words, so the reduction's lanes (one parameter each, walking the
// 64 lanes of a row) hit 64 different LDS banks.
static constexpr int kColStride = 65;
// Parameters whose columns fit in LDS (kColStride * sizeof(R) bytes each per
// wave): beyond this the columns live in a global buffer.
static constexpr size_t kLdsColBytes = 32 * 1024;

template <typename R, int K>
struct AdjointObserver {
  V3<R> delta;  // dLoss / d(sample radiance) = dLoss/dpixel / spp
  int mat[K];
  int cnt[K];
  V3<R> g[K];
  int n;
  int emat[kEmitterSlots];
  V3<R> eg[kEmitterSlots];
  int ne;
  V3<R> nee_x, nee_dle;
  int nee_mat;
  int overflow;
  R* col;  // this lane's column: parameter p at col[p * kColStride]
  const int* offsets;

  ZT_D void reset() {
    n = 0;
    ne = 0;
  }
  ZT_D void add(int off, V3<R> v) {
    col[(off + 0) * kColStride] += v.x;
    col[(off + 1) * kColStride] += v.y;
    col[(off + 2) * kColStride] += v.z;
  }
  ZT_D void event(V3<R> x) {
    const V3<R> xd = mul(delta, x);
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (i < n) g[i] = ztd::add(g[i], scale(xd, R(cnt[i])));
  }
  ZT_D void addEmitter(int e, V3<R> v) {
    bool done = false;
#pragma unroll
    for (int i = 0; i < kEmitterSlots; ++i) {
      if (!done && i < ne && emat[i] == e) {
        eg[i] = ztd::add(eg[i], v);
        done = true;
      }
    }
    if (done) return;
    if (ne < kEmitterSlots) {
#pragma unroll
      for (int i = 0; i < kEmitterSlots; ++i) {
        if (i == ne) {
          emat[i] = e;
          eg[i] = v;
        }
      }
      ++ne;
      return;
    }
    flushEmitter(e, v);  // more distinct emitters than slots: add at once
  }
  ZT_D void flushEmitter(int e, V3<R> v) {
    const int off = offsets[e];
    if (off >= 0) add(off, v);
  }
  // NoObserver hooks (see dev_path.hpp).
  ZT_D void miss(V3<R> x) { event(x); }
  ZT_D void emission(V3<R> x, V3<R> dle, int e) {
    event(x);
    addEmitter(e, mul(delta, dle));
  }
  ZT_D void scattered(int m) {
    bool found = false;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (!found && i < n && mat[i] == m) {
        cnt[i] += 1;
        found = true;
      }
    }
    if (found) return;
    if (n >= K) {
      overflow = 1;
      return;
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i == n) {
        mat[i] = m;
        cnt[i] = 1;
        g[i] = splat(R(0));
      }
    }
    ++n;
  }
  ZT_D void neePending(V3<R> x, V3<R> dle, int e) {
    nee_x = x;
    nee_dle = dle;
    nee_mat = e;
  }
  ZT_D void neeCommitted() {
    event(nee_x);
    addEmitter(nee_mat, mul(delta, nee_dle));
  }
  ZT_D void flush(const SceneView<R>& sc) {
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i < n) {
        const int off = offsets[mat[i]];
        if (off >= 0) {
          const V3<R> a = sc.mats[mat[i]].color;
          // Zero-albedo channels carry no contribution to divide (0/0): their
          // gradient is left at 0.
          if (a.x != 0) col[(off + 0) * kColStride] += g[i].x / a.x;
          if (a.y != 0) col[(off + 1) * kColStride] += g[i].y / a.y;
          if (a.z != 0) col[(off + 2) * kColStride] += g[i].z / a.z;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kEmitterSlots; ++i)
      if (i < ne) flushEmitter(emat[i], eg[i]);
  }
};

// Path replay of the pixels of 8x8 groups. A workgroup of tile_w x tile_h
// pixels (multiples of 8, one wave per group) works through the tiles of the
// image in turn (persistent grid: tile blockIdx.x, then + gridDim.x, ...);
// every group's row of partials depends only on its pixels. `gcols`: the
// per-lane columns in global memory (one block of kColStride * nparams words
// per wave of the grid) when they do not fit in LDS (kLdsCols false).
template <typename R, int K, bool kLdsCols>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 8))) void adjointKernel(
    SceneView<R> sc, Params pr, const float* __restrict__ dimg, const int* __restrict__ offsets, int nparams,
    R* __restrict__ gcols, double* __restrict__ partial, uint32_t tile_w, uint32_t tile_h, uint32_t tiles_x,
    uint32_t ntiles, uint32_t groups_x, int cap, int* spill, int* err) {
  extern __shared__ int lds[];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t waves = blockDim.x >> 6;
  const LaneStack stk = laneStack(lds, cap, spill);
  R* cols;  // this wave's columns
  if constexpr (kLdsCols) {
    cols = (R*)(lds + stackBytes(cap, (int)blockDim.x) / sizeof(int)) + (size_t)wave * kColStride * nparams;
  } else {
    cols = gcols + ((size_t)blockIdx.x * waves + wave) * kColStride * nparams;
  }
  const uint32_t gw = tile_w / 8;  // groups per tile row
  int overflow = 0;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    // This wave's group within the tile, and its lane's pixel.
    const uint32_t gx = (tile % tiles_x) * gw + wave % gw;
    const uint32_t gy = (tile / tiles_x) * (tile_h / 8) + wave / gw;
    const uint32_t x = gx * 8 + lane % 8, y = gy * 8 + lane / 8;
    for (int p = 0; p < nparams; ++p) cols[p * kColStride + lane] = R(0);
    if (x < pr.width && y < pr.height) {
      const uint64_t pix = (uint64_t)y * pr.width + x;
      Rng rng = rngForStream(pr.seed, pix);
      AdjointObserver<R, K> obs;
      const R inv_spp = R(1) / R(pr.spp);
      obs.delta = mk(R(dimg[3 * pix + 0]) * inv_spp, R(dimg[3 * pix + 1]) * inv_spp, R(dimg[3 * pix + 2]) * inv_spp);
      obs.col = cols + lane;
      obs.offsets = offsets;
      obs.overflow = 0;
      Path<R> st;
      for (uint32_t sample = 0; sample < pr.spp; ++sample) {
        obs.reset();
        int phase = beginPath<R>(sc, pr, x, y, rng, st);
        while (phase != kFinish) {
          const bool sh = phase == kShadow;
          const HitRec<R> h = query<R>(sc, stk, st.ro, sh ? st.sh_d : st.rd, R(kRayEpsA),
                                       sh ? st.sh_tmax : R(INFINITY), st.self, sh ? st.sh_excl : -1, sh, &overflow);
          phase = sh ? afterShadow<R>(h, st, &obs) : shade<R>(sc, pr, h, rng, st, &obs);
        }
        if (isFinite(st.radiance)) obs.flush(sc);  // the forward drops non-finite samples (render.zig:141-142)
      }
      overflow |= obs.overflow << 1;
    }
    __syncthreads();  // every lane's column of this group is written
    if (gx < groups_x && gy * 8 < pr.height) {
      double* row = partial + ((size_t)gy * groups_x + gx) * nparams;
      for (int p = (int)lane; p < nparams; p += 64) {
        double s = 0;
        for (int j = 0; j < 64; ++j) s += (double)cols[p * kColStride + j];
        row[p] = s;
      }
    }
    __syncthreads();  // the columns are reused by the next tile
  }
  if (overflow) atomicOr(err, overflow);
}

// dL/dparams[p] = the sum over groups of partial[g][p]: one wave per
// parameter, lane l summing groups l, l + 64, ... in order, then a fixed
// butterfly across the lanes.
__global__ __launch_bounds__(64) void reduceGroups(const double* __restrict__ partial, int nparams, uint32_t ngroups,
                                                   double* __restrict__ out) {
  const int p = (int)blockIdx.x;
  const uint32_t lane = threadIdx.x;
  double s = 0;
  for (uint32_t g = lane; g < ngroups; g += 64) s += partial[(size_t)g * nparams + p];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[p] = s;
}

// Rebuilds the light table from the emissive candidates with the current
// material colours, exactly as Lights.build does (light.zig:68-94), and
// uploads materials and lights.
void rebuildLights(DevScene& s, hipStream_t stream) {
  s.lights_host.clear();
  std::vector<double> cdf;
  std::vector<double> weights;
  for (const auto& c : s.light_candidates) {
    const zt_material& m = s.materials[c.material];
    if (m.kind != ZT_EMISSIVE) continue;
    const double lum = 0.2126 * m.color[0] + 0.7152 * m.color[1] + 0.0722 * m.color[2];
    const double w = lum * c.weight_area;
    if (!(w > 0) || !std::isfinite(w)) continue;
    Light<double> L = c;
    L.radiance = mk(m.color[0], m.color[1], m.color[2]);
    s.lights_host.push_back(L);
    weights.push_back(w);
  }
  double sum = 0;
  for (double w : weights) {
    sum += w;
    cdf.push_back(sum);
  }
  for (size_t i = 0; i < s.lights_host.size(); ++i) s.lights_host[i].select = weights[i] / sum;
  s.num_lights = (int)s.lights_host.size();
  s.light_total = sum;
  std::vector<int> prim_light(std::max(s.num_prims, 1), -1);
  for (size_t i = 0; i < s.lights_host.size(); ++i) prim_light[s.lights_host[i].prim] = (int)i;
  s.cdf.upload(cdf, stream);
  s.prim_light.upload(prim_light, stream);
  uploadMaterialsAndLights(s, stream);
}

void setParams(zt_ctx* ctx, DevScene& s, const float* params) {
  if (!params) throw Error(ZT_ERR_INVALID_ARGUMENT, "null parameters");
  for (size_t m = 0; m < s.materials.size(); ++m) {
    const int off = s.param_offset[m];
    if (off < 0) continue;
    for (int c = 0; c < 3; ++c) {
      const float v = params[off + c];
      if (!(v >= 0) || !std::isfinite(v)) throw Error(ZT_ERR_INVALID_ARGUMENT, "parameters must be finite and >= 0");
      s.materials[m].color[c] = (double)v;
    }
  }
  rebuildLights(s, ctx->stream);
  ZT_HIP(hipStreamSynchronize(ctx->stream));
}

// Traversal stack entries a query of this scene can hold at once: a
// two-child walk pushes at most depth + 1, a four-wide one 3 * depth + 1.
int stackNeed(const DevScene& s, bool four_wide) {
  if (s.brute_force || s.num_nodes == 0) return 0;
  return four_wide ? 3 * s.tree_depth4 + 1 : s.tree_depth + 1;
}

// Launch shape and buffer sizes of one backward pass.
struct AdjointPlan {
  uint32_t tw, th, threads, tiles_x, ntiles, groups_x, ngroups, grid;
  int cap;        // LDS stack entries per lane
  bool lds_cols;  // the per-lane columns fit in LDS
  const void* fn;
  size_t lds, partial_bytes, cols_bytes, spill_bytes;
};

template <typename R, int K>
AdjointPlan planAdjoint(const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, int cap) {
  AdjointPlan p{};
  p.tw = tu.tile_w ? tu.tile_w : 8;
  p.th = tu.tile_h ? tu.tile_h : 8;
  if (p.tw % 8 || p.th % 8 || p.tw * p.th > 256)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "adjoint tiles are 8x8 groups of pixels: multiples of 8, at most 256 pixels");
  p.threads = p.tw * p.th;
  const uint32_t waves = p.threads / 64;
  p.tiles_x = ((uint32_t)o.width + p.tw - 1) / p.tw;
  p.ntiles = p.tiles_x * (((uint32_t)o.height + p.th - 1) / p.th);
  p.groups_x = ((uint32_t)o.width + 7) / 8;
  p.ngroups = p.groups_x * (((uint32_t)o.height + 7) / 8);
  p.cap = cap;
  const size_t np = (size_t)std::max(s.num_params, 1);
  const size_t col_bytes = (size_t)waves * kColStride * np * sizeof(R);
  p.lds_cols = col_bytes <= kLdsColBytes;
  p.fn = p.lds_cols ? (const void*)adjointKernel<R, K, true> : (const void*)adjointKernel<R, K, false>;
  p.lds = stackBytes(cap, (int)p.threads) + (p.lds_cols ? col_bytes : 0);
  p.grid = std::min<uint32_t>(p.ntiles, residentBlocks(p.fn, p.threads, p.lds));
  p.partial_bytes = (size_t)p.ngroups * np * sizeof(double);
  p.cols_bytes = p.lds_cols ? 0 : (size_t)p.grid * col_bytes;
  p.spill_bytes = spillBytes((uint64_t)p.grid * p.threads);
  return p;
}

// The backward pass into `grad` (num_params doubles, device memory).
template <typename R, int K>
void runAdjoint(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu,
                const AdjointPlan& p, const float* ddimg, double* grad, double* partial, void* cols, int* spill) {
  SceneView<R> view = viewOf<R>(s);
  view.cam = cameraInit(s.camera, (double)o.width / (double)o.height);
  applyBvhWidth(view, s, tu);  // the replay must walk the tree the forward render walked
  const Params pr = makeParams(s, o);
  void* args[] = {&view, (void*)&pr, (void*)&ddimg, (void*)&s.param_offsets.p, (void*)&s.num_params, &cols, &partial,
                  (void*)&p.tw, (void*)&p.th, (void*)&p.tiles_x, (void*)&p.ntiles, (void*)&p.groups_x,
                  (void*)&p.cap, &spill, (void*)&ctx->errflag.p};
  ZT_HIP(hipLaunchKernel(p.fn, dim3(p.grid), dim3(p.threads), args, p.lds, ctx->stream));
  if (s.num_params > 0)
    hipLaunchKernelGGL(reduceGroups, dim3((uint32_t)s.num_params), dim3(64), 0, ctx->stream, partial, s.num_params,
                       p.ngroups, grad);
  ZT_HIP(hipGetLastError());
}

// Dispatch on precision and on the depth class of the observer's lists.
template <typename F>
void withAdjoint(bool f64, uint64_t depth, F&& f) {
  if (f64) {
    if (depth <= 8) f(double{}, std::integral_constant<int, 8>{});
    else if (depth <= 16) f(double{}, std::integral_constant<int, 16>{});
    else f(double{}, std::integral_constant<int, 32>{});
  } else {
    if (depth <= 8) f(float{}, std::integral_constant<int, 8>{});
    else if (depth <= 16) f(float{}, std::integral_constant<int, 16>{});
    else f(float{}, std::integral_constant<int, 32>{});
  }
}

void checkAdjointErr(zt_ctx* ctx) {
  int e = 0;
  ZT_HIP(hipMemcpyAsync(&e, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ZT_HIP(hipStreamSynchronize(ctx->stream));
  if (e & 1) throw Error(ZT_ERR_INTERNAL, "device traversal stack overflow");
  if (e & 2) throw Error(ZT_ERR_INTERNAL, "a path visited more distinct materials than the adjoint tracks");
}

// ---------------------------------------------------------------- allocation-free backward
// Scratch of zt_gpu_backward_noalloc: the rows of per-group partials (8 MiB,
// one copy on every GPU that loads the library, held until the process
// ends). In host code `scratch` is only a stand-in for the symbol: its device
// address on each GPU comes from hipGetSymbolAddress with that GPU current
// (initNoAllocScratch, run by zt_gpu_create) and is used everywhere instead.
static constexpr size_t kNoAllocScratchDoubles = (size_t)1 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];
static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxScratchDevices];

void initNoAllocScratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxScratchDevices) throw Error(ZT_ERR_INVALID_ARGUMENT, "device index out of range");
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  ZT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_dev[device] = (double*)p;
}

static double* noAllocScratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  double* p = device >= 0 && device < kMaxScratchDevices ? g_scratch_dev[device] : nullptr;
  if (!p) throw Error(ZT_ERR_INTERNAL, "no-alloc scratch not set up for this device");
  return p;
}

// Clears the launch status words (a kernel, not a memset node, so the call
// can sit inside a captured graph).
__global__ void clearStatus(int* status) {
  if (threadIdx.x < kStatusBytes / sizeof(int)) status[threadIdx.x] = 0;
}

}  // namespace ztg

using namespace ztg;

extern "C" {

int zt_gpu_param_count(zt_ctx* ctx, uint32_t scene_id, uint32_t* count) {
  if (!ctx || !count) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] { *count = (uint32_t)sceneOf(ctx, scene_id).num_params; });
}

int zt_gpu_param_offsets(zt_ctx* ctx, uint32_t scene_id, int32_t* offsets) {
  if (!ctx || !offsets) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    const DevScene& s = sceneOf(ctx, scene_id);
    for (size_t m = 0; m < s.param_offset.size(); ++m) offsets[m] = s.param_offset[m];
  });
}

int zt_gpu_get_params(zt_ctx* ctx, uint32_t scene_id, float* params) {
  if (!ctx || !params) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    const DevScene& s = sceneOf(ctx, scene_id);
    for (size_t m = 0; m < s.materials.size(); ++m) {
      const int off = s.param_offset[m];
      if (off < 0) continue;
      for (int c = 0; c < 3; ++c) params[off + c] = (float)s.materials[m].color[c];
    }
  });
}

int zt_gpu_set_params(zt_ctx* ctx, uint32_t scene_id, const float* params) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    DevScene& s = sceneOf(ctx, scene_id);
    setDevice(ctx);
    setParams(ctx, s, params);
  });
}

int zt_gpu_forward(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, const zt_gpu_tuning* tuning,
                   const float* params, void* out, int out_on_device) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  int rc = guarded(ctx, [&] {
    checkOptions(opts);
    DevScene& s = sceneOf(ctx, scene_id);
    setDevice(ctx);
    if (params) setParams(ctx, s, params);
  });
  if (rc != ZT_OK) return rc;
  return zt_gpu_render(ctx, scene_id, opts, tuning, out, out_on_device);
}

int zt_gpu_backward(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, const zt_gpu_tuning* tuning,
                    const float* params, const float* dL_dimage, int grad_on_device, double* dL_dparams) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    checkOptions(opts);
    if (!dL_dimage || !dL_dparams) throw Error(ZT_ERR_INVALID_ARGUMENT, "null buffer");
    if (opts->max_depth > 32) throw Error(ZT_ERR_INVALID_ARGUMENT, "the adjoint supports max_depth <= 32");
    DevScene& s = sceneOf(ctx, scene_id);
    zt_gpu_tuning tu{};
    if (tuning) tu = *tuning;
    if (tu.precision > ZT_PRECISION_FP64) throw Error(ZT_ERR_INVALID_ARGUMENT, "bad precision");
    setDevice(ctx);
    if (params) setParams(ctx, s, params);
    const bool f64 = tu.precision == ZT_PRECISION_FP64;
    const size_t npix = (size_t)opts->width * opts->height;
    const float* ddimg = dL_dimage;
    DevBuf dimg_buf;
    if (!grad_on_device) {
      dimg_buf.alloc(npix * 3 * sizeof(float));
      ZT_HIP(hipMemcpyAsync(dimg_buf.p, dL_dimage, npix * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      ddimg = dimg_buf.as<float>();
    }
    const int np = std::max(s.num_params, 1);
    ctx->adj_grad.ensure((size_t)np * sizeof(double));
    resetErr(ctx);
    ZT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    withAdjoint(f64, opts->max_depth, [&](auto tag, auto k) {
      using R = decltype(tag);
      constexpr int kList = decltype(k)::value;
      const AdjointPlan p = planAdjoint<R, kList>(s, *opts, tu, kLds);
      ctx->adj_partial.ensure(p.partial_bytes);
      ctx->adj_cols.ensure(p.cols_bytes);
      ctx->spill.ensure(p.spill_bytes);
      runAdjoint<R, kList>(ctx, s, *opts, tu, p, ddimg, ctx->adj_grad.as<double>(), ctx->adj_partial.as<double>(),
                       ctx->adj_cols.p, ctx->spill.as<int>());
    });
    ZT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> gd(np);
    ZT_HIP(hipMemcpyAsync(gd.data(), ctx->adj_grad.p, np * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    checkAdjointErr(ctx);
    float ms = 0;
    ZT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms = ms;
    for (int i = 0; i < s.num_params; ++i) dL_dparams[i] = gd[i];
  });
}

int zt_gpu_backward_noalloc(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts,
                            const zt_gpu_tuning* tuning, const float* dL_dimage, double* dL_dparams) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    checkOptions(opts);
    if (!dL_dimage || !dL_dparams) throw Error(ZT_ERR_INVALID_ARGUMENT, "null buffer");
    if (opts->max_depth > 32) throw Error(ZT_ERR_INVALID_ARGUMENT, "the adjoint supports max_depth <= 32");
    const DevScene& s = sceneOf(ctx, scene_id);
    zt_gpu_tuning tu{};
    if (tuning) tu = *tuning;
    if (tu.precision > ZT_PRECISION_FP64) throw Error(ZT_ERR_INVALID_ARGUMENT, "bad precision");
    setDevice(ctx);
    double* partial = noAllocScratch(ctx->device);
    withAdjoint(tu.precision == ZT_PRECISION_FP64, opts->max_depth, [&](auto tag, auto k) {
      using R = decltype(tag);
      constexpr int kList = decltype(k)::value;
      // The whole traversal stack in LDS, sized to the deepest the scene's
      // tree can need, so no spill buffer exists.
      SceneView<R> v = viewOf<R>(s);
      applyBvhWidth(v, s, tu);
      const int need = stackNeed(s, !K<R>::exact && v.nodes4 != nullptr);
      if (need > 64)
        throw Error(ZT_ERR_INVALID_ARGUMENT, "the scene's BVH needs more than the 64 traversal stack entries that "
                                             "fit in LDS without a spill buffer");
      const AdjointPlan p = planAdjoint<R, kList>(s, *opts, tu, std::max(8, (need + 1) & ~1));
      if (!p.lds_cols)
        throw Error(ZT_ERR_INVALID_ARGUMENT, "too many parameters for the per-lane columns in LDS");
      if ((size_t)p.ngroups * (size_t)std::max(s.num_params, 1) > kNoAllocScratchDoubles)
        throw Error(ZT_ERR_INVALID_ARGUMENT, "8x8 pixel groups x parameters exceed the 8 MiB no-alloc scratch "
                                             "(1048576 partial sums)");
      hipLaunchKernelGGL(clearStatus, dim3(1), dim3(64), 0, ctx->stream, ctx->errflag.as<int>());
      runAdjoint<R, kList>(ctx, s, *opts, tu, p, dL_dimage, dL_dparams, partial, nullptr, nullptr);
    });
    ZT_HIP(hipMemcpyAsync(ctx->status_host, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZT_HIP(hipStreamSynchronize(ctx->stream));
    const int e = ctx->status_host[0];
    if (e & 1) throw Error(ZT_ERR_INTERNAL, "device traversal stack overflow");
    if (e & 2) throw Error(ZT_ERR_INTERNAL, "a path visited more distinct materials than the adjoint tracks");
  });
}

}  // extern "C"
