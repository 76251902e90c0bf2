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
// a finite sample the list is divided through and added to a per-workgroup
// LDS accumulator, which is flushed to global memory once per workgroup.

#include <algorithm>
#include <cmath>
#include <vector>

#include "../device/dev_path.hpp"
#include "ctx.hpp"
#include "launch.hpp"

using namespace ztd;

namespace ztg {

static constexpr double kRayEpsA = 1e-4;  // render.zig:26
static constexpr int kEmitterSlots = 4;

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
  R* lds_grad;
  const int* offsets;

  ZT_D void reset() {
    n = 0;
    ne = 0;
  }
  ZT_D void event(V3<R> x) {
    const V3<R> xd = mul(delta, x);
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (i < n) g[i] = add(g[i], scale(xd, R(cnt[i])));
  }
  ZT_D void addEmitter(int e, V3<R> v) {
    bool done = false;
#pragma unroll
    for (int i = 0; i < kEmitterSlots; ++i) {
      if (!done && i < ne && emat[i] == e) {
        eg[i] = add(eg[i], v);
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
    if (off < 0) return;
    atomicAdd(&lds_grad[off + 0], v.x);
    atomicAdd(&lds_grad[off + 1], v.y);
    atomicAdd(&lds_grad[off + 2], v.z);
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
          if (a.x != 0) atomicAdd(&lds_grad[off + 0], g[i].x / a.x);
          if (a.y != 0) atomicAdd(&lds_grad[off + 1], g[i].y / a.y);
          if (a.z != 0) atomicAdd(&lds_grad[off + 2], g[i].z / a.z);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kEmitterSlots; ++i)
      if (i < ne) flushEmitter(emat[i], eg[i]);
  }
};

template <typename R, int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 8))) void adjointKernel(SceneView<R> sc, Params pr, const float* __restrict__ dimg,
                                                     const int* __restrict__ offsets, int nparams,
                                                     R* __restrict__ grad, uint32_t tile_w, uint32_t tiles_x,
                                                     uint32_t ntiles, int* spill, int* err) {
  extern __shared__ int lds[];
  const uint32_t tid = threadIdx.x;
  const LaneStack stk = laneStack<64>(lds, kLds, spill);  // workgroups of one wave (launchAdjoint)
  R* lg = (R*)(lds + 2 * kLds * blockDim.x);
  for (int i = (int)tid; i < nparams; i += (int)blockDim.x) lg[i] = R(0);
  __syncthreads();
  const uint32_t tile = blockIdx.x;
  const uint32_t tile_h = blockDim.x / tile_w;
  const uint32_t x = (tile % tiles_x) * tile_w + tid % tile_w;
  const uint32_t y = (tile / tiles_x) * tile_h + tid / tile_w;
  int overflow = 0;
  if (x < pr.width && y < pr.height) {
    const uint64_t pix = (uint64_t)y * pr.width + x;
    Rng rng = rngForStream(pr.seed, pix);
    AdjointObserver<R, K> obs;
    const R inv_spp = R(1) / R(pr.spp);
    obs.delta = mk(R(dimg[3 * pix + 0]) * inv_spp, R(dimg[3 * pix + 1]) * inv_spp, R(dimg[3 * pix + 2]) * inv_spp);
    obs.lds_grad = lg;
    obs.offsets = offsets;
    obs.overflow = 0;
    Path<R> st;
    for (uint32_t sample = 0; sample < pr.spp; ++sample) {
      obs.reset();
      int phase = beginPath<R>(sc, pr, x, y, rng, st);
      while (phase != kFinish) {
        const bool sh = phase == kShadow;
        const HitRec<R> h = query<R>(sc, stk, st.ro, sh ? st.sh_d : st.rd, R(kRayEpsA), sh ? st.sh_tmax : R(INFINITY),
                                     st.self, sh ? st.sh_excl : -1, sh, &overflow);
        phase = sh ? afterShadow<R>(h, st, &obs) : shade<R>(sc, pr, h, rng, st, &obs);
      }
      if (isFinite(st.radiance)) obs.flush(sc);  // the forward drops non-finite samples (render.zig:141-142)
    }
    overflow |= obs.overflow << 1;
  }
  __syncthreads();
  for (int i = (int)tid; i < nparams; i += (int)blockDim.x)
    if (lg[i] != R(0)) atomicAdd(&grad[i], lg[i]);
  if (overflow) atomicOr(err, overflow);
}

template <typename R>
__global__ void zeroFill(R* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = R(0);
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

template <typename R, int K>
void launchAdjoint(zt_ctx* ctx, DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, const float* ddimg,
                   R* dgrad) {
  SceneView<R> view = viewOf<R>(s);
  view.cam = cameraInit(s.camera, (double)o.width / (double)o.height);
  applyBvhWidth(view, s, tu);  // the replay must walk the tree the forward render walked
  const Params pr = makeParams(s, o);
  const uint32_t tw = 8, threads = 64;
  const uint32_t tiles_x = (pr.width + tw - 1) / tw;
  const uint32_t tiles_y = (pr.height + (threads / tw) - 1) / (threads / tw);
  const size_t lds = stackBytes(kLds, (int)threads) + (size_t)s.num_params * sizeof(R);
  ctx->spill.ensure(spillBytes((uint64_t)tiles_x * tiles_y * threads));
  hipLaunchKernelGGL((adjointKernel<R, K>), dim3(tiles_x * tiles_y), dim3(threads), lds, ctx->stream, view, pr, ddimg,
                     s.param_offsets.as<int>(), s.num_params, dgrad, tw, tiles_x, tiles_x * tiles_y,
                     ctx->spill.as<int>(), ctx->errflag.as<int>());
  ZT_HIP(hipGetLastError());
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
    DevBuf dimg_buf, grad_buf;
    if (!grad_on_device) {
      dimg_buf.alloc(npix * 3 * sizeof(float));
      ZT_HIP(hipMemcpyAsync(dimg_buf.p, dL_dimage, npix * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      ddimg = dimg_buf.as<float>();
    }
    const int np = std::max(s.num_params, 1);
    grad_buf.alloc((size_t)np * (f64 ? sizeof(double) : sizeof(float)));
    resetErr(ctx);
    ZT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (f64) hipLaunchKernelGGL(zeroFill<double>, dim3(1), dim3(256), 0, ctx->stream, grad_buf.as<double>(), np);
    else hipLaunchKernelGGL(zeroFill<float>, dim3(1), dim3(256), 0, ctx->stream, grad_buf.as<float>(), np);
    const uint64_t d = opts->max_depth;
    if (f64) {
      if (d <= 8) launchAdjoint<double, 8>(ctx, s, *opts, tu, ddimg, grad_buf.as<double>());
      else if (d <= 16) launchAdjoint<double, 16>(ctx, s, *opts, tu, ddimg, grad_buf.as<double>());
      else launchAdjoint<double, 32>(ctx, s, *opts, tu, ddimg, grad_buf.as<double>());
    } else {
      if (d <= 8) launchAdjoint<float, 8>(ctx, s, *opts, tu, ddimg, grad_buf.as<float>());
      else if (d <= 16) launchAdjoint<float, 16>(ctx, s, *opts, tu, ddimg, grad_buf.as<float>());
      else launchAdjoint<float, 32>(ctx, s, *opts, tu, ddimg, grad_buf.as<float>());
    }
    ZT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> gd(np);
    if (f64) {
      ZT_HIP(hipMemcpyAsync(gd.data(), grad_buf.p, np * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    } else {
      std::vector<float> gf(np);
      ZT_HIP(hipMemcpyAsync(gf.data(), grad_buf.p, np * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
      ZT_HIP(hipStreamSynchronize(ctx->stream));
      for (int i = 0; i < np; ++i) gd[i] = gf[i];
    }
    int e = 0;
    ZT_HIP(hipMemcpyAsync(&e, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZT_HIP(hipStreamSynchronize(ctx->stream));
    if (e & 1) throw Error(ZT_ERR_INTERNAL, "device traversal stack overflow");
    if (e & 2) throw Error(ZT_ERR_INTERNAL, "a path visited more distinct materials than the adjoint tracks");
    float ms = 0;
    ZT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms = ms;
    for (int i = 0; i < s.num_params; ++i) dL_dparams[i] = gd[i];
  });
}

}  // extern "C"
