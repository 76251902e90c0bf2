// Launch-configuration autotuner and the AUTO method's lookup.
//
// Every knob of zt_gpu_tuning leaves the image unchanged (same bits), so the
// tuner is free to pick purely on time: it renders a reduced-spp probe of the
// caller's frame under each candidate, times it with HIP events on the
// context stream, and records the fastest per (method, precision, image size,
// primitive-count class) in the context.

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.hpp"
#include "launch.hpp"

namespace ztg {

static uint32_t bitLength(uint64_t v) {
  uint32_t b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

TuneKey tuneKey(const DevScene& s, const zt_render_options& o, uint32_t method, uint32_t precision) {
  TuneKey k{};
  k.method = method;
  k.precision = precision;
  k.width = o.width;
  k.height = o.height;
  k.prim_class = bitLength((uint64_t)s.num_prims);
  k.brute = s.brute_force ? 1u : 0u;
  return k;
}

zt_gpu_tuning resolveTuning(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning* t) {
  zt_gpu_tuning tu{};
  if (t) tu = *t;
  if (tu.precision > ZT_PRECISION_FP64 || tu.method > ZT_METHOD_WAVEFRONT)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "bad tuning enum");
  if (tu.method != ZT_METHOD_AUTO) return tu;
  const TuneEntry* best = nullptr;
  for (uint32_t m : {(uint32_t)ZT_METHOD_MEGAKERNEL, (uint32_t)ZT_METHOD_WAVEFRONT}) {
    auto it = ctx->tuned.find(tuneKey(s, o, m, tu.precision));
    if (it != ctx->tuned.end() && (!best || it->second.ms < best->ms)) best = &it->second;
  }
  if (best) return best->tu;
  zt_gpu_tuning d{};
  d.method = ZT_METHOD_MEGAKERNEL;
  d.precision = tu.precision;
  return d;
}

// Candidate configurations of one method (0 = library default for a knob).
static std::vector<zt_gpu_tuning> candidates(uint32_t method, uint32_t precision, uint64_t pixels) {
  std::vector<zt_gpu_tuning> c;
  if (method == ZT_METHOD_MEGAKERNEL) {
    const uint32_t tiles[][2] = {{8, 8}, {16, 16}, {8, 32}};
    for (uint32_t variant : {1u, 2u, 3u, 4u})
      for (uint32_t cap : {8u, 16u})
        for (const auto& t : tiles) {
          zt_gpu_tuning tu{};
          tu.method = method;
          tu.precision = precision;
          tu.variant = variant;
          tu.stack_entries = cap;
          tu.tile_w = t[0];
          tu.tile_h = t[1];
          c.push_back(tu);
        }
    for (auto& tu : c) tu.schedule = 1;  // the forms above: one pixel per lane
    // Wave-scheduled forms: fp32 over the four-wide tree, fp64 over the
    // two-child tree (where variants 2-4 target 2-4 waves per SIMD).
    const bool f32 = precision == ZT_PRECISION_FP32;
    for (uint32_t variant : f32 ? std::vector<uint32_t>{1u, 2u} : std::vector<uint32_t>{2u, 3u, 4u})
      for (uint32_t refill : {24u, 32u, 40u})
        for (uint32_t cap : f32 ? std::vector<uint32_t>{8u, 12u, 16u} : std::vector<uint32_t>{12u}) {
          zt_gpu_tuning tu{};
          tu.method = method;
          tu.precision = precision;
          tu.variant = variant;
          tu.stack_entries = cap;
          tu.schedule = 2;
          tu.refill = refill;
          c.push_back(tu);
        }
  } else {
    std::vector<uint32_t> pools;
    for (uint64_t q : {1ull << 20, 1ull << 21, 1ull << 22}) {
      const uint32_t v = (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(q, pixels));
      if (std::find(pools.begin(), pools.end(), v) == pools.end()) pools.push_back(v);
    }
    auto add = [&](uint32_t sched, uint32_t block, uint32_t q, uint32_t cap, uint32_t refill) {
      zt_gpu_tuning tu{};
      tu.method = method;
      tu.precision = precision;
      tu.block_size = block;
      tu.queue_paths = q;
      tu.stack_entries = cap;
      tu.schedule = sched;
      tu.refill = refill;
      c.push_back(tu);
    };
    if (precision == ZT_PRECISION_FP32) {
      // Sample passes (their default pool, which holds every pixel; refill
      // and stack around their defaults), then one configuration of each
      // bounce-level form. The block size only shapes the one-ray-per-lane
      // kernels these do not use.
      for (uint32_t refill : {48u, 56u})
        for (uint32_t cap : {12u, 16u}) add(5, 0, 0, cap, refill);
      add(4, 0, pools.back(), 16, 40);
      for (uint32_t cap : {12u, 16u}) add(3, 0, pools.back(), cap, 0);
      add(2, 256, pools.back(), 12, 0);
    } else {  // one ray per lane
      for (uint32_t block : {128u, 256u})
        for (uint32_t q : pools)
          for (uint32_t cap : {8u, 16u, 32u}) add(0, block, q, cap, 0);
    }
  }
  return c;
}

}  // namespace ztg

using namespace ztg;

extern "C" int zt_gpu_autotune(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, uint32_t method,
                               uint32_t precision, uint32_t probe_spp, zt_gpu_tuning* best, double* best_ms) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    checkOptions(opts);
    if (method > ZT_METHOD_WAVEFRONT || precision > ZT_PRECISION_FP64)
      throw Error(ZT_ERR_INVALID_ARGUMENT, "bad method or precision");
    DevScene& s = sceneOf(ctx, scene_id);
    setDevice(ctx);
    zt_render_options probe = *opts;
    probe.spp = std::min<uint64_t>(opts->spp, probe_spp ? probe_spp : 16);
    const uint64_t pixels = (uint64_t)probe.width * probe.height;
    const size_t elem = precision == ZT_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    ctx->probe_out.ensure(pixels * 3 * elem);
    std::vector<uint32_t> methods;
    if (method == ZT_METHOD_AUTO) methods = {ZT_METHOD_MEGAKERNEL, ZT_METHOD_WAVEFRONT};
    else methods = {method};

    // One timed probe render; returns milliseconds.
    auto timeOne = [&](const zt_gpu_tuning& tu) -> double {
      resetErr(ctx);
      ZT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
      if (tu.method == ZT_METHOD_WAVEFRONT) {
        if (elem == 8) launchWavefront<double>(ctx, s, probe, tu, ctx->probe_out.as<double>());
        else launchWavefront<float>(ctx, s, probe, tu, ctx->probe_out.as<float>());
      } else {
        if (elem == 8) launchMegakernel<double>(ctx, s, probe, tu, ctx->probe_out.as<double>());
        else launchMegakernel<float>(ctx, s, probe, tu, ctx->probe_out.as<float>());
      }
      ZT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
      checkErr(ctx);
      float ms = 0;
      ZT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
      return ms;
    };

    const TuneEntry* overall = nullptr;
    for (uint32_t m : methods) {
      const std::vector<zt_gpu_tuning> cands = candidates(m, precision, pixels);
      (void)timeOne(cands.front());  // warm-up: first-touch of code and buffers
      std::vector<std::pair<double, size_t>> t;
      for (size_t i = 0; i < cands.size(); ++i) t.emplace_back(timeOne(cands[i]), i);
      std::sort(t.begin(), t.end());
      // Re-time the three fastest twice more and keep each one's minimum.
      const size_t top = std::min<size_t>(3, t.size());
      for (size_t j = 0; j < top; ++j)
        for (int r = 0; r < 2; ++r) t[j].first = std::min(t[j].first, timeOne(cands[t[j].second]));
      std::sort(t.begin(), t.begin() + top);
      TuneEntry e{cands[t[0].second], t[0].first};
      auto& slot = ctx->tuned[tuneKey(s, *opts, m, precision)];
      slot = e;
      if (!overall || slot.ms < overall->ms) overall = &slot;
    }
    if (best) *best = overall->tu;
    if (best_ms) *best_ms = overall->ms;
  });
}
