// Launch helpers shared by the renderer translation units.
#pragma once

#include "../device/dev_path.hpp"
#include "ctx.hpp"

#include <vector>

namespace ztg {

ztd::Params makeParams(const DevScene& s, const zt_render_options& o);
void checkOptions(const zt_render_options* o);
void resetErr(zt_ctx* ctx);
void checkErr(zt_ctx* ctx);
void uploadMaterialsAndLights(DevScene& s, hipStream_t stream);
void rebuildLights(DevScene& s, hipStream_t stream);

// One scene of a wave-scheduled launch (fp32, four-wide BVH): its view,
// render constants, output and its first tile in the launch's tile space.
struct WaveJob {
  ztd::SceneView<float> view;
  ztd::Params pr;
  float* out;
  uint32_t tiles_x, first_tile;
};

// Wave-scheduled megakernel over one or many scenes in one launch (tiles
// numbered scene after scene). Fills tiles_x / first_tile.
void launchWaveJobs(zt_ctx* ctx, std::vector<WaveJob>& jobs, const zt_gpu_tuning& tu);

// Longest-first order of one scene's 8x8 tiles (fp32, four-wide BVH): the
// wave-scheduled megakernel's one-sample cost probe without next-event rays
// (it writes a one-sample image to `out`) and the per-band bucket sort, on
// the context stream. Returns ctx->tile_order: each of the 8 tile bands
// sorted most expensive first.
const uint32_t* probeTileOrder(zt_ctx* ctx, const ztd::SceneView<float>& view, const ztd::Params& pr, float* out,
                               int cap, uint32_t refill);

// Blocks of `kernel` the device can hold at once (occupancy x CUs).
uint32_t residentBlocks(const void* kernel, uint32_t threads, size_t lds);

// AUTO resolution: the autotuner's record for this render class, else the
// megakernel defaults; explicit methods pass through (enums validated).
TuneKey tuneKey(const DevScene& s, const zt_render_options& o, uint32_t method, uint32_t precision);
zt_gpu_tuning resolveTuning(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning* t);

// Applies zt_gpu_tuning.bvh_width to a view (0 or 4: keep the four-wide
// nodes if the scene has them; 2: two-child traversal).
// Views come with the four-wide tree in 128-byte nodes (nodes4q null);
// tuning selects the two-child tree or the 64-byte nodes.
template <typename R>
void applyBvhWidth(ztd::SceneView<R>& v, const DevScene& s, const zt_gpu_tuning& tu) {
  if (tu.bvh_width != 0 && tu.bvh_width != 2 && tu.bvh_width != 4)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "bvh_width must be 0, 2 or 4");
  if (tu.node_bytes != 0 && tu.node_bytes != 64 && tu.node_bytes != 128)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "node_bytes must be 0, 64 or 128");
  if (tu.bvh_width == 2) v.nodes4 = nullptr;
  if (v.nodes4 && tu.node_bytes == 64 && s.nodes4q.p) v.nodes4q = s.nodes4q.as<ztd::DNode4Q>();
}

template <typename R>
void launchMegakernel(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, R* dout);
template <typename R>
void launchWavefront(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, R* dout);

}  // namespace ztg
