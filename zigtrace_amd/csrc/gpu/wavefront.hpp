// Device records of the wavefront / batch renderer.
#pragma once

#include <vector>

#include "../device/dev_path.hpp"
#include "ctx.hpp"

namespace ztg {

// One scene of a batch: its device view, render constants, output and the
// offset of its first pixel in the batch's concatenated pixel range.
template <typename R>
struct WfScene {
  ztd::SceneView<R> view;
  ztd::Params pr;
  R* out;
  uint64_t first_pixel;
};

// One path slot: the path in flight and the pixel it belongs to.
template <typename R>
struct __align__(16) WfSlot {
  ztd::Path<R> p;
  ztd::Rng rng;
  ztd::V3<R> sum;
  uint32_t sample;
  uint32_t scene;  // 0xFFFFFFFF: idle
  uint32_t pixel;  // pixel index within its scene (y * width + x)
};

template <typename R>
void runWavefront(zt_ctx* ctx, const std::vector<WfScene<R>>& scenes, const zt_gpu_tuning& tu);

}  // namespace ztg
