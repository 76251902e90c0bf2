// Launch helpers shared by the renderer translation units.
#pragma once

#include "../device/dev_path.hpp"
#include "ctx.hpp"

namespace ztg {

ztd::Params makeParams(const DevScene& s, const zt_render_options& o);
void checkOptions(const zt_render_options* o);
void resetErr(zt_ctx* ctx);
void checkErr(zt_ctx* ctx);
void uploadMaterialsAndLights(DevScene& s, hipStream_t stream);
void rebuildLights(DevScene& s, hipStream_t stream);

// AUTO resolution: the autotuner's record for this render class, else the
// megakernel defaults; explicit methods pass through (enums validated).
TuneKey tuneKey(const DevScene& s, const zt_render_options& o, uint32_t method, uint32_t precision);
zt_gpu_tuning resolveTuning(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning* t);

template <typename R>
void launchMegakernel(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, R* dout);
template <typename R>
void launchWavefront(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, R* dout);

}  // namespace ztg
