// Wavefront renderer (under construction).
#include "launch.hpp"

namespace ztg {
template <typename R>
void launchWavefront(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&, R*) {
  throw Error(ZT_ERR_INVALID_ARGUMENT, "wavefront method not available");
}
template void launchWavefront<float>(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&, float*);
template void launchWavefront<double>(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&,
                                      double*);
}  // namespace ztg
