// Output.encode on the device (image.zig:30-37): HDR framebuffer -> RGB8
// without copying the HDR image to the host. Same operations as the host
// encoder in f64 (the fp32 framebuffer is widened first, as Image holds f64).

#include "ctx.hpp"
#include "launch.hpp"

namespace ztg {

ZT_D double tonemapD(uint32_t op, double c) {  // image.zig:15-22
  const double x = fmax(c, 0.0);
  if (op == ZT_TONEMAP_REINHARD) return x / (1.0 + x);
  if (op == ZT_TONEMAP_ACES) return (x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14);
  return x;
}

template <typename T>
__global__ void encodeKernel(const T* __restrict__ hdr, uint64_t n, zt_output o, uint8_t* __restrict__ rgb) {
  const double inv_gamma = 1.0 / o.gamma;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const double linear = (double)hdr[i];
    const double v = isnan(linear) ? 0.0 : linear;
    const double mapped = fmin(fmax(tonemapD(o.tonemap, v * o.exposure), 0.0), 1.0);
    const double corrected = pow(mapped, inv_gamma);
    rgb[i] = (uint8_t)round(fmin(fmax(corrected, 0.0), 1.0) * 255.0);
  }
}

}  // namespace ztg

using namespace ztg;

extern "C" int zt_gpu_encode_rgb8(zt_ctx* ctx, const void* hdr, uint32_t precision, uint64_t count,
                                  const zt_output* out, uint8_t* rgb, int hdr_on_device, int rgb_on_device) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (count == 0) return;
    if (!hdr || !out || !rgb) throw Error(ZT_ERR_INVALID_ARGUMENT, "null argument");
    if (precision > ZT_PRECISION_FP64 || out->tonemap > ZT_TONEMAP_ACES)
      throw Error(ZT_ERR_INVALID_ARGUMENT, "bad precision or tonemap");
    setDevice(ctx);
    const size_t elem = precision == ZT_PRECISION_FP64 ? 8 : 4;
    DevBuf din, dout;
    const void* src = hdr;
    if (!hdr_on_device) {
      din.alloc(count * elem);
      ZT_HIP(hipMemcpyAsync(din.p, hdr, count * elem, hipMemcpyHostToDevice, ctx->stream));
      src = din.p;
    }
    uint8_t* dst = rgb;
    if (!rgb_on_device) {
      dout.alloc(count);
      dst = dout.as<uint8_t>();
    }
    const uint32_t threads = 256;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((count + threads - 1) / threads, 65536);
    if (elem == 8)
      hipLaunchKernelGGL(encodeKernel<double>, dim3(blocks), dim3(threads), 0, ctx->stream, (const double*)src, count,
                         *out, dst);
    else
      hipLaunchKernelGGL(encodeKernel<float>, dim3(blocks), dim3(threads), 0, ctx->stream, (const float*)src, count,
                         *out, dst);
    ZT_HIP(hipGetLastError());
    if (!rgb_on_device) ZT_HIP(hipMemcpyAsync(rgb, dst, count, hipMemcpyDeviceToHost, ctx->stream));
    ZT_HIP(hipStreamSynchronize(ctx->stream));
  });
}
