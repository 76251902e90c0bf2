// Wavefront renderer: a pool of slots, compacted queues of them, and
// persistent kernels that work through a queue, the loop over queues
// captured in a hipGraph. Forms (zt_gpu_tuning.schedule):
//
//   5 sample passes (the default for fp32 four-wide trees): the queue holds
//     pixels; one launch runs one whole sample of every queued pixel under
//     the megakernel's wave-uniform traversal passes, lanes refilling from
//     the queue as their paths end (wfPassBody)
//   4 one bounce per launch: trace, shade, next-event ray (wfStepBody)
//   3 two launches per bounce: extension, then shadow queue, each hit shaded
//     by the lane that traced it (wfBounceBody)
//   2 four launches per bounce: extension trace, shade, shadow trace, resolve
//     (wfTraceWaveBody, wfShade, wfResolve)
//   1 (and fp64) the same with one ray per lane (wfTrace)
//
// A slot owns one pixel at a time (its xoshiro256++ stream and accumulator),
// runs that pixel's samples one after the other as the reference does
// (render.zig:133-144), then claims the next unstarted pixel. Every path runs
// the same device functions as the megakernel (beginPath / shade /
// afterShadow / traverse), so the image is bit-identical to it for any form,
// pool size, block size or queue order.
//
// Queues are kept in slot order: a kernel marks each slot that goes on to a
// queue with a byte flag, and a three-launch stream compaction (count per
// range of 4096 slots, scan of the counts, ordered scatter) turns the flags
// into the queue. Slots claim pixels in order, so neighbouring queue entries
// are neighbouring pixels: a wave walks rays from one small part of the
// image, and the slot records it loads lie in a few contiguous runs. The
// segmented forms (4, 5) cut the slots into 8 segments, one per XCD, whose
// slots claim pixels from their own band of the frame.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../device/dev_path.hpp"
#include "../device/dev_wave.hpp"
#include "ctx.hpp"
#include "launch.hpp"
#include "wavefront.hpp"

using namespace ztd;

namespace ztg {

static constexpr double kRayEps = 1e-4;  // render.zig:26

// Path slots in device memory as a structure of arrays of 16-byte groups
// (32-byte in f64), so that the lanes of a wave touch one contiguous run per
// group instead of 64 records 176 bytes apart. Kernels load a slot into a
// WfSlot (registers), run the device functions on it and store it back.
template <typename R>
struct V4T;
template <>
struct V4T<float> {
  using type = float4;
};
template <>
struct V4T<double> {
  using type = double4;
};
template <typename R>
using V4 = typename V4T<R>::type;

template <typename R>
struct WfSlots {
  V4<R>* ro;     // ray origin, prev_pdf
  V4<R>* rd;     // extension direction
  V4<R>* shd;    // shadow direction, shadow t_max
  V4<R>* rad;    // path radiance
  V4<R>* thr;    // path throughput
  V4<R>* shc;    // pending next-event contribution
  V4<R>* sum;    // pixel accumulator
  int4* meta;    // self, sh_excl, depth, terminated
  uint4* where;  // scene (0xFFFFFFFF: idle), pixel, sample, -
  uint4* rng01;  // xoshiro256++ state words 0 and 1
  uint4* rng23;  // words 2 and 3
  static constexpr size_t kBytes = 7 * sizeof(V4<R>) + 4 * sizeof(uint4);  // per slot
};

template <typename R>
ZT_D V3<R> xyz(const V4<R>& v) {
  return mk(v.x, v.y, v.z);
}
template <typename R>
ZT_D V4<R> pack4(V3<R> a, R w) {
  V4<R> r;
  r.x = a.x;
  r.y = a.y;
  r.z = a.z;
  r.w = w;
  return r;
}
ZT_D uint4 packRng(uint64_t a, uint64_t b) {
  return make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
}
ZT_D uint64_t lo64(uint4 v) { return (uint64_t)v.x | ((uint64_t)v.y << 32); }
ZT_D uint64_t hi64(uint4 v) { return (uint64_t)v.z | ((uint64_t)v.w << 32); }

// kPending = false leaves out the pending next-event ray (shd, shc): a slot
// between the one-launch bounces never holds one.
template <typename R, bool kPending = true>
ZT_D void loadSlot(const WfSlots<R>& S, uint32_t i, WfSlot<R>& sl) {
  const V4<R> a = S.ro[i], b = S.rd[i], d = S.rad[i], e = S.thr[i], g = S.sum[i];
  const int4 m = S.meta[i];
  const uint4 wh = S.where[i], r01 = S.rng01[i], r23 = S.rng23[i];
  if constexpr (kPending) {
    const V4<R> c = S.shd[i], f = S.shc[i];
    sl.p.sh_d = xyz<R>(c);
    sl.p.sh_tmax = c.w;
    sl.p.sh_contrib = xyz<R>(f);
  }
  sl.p.ro = xyz<R>(a);
  sl.p.prev_pdf = a.w;
  sl.p.rd = xyz<R>(b);
  sl.p.radiance = xyz<R>(d);
  sl.p.throughput = xyz<R>(e);
  sl.sum = xyz<R>(g);
  sl.p.self = m.x;
  sl.p.sh_excl = m.y;
  sl.p.depth = m.z;
  sl.p.terminated = m.w;
  sl.scene = wh.x;
  sl.pixel = wh.y;
  sl.sample = wh.z;
  sl.rng = Rng{lo64(r01), hi64(r01), lo64(r23), hi64(r23)};
}
template <typename R, bool kPending = true>
ZT_D void storeSlot(const WfSlots<R>& S, uint32_t i, const WfSlot<R>& sl) {
  S.ro[i] = pack4<R>(sl.p.ro, sl.p.prev_pdf);
  S.rd[i] = pack4<R>(sl.p.rd, R(0));
  if constexpr (kPending) {
    S.shd[i] = pack4<R>(sl.p.sh_d, sl.p.sh_tmax);
    S.shc[i] = pack4<R>(sl.p.sh_contrib, R(0));
  }
  S.rad[i] = pack4<R>(sl.p.radiance, R(0));
  S.thr[i] = pack4<R>(sl.p.throughput, R(0));
  S.sum[i] = pack4<R>(sl.sum, R(0));
  S.meta[i] = make_int4(sl.p.self, sl.p.sh_excl, sl.p.depth, sl.p.terminated);
  S.where[i] = make_uint4(sl.scene, sl.pixel, sl.sample, 0u);
  S.rng01[i] = packRng(sl.rng.s0, sl.rng.s1);
  S.rng23[i] = packRng(sl.rng.s2, sl.rng.s3);
}

template <typename R>
struct WfShared {
  WfScene<R>* scenes;
  WfSlots<R> slots;
  float2* hits;       // per slot: (t, prim bits) of the last extension query
  double* hits_d;     // per slot: t in f64 for the fp64 path
  uint32_t* qext[2];  // extension queues (double-buffered)
  uint32_t* qsh;      // shadow queue
  uint8_t* fext;      // per slot: 1 = on the next extension queue
  uint8_t* fsh;       // per slot: 1 = on the shadow queue
  uint32_t* ranges;   // compaction: flags set per range of kRangeSlots, then the range's queue offset
  // [0..1] qext counts, [2] qsh count, [4] error flag, [6..7] next pixel (u64),
  // [8..9] extension rays traced (u64), [10..11] shadow rays traced (u64),
  // [12] / [13] read cursors of the queue-fed tracer (and the drain) over the
  // extension / shadow queue (each cleared by the kernel that runs before that
  // trace); the one-launch bounce over XCD segments: [16..23] each pixel
  // band's claim cursor (claims taken from the band), [24..31] / [32..39] the
  // segments' read cursors for bounces of even / odd parity
  uint32_t* counters;
  // Ray counts of the persistent kernels that tally per wave: kTallyLines
  // lines of 16 words (128 bytes), the extension rays in word 0, the shadow
  // rays in word 1; wfReport sums them.
  unsigned long long* tally;
  int* spill;  // traversal stack overflow of the trace kernel's threads
  uint32_t num_scenes;
  uint32_t num_slots;
  uint32_t num_ranges;  // slots rounded up to whole ranges of kRangeSlots
  uint64_t total_pixels;  // claim indices: the pixels, or 64 per 8x8 tile in `order`
  // Sample passes over one scene in longest-first order: claim index p is
  // pixel p % 64 (row by row) of 8x8 tile order[p / 64] (null: scan order).
  const uint32_t* order;
  uint32_t tiles_x;
};

constexpr int kWfCounterWords = 40;
constexpr uint32_t kTallyLines = 64;
static_assert(kWfCounterWords <= (int)kTallyLines, "the reset kernels clear one counter word per thread");

// A wave's ray counts at exit: workgroup b adds to tally line b mod
// kTallyLines, so the thousands of waves of a launch do not all add into the
// same two words (8192 device-scope atomics on one address took ~0.1 ms per
// launch, the whole cost of an empty sample pass).
template <typename R>
ZT_D void tallyRays(const WfShared<R>& w, uint64_t n_ext, uint64_t n_sh) {
  unsigned long long* t = w.tally + (size_t)(blockIdx.x % kTallyLines) * 16;
  if (n_ext) atomicAdd(t, (unsigned long long)n_ext);
  if (n_sh) atomicAdd(t + 1, (unsigned long long)n_sh);
}
constexpr uint32_t kSegs = 8;  // XCDs: workgroup b runs on XCD b mod 8

// First claim index of pixel band k of `total` (8 bands, sizes differing by
// at most one).
ZT_HD uint64_t pixelBand(uint32_t k, uint64_t total) {
  const uint64_t q = total / kSegs, r = total % kSegs;
  return k * q + (k < r ? k : r);
}

// Stream compaction of the slot flags, in slot order. A workgroup of
// kRangeThreads threads covers one range of kRangeSlots flags, 16 per
// thread (one 16-byte load).
constexpr uint32_t kRangeThreads = 256;
constexpr uint32_t kRangeSlots = kRangeThreads * 16;
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kMaxRanges = kScanThreads * 16;  // the scan's capacity: 2^26 slots

ZT_D uint32_t flagsSet(uint4 v) {  // flags are 0 or 1: one bit per set flag
  return (uint32_t)(__popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w));
}

// Inclusive prefix sum over the 64 lanes of a wave.
ZT_D uint32_t waveInclusive(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d);
    if (lane >= (uint32_t)d) v += u;
  }
  return v;
}

// Exclusive prefix sum over a workgroup of `kWaves` waves; *total gets the sum.
template <int kWaves>
ZT_D uint32_t blockExclusive(uint32_t v, uint32_t* lds, uint32_t* total) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t inc = waveInclusive(v);
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int k = 0; k < kWaves; ++k) {
    const uint32_t t = lds[k];
    if ((uint32_t)k < wave) before += t;
    all += t;
  }
  *total = all;
  return before + inc - v;
}

__global__ __launch_bounds__(kRangeThreads) void wfCount(const uint8_t* flags, uint32_t* ranges) {
  __shared__ uint32_t lds[kRangeThreads / 64];
  const uint4 v = ((const uint4*)flags)[(size_t)blockIdx.x * kRangeThreads + threadIdx.x];
  uint32_t total = 0;
  (void)blockExclusive<kRangeThreads / 64>(flagsSet(v), lds, &total);
  if (threadIdx.x == 0) ranges[blockIdx.x] = total;
}

// Exclusive scan of the ranges' counts in place (one workgroup); the total is
// the queue's length.
__global__ __launch_bounds__(kScanThreads) void wfScan(uint32_t* ranges, uint32_t n, uint32_t* count) {
  __shared__ uint32_t lds[kScanThreads / 64];
  const uint32_t per = (n + kScanThreads - 1) / kScanThreads;  // <= 16
  const uint32_t first = threadIdx.x * per;
  uint32_t c[16];
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t i = 0; i < 16; ++i) {
    c[i] = (i < per && first + i < n) ? ranges[first + i] : 0u;
    mine += c[i];
  }
  uint32_t total = 0;
  uint32_t off = blockExclusive<kScanThreads / 64>(mine, lds, &total);
#pragma unroll
  for (uint32_t i = 0; i < 16; ++i) {
    if (i < per && first + i < n) ranges[first + i] = off;
    off += c[i];
  }
  if (threadIdx.x == 0) *count = total;
}

// Writes every flagged slot to q at its range's offset plus its rank in the
// range, and clears the flags.
__global__ __launch_bounds__(kRangeThreads) void wfScatter(uint8_t* flags, const uint32_t* ranges, uint32_t* q) {
  __shared__ uint32_t lds[kRangeThreads / 64];
  const size_t word = (size_t)blockIdx.x * kRangeThreads + threadIdx.x;
  const uint4 v = ((const uint4*)flags)[word];
  const uint32_t c = flagsSet(v);
  uint32_t total = 0;
  uint32_t pos = ranges[blockIdx.x] + blockExclusive<kRangeThreads / 64>(c, lds, &total);
  if (c == 0) return;
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  const uint32_t slot0 = (uint32_t)word * 16u;
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if ((w[k >> 2] >> (8 * (k & 3))) & 1u) q[pos++] = slot0 + (uint32_t)k;
  ((uint4*)flags)[word] = make_uint4(0u, 0u, 0u, 0u);
}

template <typename R>
ZT_D uint32_t findScene(const WfShared<R>& w, uint64_t p) {
  uint32_t lo = 0, hi = w.num_scenes - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (w.scenes[mid].first_pixel <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The r-th pixel (row-major index) of a width x height image in 8x8-tile
// order: strips of 8 rows from the top, the tiles of a strip from the left,
// a tile's pixels row by row (tiles on the right and bottom edges narrower or
// shorter). Slots are filled in claim order, so the 64 lanes of a wave that
// take consecutive queue entries hold the pixels of one tile.
ZT_D uint32_t tileOrderPixel(uint32_t r, uint32_t width, uint32_t height) {
  const uint32_t y0 = r / (8 * width) * 8;
  const uint32_t h = min(8u, height - y0);
  const uint32_t i = r - y0 * width;  // within the strip
  const uint32_t x0 = i / (8 * h) * 8;
  const uint32_t tw = min(8u, width - x0);
  const uint32_t j = i - x0 * h;  // within the tile
  return (y0 + j / tw) * width + x0 + j % tw;
}

// Starts the pixel of global claim index p in slot `sl` (its stream,
// accumulator and first sample). Returns false only if the pixel completed at
// once (max_depth == 0).
template <typename R>
ZT_D bool startPixel(const WfShared<R>& w, WfSlot<R>& sl, uint64_t p) {
  const uint32_t si = findScene(w, p);
  const WfScene<R>& sc = w.scenes[si];
  const uint32_t local = tileOrderPixel((uint32_t)(p - sc.first_pixel), sc.pr.width, sc.pr.height);
  sl.scene = si;
  sl.pixel = local;
  sl.sample = 0;
  sl.sum = splat(R(0));
  sl.rng = rngForStream(sc.pr.seed, local);  // render.zig:135
  const uint32_t x = local % sc.pr.width, y = local / sc.pr.width;
  if (beginPath<R>(sc.view, sc.pr, x, y, sl.rng, sl.p) == kExtend) return true;
  // max_depth == 0: every sample is empty and the pixel is zero (render.zig:72).
  sc.out[3 * (uint64_t)local + 0] = R(0);
  sc.out[3 * (uint64_t)local + 1] = R(0);
  sc.out[3 * (uint64_t)local + 2] = R(0);
  return false;
}

// The sample passes' form of startPixel: the slot takes the pixel of claim
// index p with its stream and an empty accumulator, its first sample not yet
// begun. False if the claim index is no pixel.
template <typename R>
ZT_D bool assignPixel(const WfShared<R>& w, WfSlot<R>& sl, uint64_t p) {
  uint32_t si = 0, local = 0;
  if (w.order) {  // one scene; a claim index past the image's right or bottom edge is no pixel
    const WfScene<R>& sc = w.scenes[0];
    const uint32_t t = w.order[p >> 6], j = (uint32_t)p & 63u;
    const uint32_t x = (t % w.tiles_x) * 8 + (j & 7u), y = (t / w.tiles_x) * 8 + (j >> 3);
    if (x >= sc.pr.width || y >= sc.pr.height) return false;
    local = y * sc.pr.width + x;
  } else {
    si = findScene(w, p);
    const WfScene<R>& sc = w.scenes[si];
    local = tileOrderPixel((uint32_t)(p - sc.first_pixel), sc.pr.width, sc.pr.height);
  }
  const WfScene<R>& sc = w.scenes[si];
  sl.scene = si;
  sl.pixel = local;
  sl.sample = 0;
  sl.sum = splat(R(0));
  sl.rng = rngForStream(sc.pr.seed, local);  // render.zig:135
  return true;
}

// A slot between sample passes: its pixel, the next sample's index, the
// pixel's stream and its running sum (64 bytes).
ZT_D void loadPixelSlot(const WfSlots<float>& S, uint32_t i, WfSlot<float>& sl) {
  const uint4 wh = S.where[i], r01 = S.rng01[i], r23 = S.rng23[i];
  const float4 g = S.sum[i];
  sl.scene = wh.x;
  sl.pixel = wh.y;
  sl.sample = wh.z;
  sl.rng = Rng{lo64(r01), hi64(r01), lo64(r23), hi64(r23)};
  sl.sum = xyz<float>(g);
}
ZT_D void storePixelSlot(const WfSlots<float>& S, uint32_t i, const WfSlot<float>& sl) {
  S.where[i] = make_uint4(sl.scene, sl.pixel, sl.sample, 0u);
  S.rng01[i] = packRng(sl.rng.s0, sl.rng.s1);
  S.rng23[i] = packRng(sl.rng.s2, sl.rng.s3);
  S.sum[i] = pack4<float>(sl.sum, 0.0f);
}

// Claims pixels from the global counter until one has a ray to trace.
template <typename R>
ZT_D bool claimPixel(const WfShared<R>& w, WfSlot<R>& sl) {
  for (;;) {
    const uint64_t p = atomicAdd((unsigned long long*)&w.counters[6], 1ull);
    if (p >= w.total_pixels) {
      sl.scene = 0xFFFFFFFFu;
      return false;
    }
    if (startPixel(w, sl, p)) return true;
  }
}

// The one-launch bounce's claims: from the slot's own pixel band `home`
// first, then from the following bands once it is used up. `used_up` has a
// bit per band known to be fully claimed (read once per launch by
// bandsUsedUp, and set here when a claim comes back past the band's end), so
// a finished pixel costs no memory round trip once every band is claimed.
// kBegin = false (sample passes) assigns the pixel without starting its first
// sample.
template <typename R, bool kBegin = true>
ZT_D bool claimPixelBand(const WfShared<R>& w, WfSlot<R>& sl, uint32_t home, uint32_t& used_up) {
  for (uint32_t j = 0; j < kSegs; ++j) {
    const uint32_t b = (home + j) % kSegs;
    if ((used_up >> b) & 1u) continue;
    const uint64_t lo = pixelBand(b, w.total_pixels), len = pixelBand(b + 1, w.total_pixels) - lo;
    for (;;) {
      const uint32_t k = atomicAdd(&w.counters[16 + b], 1u);
      if (k >= len) {
        used_up |= 1u << b;
        break;
      }
      if constexpr (kBegin) {
        if (startPixel(w, sl, lo + k)) return true;
      } else {
        if (assignPixel(w, sl, lo + k)) return true;
      }
    }
  }
  sl.scene = 0xFFFFFFFFu;
  return false;
}

// The bands whose claim cursors had passed their end when the launch began
// (the cursors grow only by atomics of earlier launches here, so plain loads
// see them).
template <typename R>
ZT_D uint32_t bandsUsedUp(const WfShared<R>& w) {
  uint32_t m = 0;
  for (uint32_t b = 0; b < kSegs; ++b)
    if (w.counters[16 + b] >= pixelBand(b + 1, w.total_pixels) - pixelBand(b, w.total_pixels)) m |= 1u << b;
  return m;
}

// A path of the slot ended: accumulate it (render.zig:141-142) and start the
// pixel's next sample, or write the pixel (render.zig:144) and claim another
// (from the global counter, or from pixel band `home` when it is not
// negative). Returns true if the slot has an extension ray to trace.
template <typename R>
ZT_D bool finishPath(const WfShared<R>& w, WfSlot<R>& sl, int home = -1, uint32_t* used_up = nullptr) {
  const WfScene<R>& sc = w.scenes[sl.scene];
  if (isFinite(sl.p.radiance)) sl.sum = add(sl.sum, sl.p.radiance);
  if (++sl.sample < sc.pr.spp) {
    const uint32_t x = sl.pixel % sc.pr.width, y = sl.pixel / sc.pr.width;
    beginPath<R>(sc.view, sc.pr, x, y, sl.rng, sl.p);  // max_depth > 0 here: always an extension ray
    return true;
  }
  const R inv_spp = R(1) / R(sc.pr.spp);
  sc.out[3 * (uint64_t)sl.pixel + 0] = sl.sum.x * inv_spp;
  sc.out[3 * (uint64_t)sl.pixel + 1] = sl.sum.y * inv_spp;
  sc.out[3 * (uint64_t)sl.pixel + 2] = sl.sum.z * inv_spp;
  return home < 0 ? claimPixel(w, sl) : claimPixelBand(w, sl, (uint32_t)home, *used_up);
}

// Starts the first pixels in the slots and flags those with a ray for the
// first extension queue; clears the flags of the padding to whole ranges.
template <typename R>
__global__ void wfInit(WfShared<R> w) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < w.num_ranges * kRangeSlots; s += stride) {
    bool go = false;
    if (s < w.num_slots) {
      WfSlot<R> sl{};
      sl.scene = 0xFFFFFFFFu;
      if (s < w.total_pixels) go = startPixel(w, sl, s) || claimPixel(w, sl);
      storeSlot(w.slots, s, sl);
    }
    w.fext[s] = go ? 1 : 0;
    w.fsh[s] = 0;
  }
}

// wfInit for the one-launch bounce: slot j of segment k starts pixel j of
// band k (the band cursors were set past those pixels by wfResetSegments);
// slots beyond their band's length stay idle.
template <typename R>
__global__ void wfInitSegments(WfShared<R> w, uint32_t seg_slots) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < w.num_slots; s += stride) {
    const uint32_t k = s / seg_slots, j = s % seg_slots;
    const uint64_t lo = pixelBand(k, w.total_pixels), len = pixelBand(k + 1, w.total_pixels) - lo;
    WfSlot<R> sl{};
    sl.scene = 0xFFFFFFFFu;
    bool go = false;
    uint32_t used_up = 0;
    if (j < len) go = startPixel(w, sl, lo + j) || claimPixelBand(w, sl, k, used_up);
    storeSlot<R, false>(w.slots, s, sl);
    w.fext[s] = go ? 1 : 0;
  }
}

// wfInitSegments for the sample passes (fp32): the slots take their pixels
// without starting a sample.
__global__ void wfInitPasses(WfShared<float> w, uint32_t seg_slots) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < w.num_slots; s += stride) {
    const uint32_t k = s / seg_slots, j = s % seg_slots;
    const uint64_t lo = pixelBand(k, w.total_pixels), len = pixelBand(k + 1, w.total_pixels) - lo;
    bool go = false;
    if (j < len) {
      WfSlot<float> sl{};
      uint32_t used_up = 0;
      go = assignPixel(w, sl, lo + j) || claimPixelBand<float, false>(w, sl, k, used_up);
      if (go) storePixelSlot(w.slots, s, sl);
    }
    w.fext[s] = go ? 1 : 0;
  }
}

// One workgroup of kTallyLines (64) threads.
__global__ void wfResetSegments(uint32_t* counters, unsigned long long* tally, uint64_t total, uint32_t seg_slots) {
  const uint32_t i = threadIdx.x;
  tally[i * 16] = 0;
  tally[i * 16 + 1] = 0;
  if (i < kWfCounterWords) counters[i] = 0;
  __syncthreads();
  if (i < kSegs) {
    const uint64_t len = pixelBand(i + 1, total) - pixelBand(i, total);
    counters[16 + i] = (uint32_t)(len < seg_slots ? len : seg_slots);
  }
}

// Closest-hit (extension queue) or any-hit (shadow queue) query of every
// queued ray, one ray per lane over a grid-stride loop.
template <typename R>
__global__ __launch_bounds__(256) void wfTrace(WfShared<R> w, int cur, int shadow, int cap) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack(lds, cap, w.spill);
  const uint32_t n = shadow ? w.counters[2] : w.counters[cur];
  if (blockIdx.x == 0 && threadIdx.x == 0) *(unsigned long long*)&w.counters[shadow ? 10 : 8] += n;
  const uint32_t* q = shadow ? w.qsh : w.qext[cur];
  const uint32_t stride = gridDim.x * blockDim.x;
  int overflow = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t slot = q[i];
    const SceneView<R>& v = w.scenes[w.slots.where[slot].x].view;
    const V3<R> ro = xyz<R>(w.slots.ro[slot]);
    const int4 m = w.slots.meta[slot];
    HitRec<R> h;
    if (shadow) {
      const V4<R> c = w.slots.shd[slot];
      h = query<R>(v, stk, ro, xyz<R>(c), R(kRayEps), c.w, m.x, m.y, true, &overflow);
    } else {
      h = query<R>(v, stk, ro, xyz<R>(w.slots.rd[slot]), R(kRayEps), R(INFINITY), m.x, -1, false, &overflow);
    }
    if constexpr (K<R>::exact) {
      if (!shadow) w.hits_d[slot] = h.t;
    } else {
      if (!shadow) w.hits[slot].x = h.t;
    }
    w.hits[slot].y = __int_as_float(h.prim);
  }
  if (overflow) atomicOr((int*)&w.counters[4], 1);
}

// The scenes of a trace launch, for the wave-level traversal passes
// (dev_wave.hpp): one scene in kernel arguments, or the batch's table.
struct WfOne {
  static constexpr bool kSingle = true;
  SceneView<float> sc;
  Params pr;
  ZT_D const SceneView<float>& view(uint32_t) const { return sc; }
  ZT_D const Params& params(uint32_t) const { return pr; }
};
struct WfMany {
  static constexpr bool kSingle = false;
  const WfScene<float>* scenes;
  ZT_D const SceneView<float>& view(uint32_t si) const { return scenes[si].view; }
  ZT_D const Params& params(uint32_t si) const { return scenes[si].pr; }
};

// The same queries as wfTrace (fp32, four-wide trees), traced by persistent
// one-wave workgroups under the megakernel's wave-uniform node and leaf
// passes: lanes without a ray take the next queue entries (one atomic per
// wave on the queue's cursor), and the wave stops its passes to take more
// once `refill` lanes have finished, until the queue is exhausted. (Cutting
// the slot-ordered queue into eight parts, one per XCD, each XCD's waves
// taking from their own part first, was measured 1.44x slower on the 1080p
// mesh field: 424 vs 294 ms per frame.)
template <bool Q, typename Src>
ZT_D void wfTraceWaveBody(const Src& src, const WfShared<float>& w, int cur, int shadow, int cap, uint32_t refill) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack<64>(lds, cap, w.spill);  // workgroups of one wave
  unsigned long long* keys = (unsigned long long*)(lds + stackBytes(cap, 64) / sizeof(int));
  uint8_t* owner = (uint8_t*)(keys + 64);
  const uint32_t lane = threadIdx.x;
  const bool any = shadow != 0;
  const uint32_t n = any ? w.counters[2] : w.counters[cur];
  if (blockIdx.x == 0 && lane == 0) *(unsigned long long*)&w.counters[any ? 10 : 8] += n;
  const uint32_t* q = any ? w.qsh : w.qext[cur];
  uint32_t* cursor = &w.counters[any ? 13 : 12];
  bool tracing = false, drained = false;
  uint32_t slot = 0, si = 0;
  Query4 qy{};
  Trav4 tr;
  int overflow = 0;
  for (;;) {
    if (!drained) {
      const uint64_t need = __ballot(!tracing);
      if (need) {
        const uint32_t k = (uint32_t)__popcll(need);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(cursor, k);
        base = __shfl(base, 0);
        if (base + k >= n) drained = true;
        const uint32_t idx = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
        if (!tracing && idx < n) {
          slot = q[idx];
          si = w.slots.where[slot].x;
          const float4 o4 = w.slots.ro[slot], d4 = any ? w.slots.shd[slot] : w.slots.rd[slot];
          const int4 m = w.slots.meta[slot];
          qy = Query4{xyz<float>(o4), xyz<float>(d4), (float)kRayEps, m.x, any ? m.y : -1, any};
          tr.start(qy, any ? d4.w : INFINITY);
          tracing = true;
        }
      }
    }
    if (__ballot(tracing) == 0) {
      if (drained) break;
      continue;
    }
    for (;;) {
      const bool at_leaf = tracing && tr.atLeaf();
      const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
      const uint32_t n_node = (uint32_t)__popcll(__ballot(tracing && !at_leaf));
      if (n_leaf + n_node == 0) break;
      if (!drained && (uint32_t)__popcll(__ballot(!tracing)) >= refill) break;
      if (wavePass<Q>(src, si, tr, qy, tracing, at_leaf, n_leaf, n_node, stk, &overflow, keys, owner)) {
        const HitRec<float> h = tr.finish(src.view(si), qy);
        if (!any) w.hits[slot].x = h.t;
        w.hits[slot].y = __int_as_float(h.prim);
        tracing = false;
      }
    }
  }
  if (overflow) atomicOr((int*)&w.counters[4], 1);
}

template <bool Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfTraceWaveOne(
    WfOne src, WfShared<float> w, int cur, int shadow, int cap, uint32_t refill) {
  wfTraceWaveBody<Q>(src, w, cur, shadow, cap, refill);
}
template <bool Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfTraceWaveMany(
    WfMany src, WfShared<float> w, int cur, int shadow, int cap, uint32_t refill) {
  wfTraceWaveBody<Q>(src, w, cur, shadow, cap, refill);
}

// The two-launch bounce (fp32, four-wide trees): the queue-fed wave tracer
// above, with each path's bounce applied by its lane in the same kernel.
// Lanes whose query ended wait, holding the hit, until the wave stops its
// passes to take more rays; then they load their slots together, apply the
// bounce, store the slots and flag them for the next queues.
//   extension launch: a finished walk is shaded (render.zig:73-99) and the
//     slot flagged for the shadow queue if a next-event ray is pending, for
//     the next extension queue if the path continues or its pixel starts
//     another sample;
//   shadow launch: a finished any-hit query is applied (render.zig:51), and
//     a path that ended at that bounce moves on to its pixel's next sample or
//     to a new pixel.
template <bool Q, bool kShadowPass, typename Src>
ZT_D void wfBounceBody(const Src& src, const WfShared<float>& w, int cur, int cap, uint32_t refill) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack<64>(lds, cap, w.spill);  // workgroups of one wave
  unsigned long long* keys = (unsigned long long*)(lds + stackBytes(cap, 64) / sizeof(int));
  uint8_t* owner = (uint8_t*)(keys + 64);
  const uint32_t lane = threadIdx.x;
  const uint32_t n = kShadowPass ? w.counters[2] : w.counters[cur];
  if (blockIdx.x == 0 && lane == 0) {
    *(unsigned long long*)&w.counters[kShadowPass ? 10 : 8] += n;
    w.counters[kShadowPass ? 12 : 13] = 0;  // the cursor of the launch after this one
  }
  const uint32_t* q = kShadowPass ? w.qsh : w.qext[cur];
  uint32_t* cursor = &w.counters[kShadowPass ? 13 : 12];
  bool tracing = false, pending = false, drained = false;
  bool ends = false;  // shadow launch: the path ends at this bounce
  uint32_t slot = 0, si = 0;
  HitRec<float> hit{0.0f, -1};
  Query4 qy{};
  Trav4 tr;
  int overflow = 0;
  for (;;) {
    if (pending) {
      pending = false;
      if constexpr (kShadowPass) {
        if (ends) {  // the whole slot moves on
          WfSlot<float> sl;
          loadSlot(w.slots, slot, sl);
          (void)afterShadow<float>(hit, sl.p);
          const bool to_ext = finishPath(w, sl);
          storeSlot(w.slots, slot, sl);
          if (to_ext) w.fext[slot] = 1;
        } else if (hit.prim < 0) {  // only the radiance changes
          Path<float> p;
          p.radiance = xyz<float>(w.slots.rad[slot]);
          p.sh_contrib = xyz<float>(w.slots.shc[slot]);
          p.terminated = 0;
          (void)afterShadow<float>(hit, p);
          w.slots.rad[slot] = pack4<float>(p.radiance, 0.0f);
        }
      } else {
        WfSlot<float> sl;
        loadSlot(w.slots, slot, sl);
        const int ph = shade<float>(src.view(si), w.scenes[Src::kSingle ? 0 : si].pr, hit, sl.rng, sl.p);
        bool to_ext = ph == kExtend || (ph == kShadow && !sl.p.terminated);
        if (ph == kFinish) to_ext = finishPath(w, sl);
        storeSlot(w.slots, slot, sl);
        if (ph == kShadow) w.fsh[slot] = 1;
        if (to_ext) w.fext[slot] = 1;
      }
    }
    if (!drained) {
      const uint64_t need = __ballot(!tracing);
      if (need) {
        const uint32_t k = (uint32_t)__popcll(need);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(cursor, k);
        base = __shfl(base, 0);
        if (base + k >= n) drained = true;
        const uint32_t idx = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
        if (!tracing && idx < n) {
          slot = q[idx];
          si = w.slots.where[slot].x;
          const float4 o4 = w.slots.ro[slot], d4 = kShadowPass ? w.slots.shd[slot] : w.slots.rd[slot];
          const int4 m = w.slots.meta[slot];
          ends = m.w != 0;
          qy = Query4{xyz<float>(o4), xyz<float>(d4), (float)kRayEps, m.x, kShadowPass ? m.y : -1, kShadowPass};
          tr.start(qy, kShadowPass ? d4.w : INFINITY);
          tracing = true;
        }
      }
    }
    if (__ballot(tracing) == 0) {
      if (drained) break;
      continue;
    }
    for (;;) {
      const bool at_leaf = tracing && tr.atLeaf();
      const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
      const uint32_t n_node = (uint32_t)__popcll(__ballot(tracing && !at_leaf));
      if (n_leaf + n_node == 0) break;
      if (!drained && (uint32_t)__popcll(__ballot(!tracing)) >= refill) break;
      if (wavePass<Q>(src, si, tr, qy, tracing, at_leaf, n_leaf, n_node, stk, &overflow, keys, owner)) {
        hit = tr.finish(src.view(si), qy);
        tracing = false;
        pending = true;
      }
    }
  }
  if (overflow) atomicOr((int*)&w.counters[4], 1);
}

template <bool Q, bool kShadowPass>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void wfBounceOne(
    WfOne src, WfShared<float> w, int cur, int cap, uint32_t refill) {
  wfBounceBody<Q, kShadowPass>(src, w, cur, cap, refill);
}
template <bool Q, bool kShadowPass>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void wfBounceMany(
    WfMany src, WfShared<float> w, int cur, int cap, uint32_t refill) {
  wfBounceBody<Q, kShadowPass>(src, w, cur, cap, refill);
}

// The one-launch bounce (fp32, four-wide trees): every queued path advances
// by one bounce inside one persistent launch. A lane takes a slot, traces its
// extension ray, shades the hit (render.zig:73-99), traces the next-event ray
// if one is pending and applies it (render.zig:51), and stores the slot
// flagged for the next extension queue -- with the path's next ray, its
// pixel's next sample, or a newly claimed pixel -- or leaves it idle. All
// three steps run under the wave-uniform passes, lanes served together once
// `refill` of them wait. The pending next-event ray never leaves the lane, so
// the slot traffic per bounce is one load and one store without it.
//
// XCD locality: the slot pool is 8 segments of `rps` compaction ranges each;
// segment k holds the pixels of band k (its slots claim from that band
// first), and its queue entries are contiguous in the slot-ordered queue
// (from ranges[k * rps]). Workgroup b (on XCD b mod 8) takes entries from
// segment b mod 8, then from the following segments once that is used up,
// so each XCD's L2 serves the rays of one band of the image.
template <bool Q, typename Src>
ZT_D void wfStepBody(const Src& src, const WfShared<float>& w, int cur, int cap, uint32_t refill, uint32_t rps) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack<64>(lds, cap, w.spill);  // workgroups of one wave
  unsigned long long* keys = (unsigned long long*)(lds + stackBytes(cap, 64) / sizeof(int));
  uint8_t* owner = (uint8_t*)(keys + 64);
  const uint32_t lane = threadIdx.x;
  const uint32_t home = blockIdx.x % kSegs;
  const uint32_t n = w.counters[cur];
  uint32_t* cursors = w.counters + 24 + 8 * cur;
  if (blockIdx.x == 0 && lane < kSegs) w.counters[24 + 8 * (1 - cur) + lane] = 0;  // the next bounce's
  const uint32_t* q = w.qext[cur];
  // Wave-uniform queue position: the segment being taken from and its entries.
  uint32_t seg_k = 0, lo = 0, hi = 0;
  auto bounds = [&](uint32_t s) {
    lo = w.ranges[s * rps];
    hi = s + 1 < kSegs ? w.ranges[(s + 1) * rps] : n;
  };
  bool done = n == 0;
  if (!done) bounds(home);
  uint32_t used_up = bandsUsedUp(w);
  bool has = false, tracing = false, sh = false;
  uint32_t slot = 0, si = 0;
  WfSlot<float> sl{};
  Trav4 tr;
  int overflow = 0;
  uint64_t n_ext = 0, n_sh = 0;  // rays started by the wave (wave-uniform)
  for (;;) {
    // Serve the lanes whose query ended.
    const bool serve = has && !tracing;
    if (serve) {
      const HitRec<float> hit{tr.closest, tr.best};  // what tr.finish left
      const int ph = sh ? afterShadow<float>(hit, sl.p) : shade<float>(src.view(si), src.params(si), hit, sl.rng, sl.p);
      if (ph == kShadow) {
        sh = true;
        tr.start(Query4{sl.p.ro, sl.p.sh_d, (float)kRayEps, sl.p.self, sl.p.sh_excl, true}, sl.p.sh_tmax);
        tracing = true;
      } else {
        const bool to_ext = ph == kExtend || finishPath(w, sl, (int)home, &used_up);
        if (sl.scene != 0xFFFFFFFFu) storeSlot<float, false>(w.slots, slot, sl);
        if (to_ext) w.fext[slot] = 1;
        has = false;
      }
    }
    n_sh += (uint64_t)__popcll(__ballot(serve && tracing));
    // Slots for the free lanes, from this XCD's segment first.
    while (!done) {
      const uint64_t need = __ballot(!has);
      if (need == 0) break;
      const uint32_t k = (uint32_t)__popcll(need);
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&cursors[(home + seg_k) % kSegs], k);
      base = __shfl(base, 0);
      const uint32_t idx = lo + base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
      const bool take = !has && idx < hi;
      if (take) {
        slot = q[idx];
        loadSlot<float, false>(w.slots, slot, sl);
        si = Src::kSingle ? 0u : sl.scene;
        sh = false;
        tr.start(Query4{sl.p.ro, sl.p.rd, (float)kRayEps, sl.p.self, -1, false}, INFINITY);
        has = tracing = true;
      }
      n_ext += (uint64_t)__popcll(__ballot(take));
      if (lo + base + k >= hi) {
        if (++seg_k == kSegs) done = true;
        else bounds((home + seg_k) % kSegs);
      }
    }
    if (__ballot(tracing) == 0) {
      if (done) break;
      continue;
    }
    for (;;) {
      const Query4 qy{sl.p.ro, sh ? sl.p.sh_d : sl.p.rd, (float)kRayEps, sl.p.self, sh ? sl.p.sh_excl : -1, sh};
      const bool at_leaf = tracing && tr.atLeaf();
      const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
      const uint32_t n_node = (uint32_t)__popcll(__ballot(tracing && !at_leaf));
      if (n_leaf + n_node == 0) break;
      if ((uint32_t)__popcll(__ballot(!tracing && (has || !done))) >= refill) break;
      if (wavePass<Q>(src, si, tr, qy, tracing, at_leaf, n_leaf, n_node, stk, &overflow, keys, owner)) {
        (void)tr.finish(src.view(si), qy);
        tracing = false;
      }
    }
  }
  if (lane == 0) tallyRays(w, n_ext, n_sh);
  if (overflow) atomicOr((int*)&w.counters[4], 1);
}

template <bool Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfStepOne(
    WfOne src, WfShared<float> w, int cur, int cap, uint32_t refill, uint32_t rps) {
  wfStepBody<Q>(src, w, cur, cap, refill, rps);
}
template <bool Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfStepMany(
    WfMany src, WfShared<float> w, int cur, int cap, uint32_t refill, uint32_t rps) {
  wfStepBody<Q>(src, w, cur, cap, refill, rps);
}

// Sample passes (fp32, four-wide trees): the queue holds pixels, and a take
// runs one whole sample of the pixel -- the camera ray (render.zig:136-139),
// every bounce with its next-event ray, the accumulation (render.zig:141-142)
// -- under the wave-uniform passes, lanes served together once `refill` of
// them wait; then the slot goes back on the next pass's queue with its
// stream, sum and sample index, or the finished pixel is written
// (render.zig:144) and the slot claims a new one for the next pass. Every
// pixel takes exactly spp passes, so with a pool as large as the frame no
// pixel trails the others; a pixel's samples still run in order from its own
// stream, so the image is the megakernel's bits. Segments, bands and queue
// cursors as in the one-launch bounce.
template <bool Q, bool kClaims, typename Src>
ZT_D void wfPassBody(const Src& src, const WfShared<float>& w, int cur, int cap, uint32_t refill, uint32_t rps) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack<64>(lds, cap, w.spill);  // workgroups of one wave
  unsigned long long* keys = (unsigned long long*)(lds + stackBytes(cap, 64) / sizeof(int));
  uint8_t* owner = (uint8_t*)(keys + 64);
  const uint32_t lane = threadIdx.x;
  const uint32_t home = blockIdx.x % kSegs;
  const uint32_t n = w.counters[cur];
  uint32_t* cursors = w.counters + 24 + 8 * cur;
  if (blockIdx.x == 0 && lane < kSegs) w.counters[24 + 8 * (1 - cur) + lane] = 0;  // the next pass's
  const uint32_t* q = w.qext[cur];
  uint32_t seg_k = 0, lo = 0, hi = 0;
  auto bounds = [&](uint32_t s) {
    lo = w.ranges[s * rps];
    hi = s + 1 < kSegs ? w.ranges[(s + 1) * rps] : n;
  };
  bool done = n == 0;
  if (!done) bounds(home);
  uint32_t used_up = kClaims ? bandsUsedUp(w) : 0u;
  bool has = false, tracing = false, sh = false;
  uint32_t slot = 0, si = 0;
  WfSlot<float> sl{};
  Trav4 tr;
  int overflow = 0;
  uint64_t n_ext = 0, n_sh = 0;  // rays started by the wave (wave-uniform)
  // The lane's sample ended: accumulate it and put the slot on the next
  // pass's queue, or write the pixel and claim another.
  auto endSample = [&]() {
    if (isFinite(sl.p.radiance)) sl.sum = add(sl.sum, sl.p.radiance);  // render.zig:141-142
    const Params& pr = src.params(si);
    bool again = ++sl.sample < pr.spp;
    if (!again) {
      float* out = w.scenes[si].out;
      const float inv_spp = 1.0f / (float)pr.spp;
      out[3 * (uint64_t)sl.pixel + 0] = sl.sum.x * inv_spp;
      out[3 * (uint64_t)sl.pixel + 1] = sl.sum.y * inv_spp;
      out[3 * (uint64_t)sl.pixel + 2] = sl.sum.z * inv_spp;
      if constexpr (kClaims) again = claimPixelBand<float, false>(w, sl, home, used_up);
    }
    if (again) {
      storePixelSlot(w.slots, slot, sl);
      w.fext[slot] = 1;
    }
    has = false;
  };
  for (;;) {
    // Serve the lanes whose query ended.
    const bool serve = has && !tracing;
    if (serve) {
      const HitRec<float> hit{tr.closest, tr.best};  // what tr.finish left
      const int ph = sh ? afterShadow<float>(hit, sl.p) : shade<float>(src.view(si), src.params(si), hit, sl.rng, sl.p);
      if (ph == kFinish) {
        endSample();
      } else {
        sh = ph == kShadow;
        tr.start(Query4{sl.p.ro, sh ? sl.p.sh_d : sl.p.rd, (float)kRayEps, sl.p.self, sh ? sl.p.sh_excl : -1, sh},
                 sh ? sl.p.sh_tmax : INFINITY);
        tracing = true;
      }
    }
    {
      const uint64_t started = __ballot(serve && tracing), shadow = __ballot(serve && tracing && sh);
      n_ext += (uint64_t)__popcll(started & ~shadow);
      n_sh += (uint64_t)__popcll(shadow);
    }
    // Pixels for the free lanes, from this XCD's segment first; each starts
    // its next sample at once.
    while (!done) {
      const uint64_t need = __ballot(!has);
      if (need == 0) break;
      const uint32_t k = (uint32_t)__popcll(need);
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&cursors[(home + seg_k) % kSegs], k);
      base = __shfl(base, 0);
      const uint32_t idx = lo + base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
      bool take = !has && idx < hi;
      if (take) {
        slot = q[idx];
        loadPixelSlot(w.slots, slot, sl);
        si = Src::kSingle ? 0u : sl.scene;
        const Params& pr = src.params(si);
        has = true;
        if (beginPath<float>(src.view(si), pr, sl.pixel % pr.width, sl.pixel / pr.width, sl.rng, sl.p) == kExtend) {
          sh = false;
          tr.start(Query4{sl.p.ro, sl.p.rd, (float)kRayEps, sl.p.self, -1, false}, INFINITY);
          tracing = true;
        } else {  // max_depth == 0: an empty sample
          endSample();
          take = false;
        }
      }
      n_ext += (uint64_t)__popcll(__ballot(take));
      if (lo + base + k >= hi) {
        if (++seg_k == kSegs) done = true;
        else bounds((home + seg_k) % kSegs);
      }
    }
    if (__ballot(tracing) == 0) {
      if (done) break;
      continue;
    }
    for (;;) {
      const Query4 qy{sl.p.ro, sh ? sl.p.sh_d : sl.p.rd, (float)kRayEps, sl.p.self, sh ? sl.p.sh_excl : -1, sh};
      const bool at_leaf = tracing && tr.atLeaf();
      const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
      const uint32_t n_node = (uint32_t)__popcll(__ballot(tracing && !at_leaf));
      if (n_leaf + n_node == 0) break;
      if ((uint32_t)__popcll(__ballot(!tracing && (has || !done))) >= refill) break;
      if (wavePass<Q>(src, si, tr, qy, tracing, at_leaf, n_leaf, n_node, stk, &overflow, keys, owner)) {
        (void)tr.finish(src.view(si), qy);
        tracing = false;
      }
    }
  }
  if (lane == 0) tallyRays(w, n_ext, n_sh);
  if (overflow) atomicOr((int*)&w.counters[4], 1);
}

template <bool Q, bool kClaims>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfPassOne(
    WfOne src, WfShared<float> w, int cur, int cap, uint32_t refill, uint32_t rps) {
  wfPassBody<Q, kClaims>(src, w, cur, cap, refill, rps);
}
template <bool Q, bool kClaims>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfPassMany(
    WfMany src, WfShared<float> w, int cur, int cap, uint32_t refill, uint32_t rps) {
  wfPassBody<Q, kClaims>(src, w, cur, cap, refill, rps);
}

// The tail of a frame: once every pixel has been claimed and the live paths
// no longer fill the GPU, each remaining slot -- whose next extension ray is
// pending after a full bounce iteration -- is finished by a persistent
// wave-scheduled kernel: a lane takes a slot from `list` (the live extension
// queue), continues its path and its pixel's remaining samples from the
// slot's state with the same device functions, writes the pixel, and takes
// the next slot. The bounce loop then stops; each pixel still runs its
// samples in order from its own stream, so the image is the same bits.
template <bool S>
struct DrainSrc {
  static constexpr bool kSingle = S;
  const WfScene<float>* scenes;
  ZT_D const WfScene<float>& scene(uint32_t si) const { return scenes[S ? 0 : si]; }
  ZT_D const SceneView<float>& view(uint32_t si) const { return scene(si).view; }
};

// kPass: the slots hold pixels between sample passes (the next sample not yet
// begun) instead of paths with a pending extension ray.
template <bool Q, bool kSingle, bool kPass>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void wfDrain(
    WfShared<float> w, int cap, uint32_t refill) {
  constexpr int kResolve = 4;  // a traversal result is waiting to be applied
  extern __shared__ int lds[];
  const LaneStack stk = laneStack<64>(lds, cap, w.spill);  // workgroups of one wave
  unsigned long long* keys = (unsigned long long*)(lds + stackBytes(cap, 64) / sizeof(int));
  uint8_t* owner = (uint8_t*)(keys + 64);
  const DrainSrc<kSingle> src{w.scenes};
  const uint32_t* list = w.qext[0];
  const uint32_t n = w.counters[0];  // the live extension queue after an even number of bounces
  uint32_t* cursor = &w.counters[12];
  const uint32_t lane = threadIdx.x;
  bool done = false;  // the list is used up (wave-uniform)
  bool has = false, tracing = false, sh = false;
  uint32_t si = 0, pixel = 0, sample = 0;
  int phase = kExtend;
  Rng rng{};
  V3<float> sum = splat(0.0f);
  Path<float> st;
  Trav4 tr;
  int overflow = 0;
  uint64_t n_ext = 0, n_sh = 0;
  for (;;) {
    if (!done) {
      const uint64_t need = __ballot(!has);
      if (need) {
        const uint32_t k = (uint32_t)__popcll(need);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(cursor, k);
        base = __shfl(base, 0);
        if (base + k >= n) done = true;
        const uint32_t idx = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
        if (!has && idx < n) {
          WfSlot<float> sl;
          if constexpr (kPass) loadPixelSlot(w.slots, list[idx], sl);
          else loadSlot(w.slots, list[idx], sl);
          si = sl.scene;
          pixel = sl.pixel;
          sample = sl.sample;
          rng = sl.rng;
          sum = sl.sum;
          st = sl.p;
          phase = kPass ? (int)kCamera : (int)kExtend;
          has = true;
        }
      }
    }
    const bool idle = has && !tracing;
    if (idle) {
      const WfScene<float>& sc = src.scene(si);
      for (;;) {
        if (phase == kResolve) {
          const HitRec<float> hit{tr.closest, tr.best};
          phase = sh ? afterShadow<float>(hit, st) : shade<float>(sc.view, sc.pr, hit, rng, st);
        }
        if (phase == kFinish) {
          if (isFinite(st.radiance)) sum = add(sum, st.radiance);  // render.zig:141-142
          ++sample;
          phase = kCamera;
        }
        if (phase == kCamera) {
          if (sample == sc.pr.spp) {
            const float inv_spp = 1.0f / (float)sc.pr.spp;
            sc.out[3 * (uint64_t)pixel + 0] = sum.x * inv_spp;
            sc.out[3 * (uint64_t)pixel + 1] = sum.y * inv_spp;
            sc.out[3 * (uint64_t)pixel + 2] = sum.z * inv_spp;
            has = false;
            break;
          }
          phase = beginPath<float>(sc.view, sc.pr, pixel % sc.pr.width, pixel / sc.pr.width, rng, st);
          continue;  // max_depth > 0 here: an extension ray follows
        }
        sh = phase == kShadow;
        tr.start(Query4{st.ro, sh ? st.sh_d : st.rd, (float)kRayEps, st.self, sh ? st.sh_excl : -1, sh},
                 sh ? st.sh_tmax : INFINITY);
        tracing = true;
        break;
      }
    }
    {
      const uint64_t started = __ballot(idle && tracing), shadow = __ballot(idle && tracing && sh);
      n_ext += (uint64_t)__popcll(started & ~shadow);
      n_sh += (uint64_t)__popcll(shadow);
    }
    if (__ballot(tracing) == 0) {
      if (done) break;
      continue;
    }
    for (;;) {
      const Query4 q{st.ro, sh ? st.sh_d : st.rd, (float)kRayEps, st.self, sh ? st.sh_excl : -1, sh};
      const bool at_leaf = tracing && tr.atLeaf();
      const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
      const uint32_t n_node = (uint32_t)__popcll(__ballot(tracing && !at_leaf));
      if (n_leaf + n_node == 0) break;
      if ((uint32_t)__popcll(__ballot(!tracing && (has || !done))) >= refill) break;
      if (wavePass<Q>(src, si, tr, q, tracing, at_leaf, n_leaf, n_node, stk, &overflow, keys, owner)) {
        (void)tr.finish(src.view(si), q);
        tracing = false;
        phase = kResolve;
      }
    }
  }
  if (lane == 0) tallyRays(w, n_ext, n_sh);
  if (overflow) atomicOr((int*)&w.counters[4], 1);
}

template <typename R>
__global__ __launch_bounds__(256) void wfShade(WfShared<R> w, int cur) {
  const uint32_t n = w.counters[cur];
  if (blockIdx.x == 0 && threadIdx.x == 0) w.counters[13] = 0;  // the shadow trace's cursor
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t s = w.qext[cur][i];
    WfSlot<R> sl;
    loadSlot(w.slots, s, sl);
    const WfScene<R>& sc = w.scenes[sl.scene];
    HitRec<R> h;
    if constexpr (K<R>::exact) h.t = w.hits_d[s];
    else h.t = w.hits[s].x;
    h.prim = __float_as_int(w.hits[s].y);
    const int ph = shade<R>(sc.view, sc.pr, h, sl.rng, sl.p);
    bool to_ext = ph == kExtend || (ph == kShadow && !sl.p.terminated);
    if (ph == kFinish) to_ext = finishPath(w, sl);
    storeSlot(w.slots, s, sl);
    if (ph == kShadow) w.fsh[s] = 1;
    if (to_ext) w.fext[s] = 1;
  }
}

// Applies each shadow ray's outcome (render.zig:51) and finishes the paths
// that ended at that bounce.
template <typename R>
__global__ __launch_bounds__(256) void wfResolve(WfShared<R> w) {
  const uint32_t n = w.counters[2];
  if (blockIdx.x == 0 && threadIdx.x == 0) w.counters[12] = 0;  // the next extension trace's cursor
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t s = w.qsh[i];
    const HitRec<R> h{R(0), __float_as_int(w.hits[s].y)};
    if (w.slots.meta[s].w) {  // the path ends at this bounce: the whole slot moves on
      WfSlot<R> sl;
      loadSlot(w.slots, s, sl);
      (void)afterShadow<R>(h, sl.p);
      const bool to_ext = finishPath(w, sl);
      storeSlot(w.slots, s, sl);
      if (to_ext) w.fext[s] = 1;
    } else if (h.prim < 0) {  // only the radiance changes (afterShadow, render.zig:51)
      Path<R> p;
      p.radiance = xyz<R>(w.slots.rad[s]);
      p.sh_contrib = xyz<R>(w.slots.shc[s]);
      p.terminated = 0;
      (void)afterShadow<R>(h, p);
      w.slots.rad[s] = pack4<R>(p.radiance, R(0));
    }
  }
}

// Adds the loop's ray counts (the counter words and the tally lines) and
// error flag to the launch status words. One workgroup of kTallyLines threads.
__global__ void wfReport(const uint32_t* counters, const unsigned long long* tally, int* status) {
  __shared__ unsigned long long sum[2][kTallyLines];
  const uint32_t i = threadIdx.x;
  sum[0][i] = tally[i * 16];
  sum[1][i] = tally[i * 16 + 1];
  __syncthreads();
  if (i == 0) {
    unsigned long long e = *(const unsigned long long*)&counters[8], sh = *(const unsigned long long*)&counters[10];
    for (uint32_t k = 0; k < kTallyLines; ++k) {
      e += sum[0][k];
      sh += sum[1][k];
    }
    addRayCounts(status, e, sh);
    if (counters[4]) atomicOr(status, 1);
  }
}

// One workgroup of kTallyLines threads.
__global__ void wfReset(uint32_t* counters, unsigned long long* tally, uint64_t first_free_pixel) {
  const uint32_t i = threadIdx.x;
  tally[i * 16] = 0;
  tally[i * 16 + 1] = 0;
  if (i == 0) {
    for (int k = 0; k < kWfCounterWords; ++k) counters[k] = 0;
    *(unsigned long long*)&counters[6] = first_free_pixel;
  }
}

template <typename R>
static int gridFor(const void* kernel, uint32_t threads, size_t lds) {
  return (int)residentBlocks(kernel, threads, lds);
}

template <typename R>
void runWavefront(zt_ctx* ctx, const std::vector<WfScene<R>>& scenes, const zt_gpu_tuning& tu) {
  uint64_t total = 0;
  for (const auto& s : scenes) total += (uint64_t)s.pr.width * s.pr.height;
  if (total == 0) return;
  uint32_t threads = tu.block_size ? tu.block_size : 256;
  if (threads < 64 || threads > 256 || threads % 64)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "wavefront block size must be 64, 128, 192 or 256");
  // fp32 scenes that all have four-wide trees are traced under the
  // wave-uniform passes (schedule 1 asks for one ray per lane instead):
  // schedule 2 by the queue-fed tracer with shading and shadow resolution in
  // kernels of their own, 3 with the tracer shading each hit itself (two
  // launches per bounce), 4 with the whole bounce in one launch over XCD
  // segments, 5 (the default) in sample passes (one launch advances every
  // queued pixel by a whole sample).
  if (tu.schedule > 5) throw Error(ZT_ERR_INVALID_ARGUMENT, "wavefront schedule must be 0..5");
  bool wave_trace = false, quant = false, fused = false, step = false;
  if constexpr (!K<R>::exact) {
    wave_trace = tu.schedule != 1;
    quant = scenes[0].view.nodes4q != nullptr;
    for (const auto& sc : scenes) {
      wave_trace = wave_trace && sc.view.nodes4 != nullptr;
      if ((sc.view.nodes4q != nullptr) != quant) throw Error(ZT_ERR_INTERNAL, "mixed node formats in one launch");
    }
  }
  const uint32_t sched = tu.schedule ? tu.schedule : 5;  // fp32 four-wide default: sample passes
  fused = wave_trace && sched == 3;
  step = wave_trace && sched == 4;
  const bool pass = wave_trace && sched == 5;
  const bool segmented = step || pass;  // slots in XCD segments, pixels in bands
  if (tu.tile_order > 2) throw Error(ZT_ERR_INVALID_ARGUMENT, "tile_order must be 0, 1 or 2");
  // Sample passes over one scene take its pixels longest-first (tile_order 0:
  // from 16 samples per pixel, so the one-sample cost probe adds at most
  // 1/16; 1: scan order; 2: always): the slots of each band hold its most
  // expensive 8x8 tiles first, so every pass ends on cheap pixels.
  const bool ordered = pass && scenes.size() == 1 &&
                       (tu.tile_order == 2 || (tu.tile_order == 0 && scenes[0].pr.spp >= 16));
  uint32_t tiles_x = 0;
  uint64_t claims = total;  // claim indices
  if (ordered) {
    tiles_x = (scenes[0].pr.width + 7) / 8;
    claims = (uint64_t)tiles_x * ((scenes[0].pr.height + 7) / 8) * 64;
  }
  if (segmented && claims >= (1ull << 31)) throw Error(ZT_ERR_IMAGE_TOO_LARGE, "segmented schedules take < 2^31 pixels");

  // Default pool: sample passes hold every pixel from the start (64 B each:
  // 2^26 pixels take 4 GiB), so no slot ever takes a second pixel and the
  // queue stays in claim order -- neighbouring entries are neighbouring
  // pixels of one scene (the 256-scene batch: a 4M-slot pool whose slots
  // take new pixels as theirs finish mixes scenes within waves, 487 against
  // 618 Mpaths/s); the bounce forms hold up to 4M paths (176 B each).
  uint64_t want = tu.queue_paths ? tu.queue_paths : pass ? claims : std::min<uint64_t>(claims, 1ull << 22);
  uint32_t nslots = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 64), std::min<uint64_t>(claims, 1ull << 26));
  // Segmented schedules: 8 segments of whole compaction ranges.
  uint32_t seg_slots = 0, rps = 0;
  if (segmented) {
    seg_slots = ((nslots + kSegs - 1) / kSegs + kRangeSlots - 1) / kRangeSlots * kRangeSlots;
    rps = seg_slots / kRangeSlots;
    nslots = kSegs * seg_slots;
  }
  const uint32_t nranges = (nslots + kRangeSlots - 1) / kRangeSlots;
  if (nranges > kMaxRanges) throw Error(ZT_ERR_INTERNAL, "wavefront slot pool exceeds the compaction's capacity");

  WfBuffers& b = ctx->wf;
  b.scenes.ensure(scenes.size() * sizeof(WfScene<R>));
  // Sample passes keep only a slot's pixel, stream and sum (4 arrays of 16
  // bytes) and use neither hit records nor the shadow queue.
  const size_t path_slots = pass ? 1 : nslots;
  b.slots.ensure((size_t)nslots * (pass ? 4 * sizeof(uint4) : WfSlots<R>::kBytes) + 11 * 256);
  b.hits.ensure(path_slots * sizeof(float2));
  b.hits_d.ensure(path_slots * sizeof(double));
  b.qext0.ensure((size_t)nslots * 4);
  b.qext1.ensure((size_t)nslots * 4);
  b.qsh.ensure(path_slots * 4);
  b.fext.ensure((size_t)nranges * kRangeSlots);
  b.fsh.ensure(pass ? 4096 : (size_t)nranges * kRangeSlots);
  b.ranges.ensure((size_t)nranges * 4);
  b.counters.ensure(kWfCounterWords * sizeof(uint32_t));
  b.tally.ensure(kTallyLines * 128);
  b.ensureHost();
  ZT_HIP(hipMemcpyAsync(b.scenes.p, scenes.data(), scenes.size() * sizeof(WfScene<R>), hipMemcpyHostToDevice,
                        ctx->stream));

  WfShared<R> w;
  std::memset((void*)&w, 0, sizeof w);  // padding included: its bytes key the cached graph
  w.scenes = b.scenes.as<WfScene<R>>();
  {
    char* base = (char*)b.slots.p;
    size_t off = 0;
    auto carve = [&](size_t elem) {
      void* p = base + off;
      off += ((size_t)nslots * elem + 255) & ~(size_t)255;
      return p;
    };
    if (!pass) {
      w.slots.ro = (V4<R>*)carve(sizeof(V4<R>));
      w.slots.rd = (V4<R>*)carve(sizeof(V4<R>));
      w.slots.shd = (V4<R>*)carve(sizeof(V4<R>));
      w.slots.rad = (V4<R>*)carve(sizeof(V4<R>));
      w.slots.thr = (V4<R>*)carve(sizeof(V4<R>));
      w.slots.shc = (V4<R>*)carve(sizeof(V4<R>));
      w.slots.meta = (int4*)carve(sizeof(int4));
    }
    w.slots.sum = (V4<R>*)carve(pass ? sizeof(uint4) : sizeof(V4<R>));
    w.slots.where = (uint4*)carve(sizeof(uint4));
    w.slots.rng01 = (uint4*)carve(sizeof(uint4));
    w.slots.rng23 = (uint4*)carve(sizeof(uint4));
  }
  w.hits = b.hits.as<float2>();
  w.hits_d = b.hits_d.as<double>();
  w.qext[0] = b.qext0.as<uint32_t>();
  w.qext[1] = b.qext1.as<uint32_t>();
  w.qsh = b.qsh.as<uint32_t>();
  w.fext = b.fext.as<uint8_t>();
  w.fsh = b.fsh.as<uint8_t>();
  w.ranges = b.ranges.as<uint32_t>();
  w.counters = b.counters.as<uint32_t>();
  w.tally = b.tally.as<unsigned long long>();
  w.num_scenes = (uint32_t)scenes.size();
  w.num_slots = nslots;
  w.num_ranges = nranges;
  w.total_pixels = claims;
  w.tiles_x = tiles_x;

  const int cap = tu.stack_entries ? (int)tu.stack_entries : (wave_trace ? 12 : kLds);
  if (cap < 8 || cap > 64 || (cap & 1)) throw Error(ZT_ERR_INVALID_ARGUMENT, "stack entries must be even, 8..64");
  // Sample passes serve later than the other forms: a take waits on the queue
  // cursor, the queue entry and the slot, one after the other, and serving
  // more lanes at once spreads those waits. One scene: 56 of 64 lanes waiting
  // (1080p mesh field, scan order: 32 215 ms, 48 183 ms, 56 179 ms, 64 197
  // ms per frame); many scenes, whose shorter paths make the waiting lanes'
  // idle time dearer: 40 (the 256-scene batch: 36-44 675-678 Mpaths/s, 48
  // 662, 56 617).
  const uint32_t refill = tu.refill ? tu.refill : pass ? (scenes.size() == 1 ? 56 : 40) : 32;
  if (refill > 64) throw Error(ZT_ERR_INVALID_ARGUMENT, "refill must be 1..64");
  const size_t lds = stackBytes(cap, (int)threads);
  const bool single = scenes.size() == 1;
  const void* wfn = nullptr;
  size_t wlds = 0;
  int g_wave = 0;
  if (wave_trace) {
    wlds = stackBytes(cap, 64) + kLeafPassBytes;
    wfn = single ? (quant ? (const void*)wfTraceWaveOne<true> : (const void*)wfTraceWaveOne<false>)
                 : (quant ? (const void*)wfTraceWaveMany<true> : (const void*)wfTraceWaveMany<false>);
    g_wave = (int)residentBlocks(wfn, 64, wlds);
  }
  const void* bfn[2] = {nullptr, nullptr};  // the two-launch bounce: extension, shadow
  int g_bounce[2] = {0, 0};
  if (fused) {
    if (single) {
      bfn[0] = quant ? (const void*)wfBounceOne<true, false> : (const void*)wfBounceOne<false, false>;
      bfn[1] = quant ? (const void*)wfBounceOne<true, true> : (const void*)wfBounceOne<false, true>;
    } else {
      bfn[0] = quant ? (const void*)wfBounceMany<true, false> : (const void*)wfBounceMany<false, false>;
      bfn[1] = quant ? (const void*)wfBounceMany<true, true> : (const void*)wfBounceMany<false, true>;
    }
    for (int k = 0; k < 2; ++k) g_bounce[k] = (int)residentBlocks(bfn[k], 64, wlds);
  }
  const void* sfn = nullptr;  // the one-launch bounce, or the sample pass
  int g_step = 0;
  if (step) {
    sfn = single ? (quant ? (const void*)wfStepOne<true> : (const void*)wfStepOne<false>)
                 : (quant ? (const void*)wfStepMany<true> : (const void*)wfStepMany<false>);
  } else if (pass) {
    // A pool that holds every pixel never claims another: the passes are
    // compiled without the claim path.
    if (seg_slots >= (claims + kSegs - 1) / kSegs)
      sfn = single ? (quant ? (const void*)wfPassOne<true, false> : (const void*)wfPassOne<false, false>)
                   : (quant ? (const void*)wfPassMany<true, false> : (const void*)wfPassMany<false, false>);
    else
      sfn = single ? (quant ? (const void*)wfPassOne<true, true> : (const void*)wfPassOne<false, true>)
                   : (quant ? (const void*)wfPassMany<true, true> : (const void*)wfPassMany<false, true>);
  }
  if (sfn) g_step = (int)residentBlocks(sfn, 64, wlds);
  // The tail hand-off (wfDrain): once every pixel is claimed and no more
  // live paths remain than the drain kernel holds lanes.
  const void* dfn = nullptr;
  int g_drain = 0;
  if (wave_trace) {
    if (pass)
      dfn = single ? (quant ? (const void*)wfDrain<true, true, true> : (const void*)wfDrain<false, true, true>)
                   : (quant ? (const void*)wfDrain<true, false, true> : (const void*)wfDrain<false, false, true>);
    else
      dfn = single ? (quant ? (const void*)wfDrain<true, true, false> : (const void*)wfDrain<false, true, false>)
                   : (quant ? (const void*)wfDrain<true, false, false> : (const void*)wfDrain<false, false, false>);
    g_drain = (int)residentBlocks(dfn, 64, wlds);
  }
  const uint64_t drain_at = (uint64_t)g_drain * 64;
  const int g_tr = wave_trace ? 0 : gridFor<R>((const void*)wfTrace<R>, threads, lds);
  const int g_shade = gridFor<R>((const void*)wfShade<R>, threads, 0);
  const int g_res = gridFor<R>((const void*)wfResolve<R>, threads, 0);
  const int g_init = std::max(1, (int)std::min<uint64_t>(((uint64_t)nranges * kRangeSlots + threads - 1) / threads, 65535));
  ctx->spill.ensure(spillBytes(std::max<uint64_t>({(uint64_t)g_tr * threads, (uint64_t)g_wave * 64,
                                                    (uint64_t)g_drain * 64, (uint64_t)g_bounce[0] * 64,
                                                    (uint64_t)g_bounce[1] * 64, (uint64_t)g_step * 64})));
  w.spill = ctx->spill.as<int>();
  // The cost probe after the spill buffer has its final size (it runs on the
  // same buffer).
  if constexpr (!K<R>::exact) {
    if (ordered) w.order = probeTileOrder(ctx, scenes[0].view, scenes[0].pr, scenes[0].out, cap, 32);
  }
  WfOne one;
  WfMany many;
  std::memset((void*)&one, 0, sizeof one);  // their bytes key the cached graph too
  std::memset((void*)&many, 0, sizeof many);
  if constexpr (!K<R>::exact) {
    one.sc = scenes[0].view;
    one.pr = scenes[0].pr;
    many.scenes = w.scenes;
  }
  const hipStream_t st = ctx->stream;
  // Ordered compaction of a flag array into queue q, its length into counters[c].
  auto compact = [&](uint8_t* flags, uint32_t* q, int c) {
    hipLaunchKernelGGL(wfCount, dim3(nranges), dim3(kRangeThreads), 0, st, flags, w.ranges);
    hipLaunchKernelGGL(wfScan, dim3(1), dim3(kScanThreads), 0, st, w.ranges, nranges, w.counters + c);
    hipLaunchKernelGGL(wfScatter, dim3(nranges), dim3(kRangeThreads), 0, st, flags, w.ranges, q);
  };
  // One bounce: the kernels of the chosen schedule.
  auto bounce = [&](int cur) {
    int c = cur, cp = cap;
    uint32_t r = refill;
    void* src = single ? (void*)&one : (void*)&many;
    if (sfn) {
      uint32_t rp = rps;
      void* args[] = {src, (void*)&w, &c, &cp, &r, &rp};
      ZT_HIP(hipLaunchKernel(sfn, dim3(g_step), dim3(64), args, wlds, st));
      compact(w.fext, w.qext[1 - cur], 1 - cur);
      return;
    }
    if (fused) {
      void* args[] = {src, (void*)&w, &c, &cp, &r};
      ZT_HIP(hipLaunchKernel(bfn[0], dim3(g_bounce[0]), dim3(64), args, wlds, st));
      compact(w.fsh, w.qsh, 2);
      ZT_HIP(hipLaunchKernel(bfn[1], dim3(g_bounce[1]), dim3(64), args, wlds, st));
      compact(w.fext, w.qext[1 - cur], 1 - cur);
      return;
    }
    auto trace = [&](int shadow) {
      if (wave_trace) {
        void* args[] = {src, (void*)&w, &c, &shadow, &cp, &r};
        ZT_HIP(hipLaunchKernel(wfn, dim3(g_wave), dim3(64), args, wlds, st));
      } else {
        hipLaunchKernelGGL(wfTrace<R>, dim3(g_tr), dim3(threads), lds, st, w, cur, shadow, cap);
      }
    };
    trace(0);
    hipLaunchKernelGGL(wfShade<R>, dim3(g_shade), dim3(threads), 0, st, w, cur);
    compact(w.fsh, w.qsh, 2);
    trace(1);
    hipLaunchKernelGGL(wfResolve<R>, dim3(g_res), dim3(threads), 0, st, w);
    compact(w.fext, w.qext[1 - cur], 1 - cur);
  };

  if (segmented) {
    hipLaunchKernelGGL(wfResetSegments, dim3(1), dim3(kTallyLines), 0, st, w.counters, w.tally, claims, seg_slots);
    if constexpr (!K<R>::exact) {
      if (pass) hipLaunchKernelGGL(wfInitPasses, dim3(g_init), dim3(threads), 0, st, w, seg_slots);
    }
    if (step) hipLaunchKernelGGL(wfInitSegments<R>, dim3(g_init), dim3(threads), 0, st, w, seg_slots);
  } else {
    hipLaunchKernelGGL(wfReset, dim3(1), dim3(kTallyLines), 0, st, w.counters, w.tally, (uint64_t)nslots);
    hipLaunchKernelGGL(wfInit<R>, dim3(g_init), dim3(threads), 0, st, w);
  }
  compact(w.fext, w.qext[0], 0);
  ZT_HIP(hipGetLastError());
  ctx->stats.launches += 5;

  // Bounce iterations in batches of an even count, captured as a graph. The
  // context keeps the instantiated graph with the bytes of everything its
  // launches were captured with, and reuses it while those match (repeated
  // renders of one scene and size).
  const int kBatch = 16;
  std::vector<unsigned char> key;
  auto put = [&](const void* p, size_t n) { key.insert(key.end(), (const unsigned char*)p, (const unsigned char*)p + n); };
  {
    const size_t elem = sizeof(R);
    const uint64_t ints[] = {elem, threads, (uint64_t)g_shade, (uint64_t)g_res, (uint64_t)g_tr, (uint64_t)g_wave,
                             lds, wlds, (uint64_t)cap, refill, (uint64_t)wave_trace, (uint64_t)single,
                             (uint64_t)fused, (uint64_t)g_bounce[0], (uint64_t)g_bounce[1], (uint64_t)step,
                             (uint64_t)pass, (uint64_t)g_step, (uint64_t)rps};
    put(ints, sizeof ints);
    put(&w, sizeof w);
    put(&wfn, sizeof wfn);
    put(bfn, sizeof bfn);
    put(&sfn, sizeof sfn);
    if (wave_trace) put(single ? (const void*)&one : (const void*)&many, single ? sizeof one : sizeof many);
  }
  if (b.graph_key != key) {
    b.dropGraph();
    ZT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < kBatch; ++j) bounce(j & 1);
    const hipError_t ce = hipStreamEndCapture(st, &b.graph);
    if (ce != hipSuccess) throw Error(ZT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    for (auto& ex : b.exec) {
      const hipError_t ie = hipGraphInstantiate(&ex, b.graph, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        b.dropGraph();
        throw Error(ZT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      }
    }
    b.graph_key = std::move(key);
  }

  // Each batch advances every live path by kBatch bounces; a frame needs at
  // most spp * (max_depth + 1) of them per pixel pass, so the loop is bounded.
  uint64_t max_batches = 1;
  for (const auto& s : scenes) {
    const uint64_t per_pixel = (uint64_t)s.pr.spp * ((uint64_t)std::min<uint32_t>(s.pr.max_depth, 1u << 20) + 2) * 2;
    const uint64_t rounds = (claims + nslots - 1) / nslots;
    max_batches = std::max<uint64_t>(max_batches, per_pixel * rounds / kBatch + 2);
  }
  // Two batches stay in flight: after each one the live count and the next
  // unclaimed pixel are copied to pinned memory behind an event, and the
  // host, once a batch has ended, decides with the next one still queued
  // whether to queue another, hand the live paths to the drain kernel
  // (which reads the live count on the device), or stop (a batch that finds
  // the queues empty does nothing).
  uint64_t launched = 0, seen = 0;
  bool drain = false, overrun = false;
  hipError_t le = hipSuccess;
  auto launchBatch = [&]() -> hipError_t {
    const int i = (int)(launched & 1);
    hipError_t e = hipGraphLaunch(b.exec[i], st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b.host + i * kWfHostWords, w.counters, kWfHostWords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                         st);
    if (e == hipSuccess) e = hipEventRecord(b.done[i], st);
    if (e == hipSuccess) ++launched;
    return e;
  };
  // Sample passes over a pool that holds every pixel from the start (each
  // band fits its segment): no pixel is claimed later, so the frame takes
  // exactly max(spp) passes and the batches are queued without waiting.
  uint32_t max_spp = 0;
  for (const auto& s : scenes) max_spp = std::max(max_spp, s.pr.spp);
  const bool all_at_once = pass && seg_slots >= (claims + kSegs - 1) / kSegs;
  if (all_at_once) {
    const uint64_t batches = (max_spp + kBatch - 1) / kBatch;
    for (uint64_t j = 0; j < batches && le == hipSuccess; ++j) {
      le = hipGraphLaunch(b.exec[0], st);
      if (le == hipSuccess) ++launched;
    }
  } else {
    le = launchBatch();
    if (le == hipSuccess) le = launchBatch();
  }
  while (le == hipSuccess && !all_at_once) {
    const int i = (int)(seen & 1);
    le = hipEventSynchronize(b.done[i]);
    if (le != hipSuccess) break;
    ++seen;
    const uint32_t* c = b.host + i * kWfHostWords;
    const uint32_t active = c[0];
    bool claimed = true;  // every pixel has been claimed
    if (segmented) {
      for (uint32_t k = 0; k < kSegs; ++k)
        claimed = claimed && c[16 + k] >= pixelBand(k + 1, claims) - pixelBand(k, claims);
    } else {
      uint64_t next_pixel = 0;
      std::memcpy(&next_pixel, c + 6, 8);
      claimed = next_pixel >= total;
    }
    if (active == 0) break;
    if (dfn && claimed && active <= drain_at) {
      int cp = cap;
      uint32_t r = refill;
      void* args[] = {(void*)&w, &cp, &r};
      le = hipLaunchKernel(dfn, dim3((uint32_t)g_drain), dim3(64), args, wlds, st);
      drain = true;
      break;
    }
    if (launched > max_batches) {
      overrun = true;
      le = hipErrorUnknown;
      break;
    }
    le = launchBatch();
  }
  if (le != hipSuccess) {
    if (overrun) throw Error(ZT_ERR_INTERNAL, "wavefront did not drain");
    throw Error(ZT_ERR_HIP, std::string("wavefront loop: ") + hipGetErrorString(le));
  }
  ctx->stats.launches += launched * kBatch * (segmented ? 4 : fused ? 8 : 10) + 1 + (drain ? 1 : 0);
  hipLaunchKernelGGL(wfReport, dim3(1), dim3(kTallyLines), 0, st, w.counters, w.tally, ctx->errflag.as<int>());
  ZT_HIP(hipGetLastError());
}
template void runWavefront<float>(zt_ctx*, const std::vector<WfScene<float>>&, const zt_gpu_tuning&);
template void runWavefront<double>(zt_ctx*, const std::vector<WfScene<double>>&, const zt_gpu_tuning&);

template <typename R>
void launchWavefront(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, R* dout) {
  WfScene<R> ws{};
  ws.view = viewOf<R>(s);
  ws.pr = makeParams(s, o);
  ws.view.cam = cameraInit(s.camera, (double)o.width / (double)o.height);
  applyBvhWidth(ws.view, s, tu);
  ws.out = dout;
  ws.first_pixel = 0;
  runWavefront<R>(ctx, std::vector<WfScene<R>>{ws}, tu);
}
template void launchWavefront<float>(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&,
                                     float*);
template void launchWavefront<double>(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&,
                                      double*);

}  // namespace ztg

using namespace ztg;

extern "C" int zt_gpu_render_batch(zt_ctx* ctx, uint32_t n, const uint32_t* scene_ids, const zt_render_options* opts,
                                   const zt_gpu_tuning* tuning, void* const* outs, int out_on_device) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (n == 0) return;
    if (!scene_ids || !opts || !outs) throw Error(ZT_ERR_INVALID_ARGUMENT, "null argument");
    zt_gpu_tuning tu{};
    if (tuning) tu = *tuning;
    if (tu.precision > ZT_PRECISION_FP64) throw Error(ZT_ERR_INVALID_ARGUMENT, "bad precision");
    for (uint32_t i = 0; i < n; ++i) {
      checkOptions(&opts[i]);
      if (!outs[i]) throw Error(ZT_ERR_INVALID_ARGUMENT, "null output");
      sceneOf(ctx, scene_ids[i]);
    }
    setDevice(ctx);
    const bool f64 = tu.precision == ZT_PRECISION_FP64;
    const size_t elem = f64 ? sizeof(double) : sizeof(float);
    std::vector<size_t> bytes(n), offs(n);
    size_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
      bytes[i] = (size_t)opts[i].width * opts[i].height * 3 * elem;
      offs[i] = total;
      total += (bytes[i] + 255) & ~(size_t)255;
    }
    if (!out_on_device) ctx->out.ensure(total);
    auto devOut = [&](uint32_t i) -> void* {
      return out_on_device ? outs[i] : (void*)((char*)ctx->out.p + offs[i]);
    };
    ctx->stats.launches = 0;
    resetErr(ctx);
    ZT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    uint64_t first = 0, paths = 0;
    auto build = [&](auto tag) {
      using R = decltype(tag);
      std::vector<WfScene<R>> scenes(n);
      for (uint32_t i = 0; i < n; ++i) {
        const DevScene& s = sceneOf(ctx, scene_ids[i]);
        scenes[i].view = viewOf<R>(s);
        scenes[i].view.cam = cameraInit(s.camera, (double)opts[i].width / (double)opts[i].height);
        applyBvhWidth(scenes[i].view, s, tu);
        scenes[i].pr = makeParams(s, opts[i]);
        scenes[i].out = (R*)devOut(i);
        scenes[i].first_pixel = first;
        first += (uint64_t)opts[i].width * opts[i].height;
        paths += (uint64_t)opts[i].width * opts[i].height * opts[i].spp;
      }
      // fp32 scenes that all have four-wide trees go through the persistent
      // wave-scheduled megakernel (one launch over every scene's tiles)
      // unless the wavefront renderer or the one-pixel-per-lane schedule is
      // asked for; everything else through the wavefront renderer.
      bool wave = !K<R>::exact && tu.method != ZT_METHOD_WAVEFRONT && tu.schedule != 1;
      for (uint32_t i = 0; wave && i < n; ++i) wave = scenes[i].view.nodes4 != nullptr;
      if constexpr (!K<R>::exact) {
        if (wave) {
          std::vector<WaveJob> jobs(n);
          for (uint32_t i = 0; i < n; ++i) jobs[i] = WaveJob{scenes[i].view, scenes[i].pr, scenes[i].out, 0, 0};
          launchWaveJobs(ctx, jobs, tu);
          return;
        }
      }
      runWavefront<R>(ctx, scenes, tu);
    };
    if (f64) build(double{});
    else build(float{});
    ZT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (!out_on_device)
      for (uint32_t i = 0; i < n; ++i)
        ZT_HIP(hipMemcpyAsync(outs[i], devOut(i), bytes[i], hipMemcpyDeviceToHost, ctx->stream));
    checkErr(ctx);
    float ms = 0;
    ZT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms = ms;
    ctx->stats.paths = paths;
  });
}
