// The plain renderer: one lane per pixel (replacing the tile loop of
// render.zig:104-180), plus the ray-level entry points used for parity
// (traceWith, Scene.intersect, the RNG streams).
//
// A lane owns its pixel's xoshiro256++ stream (render.zig:135) and its
// accumulator, so images need no atomics and are bit-identical for any launch
// shape. The lane runs a flat loop of ray queries: each iteration traces one
// extension or shadow ray, then shades or resolves it; when a path ends the
// lane starts its pixel's next sample in the same iteration structure, so a
// wave keeps tracing while its lanes' paths have different lengths.

#include <cmath>
#include <cstring>

#include "../device/dev_path.hpp"
#include "../device/dev_wave.hpp"
#include "ctx.hpp"
#include "launch.hpp"

using namespace ztd;

namespace ztg {

static constexpr double kRayEpsilon = 1e-4;  // render.zig:26

template <typename R>
ZT_D HitRec<R> rayQuery(const SceneView<R>& sc, const LaneStack& stk, V3<R> o, V3<R> d, R tmax, int self, int excl,
                        bool anyhit, int* overflow) {
  return query<R>(sc, stk, o, d, R(kRayEpsilon), tmax, self, excl, anyhit, overflow);
}

// Tiles are split into 8 contiguous bands, one per XCD: the hardware places
// block b on XCD b mod 8, so a block that works through band b mod 8 shares
// its L2 with the blocks rendering the neighbouring tiles.
ZT_D uint32_t bandStart(uint32_t band, uint32_t ntiles) {
  const uint32_t q = ntiles / 8, r = ntiles % 8;
  return band * q + min(band, r);
}

// One pixel: all its samples from its own stream, in order (render.zig:133-144).
template <typename R>
ZT_D void renderPixel(const SceneView<R>& sc, const Params& pr, const LaneStack& stk, uint32_t x, uint32_t y,
                      R* __restrict__ out, int* overflow, RayTally& rays) {
  const uint64_t pix = (uint64_t)y * pr.width + x;
  Rng rng = rngForStream(pr.seed, pix);
  V3<R> sum = splat(R(0));
  Path<R> st;
  uint32_t sample = 0;
  int phase = kCamera;
  for (;;) {
    if (phase == kCamera) {
      if (sample == pr.spp) break;
      phase = beginPath<R>(sc, pr, x, y, rng, st);
    }
    if (phase == kExtend || phase == kShadow) {
      const bool sh = phase == kShadow;
      const HitRec<R> h = rayQuery<R>(sc, stk, st.ro, sh ? st.sh_d : st.rd, sh ? st.sh_tmax : R(INFINITY), st.self,
                                      sh ? st.sh_excl : -1, sh, overflow);
      rays.count(sh);
      phase = sh ? afterShadow<R>(h, st) : shade<R>(sc, pr, h, rng, st);
    }
    if (phase == kFinish) {
      if (isFinite(st.radiance)) sum = add(sum, st.radiance);  // render.zig:141-142
      ++sample;
      phase = kCamera;
    }
  }
  const R inv_spp = R(1) / R(pr.spp);
  out[3 * pix + 0] = sum.x * inv_spp;
  out[3 * pix + 1] = sum.y * inv_spp;
  out[3 * pix + 2] = sum.z * inv_spp;
}

// Persistent megakernel: each block takes tile after tile from its XCD's band
// (one atomic per tile on band_next[band]) and, once that band is done, from
// the other bands in turn. A pixel's result does not depend on which block
// renders it.
template <typename R>
ZT_D void megakernelBody(const SceneView<R>& sc, const Params& pr, R* __restrict__ out, uint32_t tile_w,
                         uint32_t tile_h, uint32_t tiles_x, uint32_t ntiles, int cap, int* spill,
                         uint32_t* band_next, int* err) {
  extern __shared__ int lds[];
  __shared__ uint32_t s_tile;
  const uint32_t tid = threadIdx.x;
  const LaneStack stk = laneStack(lds, cap, spill);
  int overflow = 0;
  RayTally rays;
  const uint32_t home = blockIdx.x % 8;
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t band = (home + k) % 8;
    const uint32_t hi = bandStart(band + 1, ntiles);
    for (;;) {
      if (tid == 0) s_tile = bandStart(band, ntiles) + atomicAdd(&band_next[band], 1u);
      __syncthreads();
      const uint32_t tile = s_tile;
      __syncthreads();
      if (tile >= hi) break;
      const uint32_t x = (tile % tiles_x) * tile_w + tid % tile_w;
      const uint32_t y = (tile / tiles_x) * tile_h + tid / tile_w;
      if (x < pr.width && y < pr.height) renderPixel<R>(sc, pr, stk, x, y, out, &overflow, rays);
    }
  }
  rays.flush(err);
  if (overflow) atomicOr(err, 1);
}

// The same kernel compiled for different occupancy targets: W > 0 asks the
// compiler to fit at least W waves per SIMD (registers beyond that budget
// spill). Every variant renders the same bits.
template <typename R, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1, 8))) void megakernel(
    SceneView<R> sc, Params pr, R* __restrict__ out, uint32_t tile_w, uint32_t tile_h, uint32_t tiles_x,
    uint32_t ntiles, int cap, int* spill, uint32_t* band_next, int* err) {
  megakernelBody<R>(sc, pr, out, tile_w, tile_h, tiles_x, ntiles, cap, spill, band_next, err);
}

// Wave-scheduled megakernel. One wave per workgroup; every lane owns a pixel
// (its stream, accumulator and sample count, as renderPixel) and at most one
// ray in flight, walked with a resumable walker (Trav4 over the four-wide
// tree of the fp32 path, Trav2 over the two-child tree). The wave runs
// wave-uniform passes -- node steps, leaf tests -- until `refill` lanes are
// waiting for work, then those lanes shade their hit (or apply their shadow
// ray), start their next ray or sample, or take the next pixel of the wave's
// current tile, while the other lanes keep their traversal state. Each
// pixel's computation is the same sequence of device functions as
// renderPixel's, so the image is identical; only the interleaving of lanes
// changes.
// The scenes a wave-scheduled launch renders, with their tiles numbered in
// one index space. One scene: everything in kernel arguments. Many scenes:
// a table in device memory, each lane reading its own pixel's scene.
template <typename R>
struct OneSceneT {
  static constexpr bool kSingle = true;
  SceneView<R> sc;
  Params pr;
  R* out;
  uint32_t tiles_x;
  ZT_D uint32_t sceneOf(uint32_t) const { return 0; }
  ZT_D uint32_t firstTile(uint32_t) const { return 0; }
  ZT_D uint32_t tilesX(uint32_t) const { return tiles_x; }
  ZT_D const SceneView<R>& view(uint32_t) const { return sc; }
  ZT_D const Params& params(uint32_t) const { return pr; }
  ZT_D R* output(uint32_t) const { return out; }
};
using OneScene = OneSceneT<float>;

struct ManyScenes {
  static constexpr bool kSingle = false;
  const WaveJob* jobs;
  uint32_t n;
  ZT_D uint32_t sceneOf(uint32_t tile) const {  // last job whose first tile <= tile
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (jobs[mid].first_tile <= tile) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
  ZT_D uint32_t firstTile(uint32_t si) const { return jobs[si].first_tile; }
  ZT_D uint32_t tilesX(uint32_t si) const { return jobs[si].tiles_x; }
  ZT_D const SceneView<float>& view(uint32_t si) const { return jobs[si].view; }
  ZT_D const Params& params(uint32_t si) const { return jobs[si].pr; }
  ZT_D float* output(uint32_t si) const { return jobs[si].out; }
};

// `order` (may be null) lists the tiles in the order they are handed out
// within each band; the cost probe (kCost) instead adds up per tile, in
// `cost`, the cube of the rays each of the tile's pixels starts. `Walk` is the lanes' walker: Trav4 over the
// four-wide tree (fp32), Trav2<R> over the two-child tree.
template <typename R, typename Walk, bool Q, bool kCost, typename Src>
ZT_D void megakernelWaveBody(const Src& src, uint32_t tile_w, uint32_t tile_h, uint32_t ntiles, int cap,
                             uint32_t refill, int* spill, uint32_t* band_next, const uint32_t* order,
                             uint32_t* cost, int* err) {
  constexpr int kResolve = 4;  // a traversal result is waiting to be applied
  extern __shared__ int lds[];
  const LaneStack stk = laneStack<64>(lds, cap, spill);  // workgroups of one wave (64 threads)
  unsigned long long* leaf_keys = (unsigned long long*)(lds + stackBytes(cap, 64) / sizeof(int));
  uint8_t* leaf_owner = (uint8_t*)(leaf_keys + 64);
  const uint32_t lane = threadIdx.x;
  const uint32_t tile_px = tile_w * tile_h;
  const uint32_t home = blockIdx.x % 8;
  // Wave-uniform tile cursor.
  uint32_t band_k = 0, tile = 0, tile_scene = 0, next_px = tile_px;
  bool tiles_done = false;
  // Lane state.
  bool has = false, tracing = false, sh = false;
  uint32_t si = 0, x = 0, y = 0, sample = 0;
  int phase = kCamera;
  Rng rng{};
  V3<R> sum = splat(R(0));
  Path<R> st;
  Walk tr;
  int overflow = 0;
  uint64_t n_ext = 0, n_sh = 0;  // rays started by the wave (wave-uniform)
  uint32_t pix_rays = 0;          // cost probe: rays the lane's pixel started
  for (;;) {
    // Pixels for lanes without one, from the wave's tile (then the next tile).
    for (;;) {
      const uint64_t need = __ballot(!has);
      if (need == 0 || tiles_done) break;
      if (next_px >= tile_px) {
        for (;;) {
          if (band_k == 8) {
            tiles_done = true;
            break;
          }
          const uint32_t band = (home + band_k) % 8;
          uint32_t t = 0;
          if (lane == 0) t = bandStart(band, ntiles) + atomicAdd(&band_next[band], 1u);
          t = __shfl(t, 0);
          if (t < bandStart(band + 1, ntiles)) {
            tile = order ? order[t] : t;
            tile_scene = src.sceneOf(tile);
            next_px = 0;
            break;
          }
          ++band_k;
        }
        if (tiles_done) break;
      }
      const uint32_t idx = next_px + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
      next_px = min(tile_px, next_px + (uint32_t)__popcll(need));
      if (!has && idx < tile_px) {
        const Params& tp = src.params(tile_scene);
        const uint32_t lt = tile - src.firstTile(tile_scene), tx = src.tilesX(tile_scene);
        const uint32_t px = (lt % tx) * tile_w + idx % tile_w;
        const uint32_t py = (lt / tx) * tile_h + idx / tile_w;
        if (px < tp.width && py < tp.height) {
          si = tile_scene;
          x = px;
          y = py;
          has = true;
          rng = rngForStream(tp.seed, (uint64_t)py * tp.width + px);  // render.zig:135
          sum = splat(R(0));
          sample = 0;
          phase = kCamera;
        }
      }
    }
    // Lanes with no ray in flight advance their path to its next ray.
    const bool idle = has && !tracing;
    if (idle) {
      const SceneView<R>& sc = src.view(si);
      const Params& pr = src.params(si);
      for (;;) {
        if (phase == kResolve) {
          const HitRec<R> hit{tr.closest, tr.best};  // what tr.finish left
          phase = sh ? afterShadow<R>(hit, st) : shade<R>(sc, pr, hit, rng, st);
        }
        if (phase == kFinish) {
          if (isFinite(st.radiance)) sum = add(sum, st.radiance);  // render.zig:141-142
          ++sample;
          phase = kCamera;
        }
        if (phase == kCamera) {
          if (sample == pr.spp) {
            R* out = src.output(si);
            const uint64_t pix = (uint64_t)y * pr.width + x;
            const R inv_spp = R(1) / R(pr.spp);
            out[3 * pix + 0] = sum.x * inv_spp;
            out[3 * pix + 1] = sum.y * inv_spp;
            out[3 * pix + 2] = sum.z * inv_spp;
            if constexpr (kCost) {  // one add per pixel, not one per ray
              // The pixel's rays cubed: a tile holding a few long pixels ranks
              // above one whose cost is spread evenly, so the long chains that
              // end the frame start early. Below 2^20 per pixel, a tile's sum
              // (at most 4096 pixels) stays below 2^32.
              const uint64_t c = pix_rays;
              const uint32_t weight = (uint32_t)min(c * c * c, (uint64_t)((1u << 20) - 1u));
              atomicAdd(&cost[src.firstTile(si) + (y / tile_h) * src.tilesX(si) + x / tile_w], weight);
              pix_rays = 0;
            }
            has = false;
            break;
          }
          phase = beginPath<R>(sc, pr, x, y, rng, st);
          continue;  // kFinish at once when max_depth == 0
        }
        sh = phase == kShadow;
        tr.start(QueryR<R>{st.ro, sh ? st.sh_d : st.rd, R(kRayEpsilon), st.self, sh ? st.sh_excl : -1, sh},
                 sh ? st.sh_tmax : R(INFINITY));
        tracing = true;
        if constexpr (kCost) ++pix_rays;
        break;
      }
    }
    if constexpr (!kCost) {
      const uint64_t started = __ballot(idle && tracing), shadow = __ballot(idle && tracing && sh);
      n_ext += (uint64_t)__popcll(started & ~shadow);
      n_sh += (uint64_t)__popcll(shadow);
    }
    if (__ballot(tracing) == 0) {
      // Every lane is without a pixel (a lane with one always has a ray now).
      if (tiles_done) break;
      continue;
    }
    // Traverse in wave-uniform passes (wavePass) until enough lanes wait for
    // work.
    for (;;) {
      const QueryR<R> q{st.ro, sh ? st.sh_d : st.rd, R(kRayEpsilon), st.self, sh ? st.sh_excl : -1, sh};
      const bool at_leaf = tracing && tr.atLeaf();
      const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
      const uint32_t n_node = (uint32_t)__popcll(__ballot(tracing && !at_leaf));
      if (n_leaf + n_node == 0) break;
      const uint32_t waiting = (uint32_t)__popcll(__ballot(!tracing && (has || !tiles_done)));
      if (waiting >= refill) break;
      const bool ended =
          wavePass<Q>(src, si, tr, q, tracing, at_leaf, n_leaf, n_node, stk, &overflow, leaf_keys, leaf_owner);
      if (ended) {
        (void)tr.finish(src.view(si), q);
        tracing = false;
        phase = kResolve;
      }
    }
  }
  if constexpr (!kCost)
    if (lane == 0) addRayCounts(err, n_ext, n_sh);
  if (overflow) atomicOr(err, 1);
}

template <int W, bool Q, bool kCost = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, 8))) void megakernelWave(
    OneScene src, uint32_t tile_w, uint32_t tile_h, uint32_t ntiles, int cap, uint32_t refill, int* spill,
    uint32_t* band_next, const uint32_t* order, uint32_t* cost, int* err) {
  megakernelWaveBody<float, Trav4, Q, kCost>(src, tile_w, tile_h, ntiles, cap, refill, spill, band_next, order, cost,
                                            err);
}

// The wave-scheduled kernel over the two-child tree: the fp64 path's default
// (and fp32's when the scene keeps only the two-child tree).
template <typename R, int W, bool kCost = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, 8))) void megakernelWave2(
    OneSceneT<R> src, uint32_t tile_w, uint32_t tile_h, uint32_t ntiles, int cap, uint32_t refill, int* spill,
    uint32_t* band_next, const uint32_t* order, uint32_t* cost, int* err) {
  megakernelWaveBody<R, Trav2<R>, false, kCost>(src, tile_w, tile_h, ntiles, cap, refill, spill, band_next, order,
                                                cost, err);
}

// Many scenes in one persistent launch (zt_gpu_render_batch): the same
// kernel over the concatenated tiles of every scene.
template <int W, bool Q, bool kCost = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, 8))) void megakernelWaveBatch(
    ManyScenes src, uint32_t tile_w, uint32_t tile_h, uint32_t ntiles, int cap, uint32_t refill, int* spill,
    uint32_t* band_next, const uint32_t* order, uint32_t* cost, int* err) {
  megakernelWaveBody<float, Trav4, Q, kCost>(src, tile_w, tile_h, ntiles, cap, refill, spill, band_next, order, cost,
                                            err);
}

__global__ void clearBands(uint32_t* band_next) {
  if (threadIdx.x < 8) band_next[threadIdx.x] = 0;
}

__global__ void zeroWords(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Cost bucket of a tile's probe ray count: the position of its leading bit
// and the two bits below it, so buckets grow in steps of at most 25%.
ZT_D uint32_t costBucket(uint32_t c) {
  const uint32_t v = c + 1u;
  const uint32_t e = 31u - (uint32_t)__clz(v);
  const uint32_t m = e >= 2 ? (v >> (e - 2)) & 3u : (v << (2 - e)) & 3u;
  return e * 4u + m;  // < 128
}

// Longest-first tile order: one workgroup per band sorts the band's tiles by
// cost bucket, most expensive first (order within a bucket is arbitrary;
// the image does not depend on it), so a band ends with its cheapest tiles
// and the last pixels in flight are short.
__global__ __launch_bounds__(256) void orderTiles(const uint32_t* cost, uint32_t* order, uint32_t ntiles) {
  __shared__ uint32_t count[128], next[128];
  const uint32_t band = blockIdx.x, lo = bandStart(band, ntiles), hi = bandStart(band + 1, ntiles);
  for (uint32_t i = threadIdx.x; i < 128; i += blockDim.x) count[i] = 0;
  __syncthreads();
  for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) atomicAdd(&count[costBucket(cost[t])], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = lo;
    for (int b = 127; b >= 0; --b) {
      next[b] = run;
      run += count[b];
    }
  }
  __syncthreads();
  for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) order[atomicAdd(&next[costBucket(cost[t])], 1u)] = t;
}

// traceWith for caller-given rays and RNG states (render.zig:61-102).
template <typename R>
ZT_D void traceOne(const SceneView<R>& sc, const Params& pr, const LaneStack& stk, const zt_ray* rays, uint64_t i,
                   uint64_t* states, double* out, int* overflow, RayTally& tally) {
  Rng rng{states[4 * i], states[4 * i + 1], states[4 * i + 2], states[4 * i + 3]};
  Path<R> st;
  st.ro = mk((R)rays[i].origin[0], (R)rays[i].origin[1], (R)rays[i].origin[2]);
  st.rd = mk((R)rays[i].dir[0], (R)rays[i].dir[1], (R)rays[i].dir[2]);
  st.radiance = splat(R(0));
  st.throughput = splat(R(1));
  st.prev_pdf = R(0);
  st.self = -1;
  st.sh_excl = -1;
  st.depth = 0;
  st.terminated = 0;
  int phase = pr.max_depth == 0 ? kFinish : kExtend;
  while (phase != kFinish) {
    const bool sh = phase == kShadow;
    const HitRec<R> h = rayQuery<R>(sc, stk, st.ro, sh ? st.sh_d : st.rd, sh ? st.sh_tmax : R(INFINITY), st.self,
                                 sh ? st.sh_excl : -1, sh, overflow);
    tally.count(sh);
    phase = sh ? afterShadow<R>(h, st) : shade<R>(sc, pr, h, rng, st);
  }
  out[3 * i + 0] = st.radiance.x;
  out[3 * i + 1] = st.radiance.y;
  out[3 * i + 2] = st.radiance.z;
  states[4 * i + 0] = rng.s0;
  states[4 * i + 1] = rng.s1;
  states[4 * i + 2] = rng.s2;
  states[4 * i + 3] = rng.s3;
}

template <typename R>
__global__ __launch_bounds__(256) void traceKernel(SceneView<R> sc, Params pr, const zt_ray* rays, uint64_t n,
                                                   uint64_t* states, double* out, int* spill, int* err) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack(lds, kLds, spill);
  int overflow = 0;
  RayTally tally;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    traceOne<R>(sc, pr, stk, rays, i, states, out, &overflow, tally);
  tally.flush(err);
  if (overflow) atomicOr(err, 1);
}

// Scene.intersect / brute force for caller-given rays (no self exclusion:
// these rays do not start on a surface).
template <typename R>
ZT_D zt_hit intersectOne(const SceneView<R>& sc, const LaneStack& stk, const zt_ray& ray, R tmin, R tmax, int mode,
                         int* overflow) {
  const V3<R> o = mk((R)ray.origin[0], (R)ray.origin[1], (R)ray.origin[2]);
  const V3<R> d = mk((R)ray.dir[0], (R)ray.dir[1], (R)ray.dir[2]);
  const HitRec<R> h = mode == 1 ? bruteForce<R>(sc, o, d, tmin, tmax, -1, -1, false)
                                : query<R>(sc, stk, o, d, tmin, tmax, -1, -1, false, overflow);
  zt_hit r{};
  r.hit = h.prim >= 0 ? 1 : 0;
  if (h.prim >= 0) {
    const Surface<R> s = surfaceAt<R>(sc, o, d, h);
    r.t = h.t;
    r.point[0] = s.p.x;
    r.point[1] = s.p.y;
    r.point[2] = s.p.z;
    r.normal[0] = s.n.x;
    r.normal[1] = s.n.y;
    r.normal[2] = s.n.z;
    r.front_face = s.front ? 1 : 0;
    r.material = (uint32_t)s.material;
    r.prim = h.prim;
  }
  return r;
}

// Scene.intersect / brute force for caller-given rays (no self exclusion:
// these rays do not start on a surface).
template <typename R>
__global__ __launch_bounds__(256) void intersectKernel(SceneView<R> sc, const zt_ray* rays, uint64_t n, R tmin, R tmax,
                                                       int mode, zt_hit* out, int* spill, int* err) {
  extern __shared__ int lds[];
  const LaneStack stk = laneStack(lds, kLds, spill);
  if (mode == 2) sc.nodes4 = nullptr;  // the two-child traversal
  int overflow = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = intersectOne<R>(sc, stk, rays[i], tmin, tmax, mode, &overflow);
  if (overflow) atomicOr(err, 1);
}

__global__ void rngKernel(uint64_t seed, const uint64_t* streams, uint64_t n, uint32_t draws, int kind,
                          uint64_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng r = rngForStream(seed, streams[i]);
  for (uint32_t k = 0; k < draws; ++k) {
    uint64_t v;
    if (kind == 0) v = r.next();
    else v = (uint64_t)__float_as_uint(r.u32f());
    out[i * draws + k] = v;
  }
}

static size_t ldsBytes(uint32_t threads) { return stackBytes(kLds, (int)threads); }
// Grid cap of the caller-given-ray kernels (grid-stride loops).
static constexpr uint64_t kRayGridBlocks = 2048;

Params makeParams(const DevScene& s, const zt_render_options& o) {
  Params p{};
  p.seed = o.seed;
  p.width = (uint32_t)o.width;
  p.height = (uint32_t)o.height;
  p.spp = (uint32_t)o.spp;
  p.max_depth = (uint32_t)std::min<uint64_t>(o.max_depth, 0xFFFFFFFFull);
  p.use_nee = (o.light_sampling != 0 && s.num_lights > 0) ? 1 : 0;
  return p;
}

void checkOptions(const zt_render_options* o) {
  if (!o) throw Error(ZT_ERR_INVALID_ARGUMENT, "null options");
  // render.zig:153, checked before anything else.
  if (o->width == 0 || o->height == 0 || o->spp == 0 || o->tile_size == 0)
    throw Error(ZT_ERR_INVALID_OPTIONS, "width, height, spp and tile_size must be non-zero");
  if (o->width > (1u << 20) || o->height > (1u << 20) || o->width * o->height > (1ull << 31) ||
      o->spp > 0xFFFFFFFFull)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "image size or spp beyond the device limits");
}

// Launch shape of the wave-scheduled kernels, from the tuning (0 = default).
struct WaveShape {
  uint32_t tw, th, refill, variant;
  int cap;
};
static WaveShape waveShape(const zt_gpu_tuning& tu) {
  // A tile is a wave's unit of work: 8x8 gives every lane one pixel to start
  // with and keeps enough tiles for every resident wave on small images.
  WaveShape w{tu.tile_w ? tu.tile_w : 8, tu.tile_h ? tu.tile_h : 8, tu.refill ? tu.refill : 32,
              tu.variant ? tu.variant : 2u, tu.stack_entries ? (int)tu.stack_entries : 12};
  if (w.tw * w.th < 64 || w.tw * w.th > 4096)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "wave-scheduled tiles must hold 64..4096 pixels");
  if (w.refill > 64) throw Error(ZT_ERR_INVALID_ARGUMENT, "refill must be 1..64");
  if (w.cap < 8 || w.cap > 64 || (w.cap & 1)) throw Error(ZT_ERR_INVALID_ARGUMENT, "stack entries must be even, 8..64");
  if (w.variant < 1 || w.variant > 4) throw Error(ZT_ERR_INVALID_ARGUMENT, "megakernel variant must be 0..4");
  if (tu.tile_order > 2) throw Error(ZT_ERR_INVALID_ARGUMENT, "tile_order must be 0, 1 or 2");
  return w;
}

// One launch of a wave-scheduled kernel (every one takes its scene source
// first, then these arguments), after clearing the band cursors.
static void launchWave(zt_ctx* ctx, const void* f, uint32_t grid, void* src, const WaveShape& w, uint32_t nt,
                       size_t lds, const uint32_t* ord, uint32_t* cost) {
  uint32_t* bands = ctx->bands.as<uint32_t>();
  int* spill = ctx->spill.as<int>();
  int* err = ctx->errflag.as<int>();
  uint32_t tw = w.tw, th = w.th, refill = w.refill;
  int cap = w.cap;
  void* args[] = {src, (void*)&tw, (void*)&th, (void*)&nt, (void*)&cap, (void*)&refill, &spill, &bands,
                  (void*)&ord, (void*)&cost, &err};
  hipLaunchKernelGGL(clearBands, dim3(1), dim3(64), 0, ctx->stream, bands);
  ZT_HIP(hipLaunchKernel(f, dim3(grid), dim3(64), args, lds, ctx->stream));
  ctx->stats.launches += 2;
}

// Longest-first tile order: the cost probe `probe_fn` (one sample per pixel
// without shadow rays: `probe_src` carries those parameters) sums over each
// tile the cube of its pixels' ray counts, and orderTiles sorts every band's
// tiles by that sum, most expensive first. Returns the order (context-owned).
static const uint32_t* orderByProbe(zt_ctx* ctx, const void* probe_fn, uint32_t grid, void* probe_src,
                                    const WaveShape& w, uint32_t nt, size_t lds) {
  ctx->tile_cost.ensure((size_t)nt * sizeof(uint32_t));
  ctx->tile_order.ensure((size_t)nt * sizeof(uint32_t));
  uint32_t* cost = ctx->tile_cost.as<uint32_t>();
  hipLaunchKernelGGL(zeroWords, dim3((nt + 255) / 256 < 1024 ? (nt + 255) / 256 : 1024), dim3(256), 0, ctx->stream,
                     cost, nt);
  launchWave(ctx, probe_fn, grid, probe_src, w, nt, lds, nullptr, cost);
  hipLaunchKernelGGL(orderTiles, dim3(8), dim3(256), 0, ctx->stream, cost, ctx->tile_order.as<uint32_t>(), nt);
  ctx->stats.launches += 2;
  return ctx->tile_order.as<uint32_t>();
}

// Longest-first tile order (tuning tile_order: 0 = for one scene with at
// least 16 samples per pixel, so the one-sample cost probe adds at most
// 1/16, and more tiles than resident waves -- when every tile starts at once
// their order cannot matter; 1 = scan order; 2 = always). A probe of one
// pixel per 2x2 block, or with paths cut at depth 6, orders worse than it
// saves; across the scenes of a batch the order mixes scenes within each
// wave and loses more than the shorter tail gains.
static bool orderTilesFor(const zt_gpu_tuning& tu, bool many, uint64_t pixels, uint64_t paths, uint64_t ntiles,
                          uint32_t grid) {
  return tu.tile_order == 2 || (tu.tile_order == 0 && !many && paths >= 16 * pixels && ntiles > grid);
}

// Wave-scheduled fp32 render over the four-wide BVH (see megakernelWaveBody)
// of one scene (kernel arguments) or many (a scene table in device memory).
void launchWaveJobs(zt_ctx* ctx, std::vector<WaveJob>& jobs, const zt_gpu_tuning& tu) {
  const WaveShape w = waveShape(tu);
  uint64_t ntiles = 0;
  for (WaveJob& j : jobs) {
    j.tiles_x = (j.pr.width + w.tw - 1) / w.tw;
    j.first_tile = (uint32_t)ntiles;
    ntiles += (uint64_t)j.tiles_x * ((j.pr.height + w.th - 1) / w.th);
  }
  if (ntiles == 0) return;
  if (ntiles > 0xFFFFFFFFull - 8) throw Error(ZT_ERR_INVALID_ARGUMENT, "too many tiles in one launch");
  const bool many = jobs.size() > 1;
  // 64-byte nodes when the caller asked for them (then every scene has them).
  const bool quant = jobs[0].view.nodes4q != nullptr;
  for (const WaveJob& j : jobs)
    if ((j.view.nodes4q != nullptr) != quant) throw Error(ZT_ERR_INTERNAL, "mixed node formats in one launch");
  const size_t lds = stackBytes(w.cap, 64) + kLeafPassBytes;
  static const void* const kOne[2][4] = {
      {(const void*)megakernelWave<1, false>, (const void*)megakernelWave<4, false>,
       (const void*)megakernelWave<5, false>, (const void*)megakernelWave<6, false>},
      {(const void*)megakernelWave<1, true>, (const void*)megakernelWave<4, true>,
       (const void*)megakernelWave<5, true>, (const void*)megakernelWave<6, true>}};
  static const void* const kMany[2][4] = {
      {(const void*)megakernelWaveBatch<1, false>, (const void*)megakernelWaveBatch<4, false>,
       (const void*)megakernelWaveBatch<5, false>, (const void*)megakernelWaveBatch<6, false>},
      {(const void*)megakernelWaveBatch<1, true>, (const void*)megakernelWaveBatch<4, true>,
       (const void*)megakernelWaveBatch<5, true>, (const void*)megakernelWaveBatch<6, true>}};
  static const void* const kProbe[2][2] = {
      {(const void*)megakernelWave<4, false, true>, (const void*)megakernelWave<4, true, true>},
      {(const void*)megakernelWaveBatch<4, false, true>, (const void*)megakernelWaveBatch<4, true, true>}};
  const void* fn = (many ? kMany : kOne)[quant ? 1 : 0][w.variant - 1];
  uint64_t pixels = 0, paths = 0;
  for (const WaveJob& j : jobs) {
    pixels += (uint64_t)j.pr.width * j.pr.height;
    paths += (uint64_t)j.pr.width * j.pr.height * j.pr.spp;
  }
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, residentBlocks(fn, 64, lds));
  const bool ordered = orderTilesFor(tu, many, pixels, paths, ntiles, grid);
  const void* probe_fn = kProbe[many ? 1 : 0][quant ? 1 : 0];
  const uint32_t probe_grid = ordered ? (uint32_t)std::min<uint64_t>(ntiles, residentBlocks(probe_fn, 64, lds)) : 0;
  ctx->spill.ensure(spillBytes((uint64_t)std::max(grid, probe_grid) * 64));
  ctx->bands.ensure(8 * sizeof(uint32_t));
  const uint32_t nt = (uint32_t)ntiles;
  // One scene: everything in kernel arguments. Many: the scene table (and the
  // probe's copy of it, one sample per pixel) in device memory.
  std::vector<WaveJob> probe_jobs(jobs);
  for (WaveJob& j : probe_jobs) {  // one sample; extension rays only (they rank tiles as well)
    j.pr.spp = 1;
    j.pr.use_nee = 0;
  }
  OneScene one{jobs[0].view, jobs[0].pr, jobs[0].out, jobs[0].tiles_x};
  OneScene one_probe{jobs[0].view, probe_jobs[0].pr, jobs[0].out, jobs[0].tiles_x};
  ManyScenes table{}, table_probe{};
  if (many) {
    const size_t bytes = jobs.size() * sizeof(WaveJob);
    ctx->jobs.ensure(2 * bytes);
    ZT_HIP(hipMemcpyAsync(ctx->jobs.p, jobs.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    ZT_HIP(hipMemcpyAsync((char*)ctx->jobs.p + bytes, probe_jobs.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    table = ManyScenes{ctx->jobs.as<WaveJob>(), (uint32_t)jobs.size()};
    table_probe = ManyScenes{ctx->jobs.as<WaveJob>() + jobs.size(), (uint32_t)jobs.size()};
  }
  void* src = many ? (void*)&table : (void*)&one;
  void* probe_src = many ? (void*)&table_probe : (void*)&one_probe;
  const uint32_t* order = ordered ? orderByProbe(ctx, probe_fn, probe_grid, probe_src, w, nt, lds) : nullptr;
  launchWave(ctx, fn, grid, src, w, nt, lds, order, nullptr);
  ZT_HIP(hipGetLastError());
}

// Wave-scheduled render of one scene over its two-child tree (Trav2 walkers,
// the reference's traversal order): the fp64 path's default. Variants 1-4
// build for at least 1, 2, 3 or 4 waves per SIMD. Defaults measured on the
// 1080p mesh field: fp64 variant 3 (168 VGPRs and 47 spilled: 3 waves beat
// 2 without spills by 14%) with refill 40; fp32 variant 4 (128 VGPRs, no
// spills) with refill 32.
template <typename R>
static void launchWaveTwoChild(zt_ctx* ctx, const SceneView<R>& view, const Params& pr, R* out,
                               const zt_gpu_tuning& tu) {
  WaveShape w = waveShape(tu);
  if (!tu.variant) w.variant = K<R>::exact ? 3 : 4;
  if (!tu.refill) w.refill = K<R>::exact ? 40 : 32;
  const uint32_t tiles_x = (pr.width + w.tw - 1) / w.tw;
  const uint64_t ntiles = (uint64_t)tiles_x * ((pr.height + w.th - 1) / w.th);
  if (ntiles > 0xFFFFFFFFull - 8) throw Error(ZT_ERR_INVALID_ARGUMENT, "too many tiles in one launch");
  static const void* const kFn[4] = {(const void*)megakernelWave2<R, 1>, (const void*)megakernelWave2<R, 2>,
                                     (const void*)megakernelWave2<R, 3>, (const void*)megakernelWave2<R, 4>};
  const void* fn = kFn[w.variant - 1];
  const void* probe_fn = (const void*)megakernelWave2<R, 2, true>;
  const size_t lds = stackBytes(w.cap, 64) + kLeafPassBytes;
  const uint64_t pixels = (uint64_t)pr.width * pr.height;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, residentBlocks(fn, 64, lds));
  const bool ordered = orderTilesFor(tu, false, pixels, pixels * pr.spp, ntiles, grid);
  const uint32_t probe_grid = ordered ? (uint32_t)std::min<uint64_t>(ntiles, residentBlocks(probe_fn, 64, lds)) : 0;
  ctx->spill.ensure(spillBytes((uint64_t)std::max(grid, probe_grid) * 64));
  ctx->bands.ensure(8 * sizeof(uint32_t));
  const uint32_t nt = (uint32_t)ntiles;
  Params probe = pr;  // one sample, extension rays only
  probe.spp = 1;
  probe.use_nee = 0;
  OneSceneT<R> src{view, pr, out, tiles_x}, probe_src{view, probe, out, tiles_x};
  const uint32_t* order = ordered ? orderByProbe(ctx, probe_fn, probe_grid, &probe_src, w, nt, lds) : nullptr;
  launchWave(ctx, fn, grid, &src, w, nt, lds, order, nullptr);
  ZT_HIP(hipGetLastError());
}

const uint32_t* probeTileOrder(zt_ctx* ctx, const SceneView<float>& view, const Params& pr, float* out, int cap,
                               uint32_t refill) {
  zt_gpu_tuning tu{};
  tu.stack_entries = (uint32_t)cap;
  tu.refill = refill;
  const WaveShape w = waveShape(tu);  // 8x8 tiles
  const uint32_t tiles_x = (pr.width + w.tw - 1) / w.tw;
  const uint64_t ntiles = (uint64_t)tiles_x * ((pr.height + w.th - 1) / w.th);
  if (ntiles > 0xFFFFFFFFull - 8) throw Error(ZT_ERR_INVALID_ARGUMENT, "too many tiles in one launch");
  const bool quant = view.nodes4q != nullptr;
  const void* fn = quant ? (const void*)megakernelWave<4, true, true> : (const void*)megakernelWave<4, false, true>;
  const size_t lds = stackBytes(cap, 64) + kLeafPassBytes;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, residentBlocks(fn, 64, lds));
  ctx->spill.ensure(spillBytes((uint64_t)grid * 64));
  ctx->bands.ensure(8 * sizeof(uint32_t));
  Params probe = pr;  // one sample, extension rays only
  probe.spp = 1;
  probe.use_nee = 0;
  OneScene src{view, probe, out, tiles_x};
  return orderByProbe(ctx, fn, grid, &src, w, (uint32_t)ntiles, lds);
}

uint32_t residentBlocks(const void* kernel, uint32_t threads, size_t lds) {
  int dev = 0, cus = 0, per = 0;
  ZT_HIP(hipGetDevice(&dev));
  ZT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  ZT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, (int)threads, lds));
  return (uint32_t)std::max(1, cus * std::max(per, 1));
}

template <typename R>
void launchMegakernel(zt_ctx* ctx, const DevScene& s, const zt_render_options& o, const zt_gpu_tuning& tu, R* dout) {
  SceneView<R> view = viewOf<R>(s);
  view.cam = cameraInit(s.camera, (double)o.width / (double)o.height);  // render.zig:157,162
  applyBvhWidth(view, s, tu);
  Params pr = makeParams(s, o);
  // One pixel per lane. Defaults measured on the 1080p mesh field (1.3M
  // triangles): 8x8 tiles (one wave each; many tiles per resident block
  // balance the tail); fp32 compiled for 6 waves per SIMD with 8 LDS stack
  // entries per lane, fp64 unconstrained with 16 entries.
  constexpr bool kF64 = K<R>::exact;
  if (tu.schedule > 2) throw Error(ZT_ERR_INVALID_ARGUMENT, "schedule must be 0..2");
  int* err = ctx->errflag.as<int>();
  // Wave-scheduled unless asked otherwise. fp32 over the four-wide tree:
  // defaults measured on the 1080p mesh field, a build for >= 4 waves per
  // SIMD (128 VGPRs, no spills), 12 LDS stack entries, refill 32. Over the
  // two-child tree (fp64, or fp32 without the four-wide tree) the same
  // scheduling with the reference's traversal order.
  if (tu.schedule != 1) {
    if constexpr (!kF64) {
      if (view.nodes4) {
        std::vector<WaveJob> jobs{WaveJob{view, pr, dout, 0, 0}};
        launchWaveJobs(ctx, jobs, tu);
        return;
      }
    }
    if (view.num_nodes > 0) {
      launchWaveTwoChild<R>(ctx, view, pr, dout, tu);
      return;
    }
  }
  const int cap = tu.stack_entries ? (int)tu.stack_entries : (kF64 ? 16 : 8);
  if (cap < 8 || cap > 64 || (cap & 1)) throw Error(ZT_ERR_INVALID_ARGUMENT, "stack entries must be even, 8..64");
  const uint32_t variant = tu.variant ? tu.variant : (kF64 ? 1u : 4u);
  uint32_t tw = tu.tile_w ? tu.tile_w : 8;
  uint32_t th = tu.tile_h ? tu.tile_h : 8;
  const uint32_t threads = tw * th;
  if (threads == 0 || threads > 256 || threads % 64 != 0)
    throw Error(ZT_ERR_INVALID_ARGUMENT, "megakernel tile must hold 64, 128, 192 or 256 pixels");
  const uint32_t tiles_x = (pr.width + tw - 1) / tw;
  const uint32_t tiles_y = (pr.height + th - 1) / th;
  const uint32_t ntiles = tiles_x * tiles_y;
  const size_t lds = stackBytes(cap, (int)threads);
  const void* fn = nullptr;
  switch (variant) {
    case 1: fn = (const void*)megakernel<R, 0>; break;
    case 2: fn = (const void*)megakernel<R, 4>; break;
    case 3: fn = (const void*)megakernel<R, 5>; break;
    case 4: fn = (const void*)megakernel<R, 6>; break;
    default: throw Error(ZT_ERR_INVALID_ARGUMENT, "megakernel variant must be 0..4");
  }
  // Persistent grid: as many blocks as can be resident (never more than tiles).
  const uint32_t grid = std::min<uint32_t>(ntiles, residentBlocks(fn, threads, lds));
  ctx->spill.ensure(spillBytes((uint64_t)grid * threads));
  ctx->bands.ensure(8 * sizeof(uint32_t));
  uint32_t* bands = ctx->bands.as<uint32_t>();
  int* spill = ctx->spill.as<int>();
  hipLaunchKernelGGL(clearBands, dim3(1), dim3(64), 0, ctx->stream, bands);
  switch (variant) {
    case 1:
      hipLaunchKernelGGL((megakernel<R, 0>), dim3(grid), dim3(threads), lds, ctx->stream, view, pr, dout, tw, th,
                         tiles_x, ntiles, cap, spill, bands, err);
      break;
    case 2:
      hipLaunchKernelGGL((megakernel<R, 4>), dim3(grid), dim3(threads), lds, ctx->stream, view, pr, dout, tw, th,
                         tiles_x, ntiles, cap, spill, bands, err);
      break;
    case 3:
      hipLaunchKernelGGL((megakernel<R, 5>), dim3(grid), dim3(threads), lds, ctx->stream, view, pr, dout, tw, th,
                         tiles_x, ntiles, cap, spill, bands, err);
      break;
    default:
      hipLaunchKernelGGL((megakernel<R, 6>), dim3(grid), dim3(threads), lds, ctx->stream, view, pr, dout, tw, th,
                         tiles_x, ntiles, cap, spill, bands, err);
      break;
  }
  ZT_HIP(hipGetLastError());
  ctx->stats.launches += 2;
}
template void launchMegakernel<float>(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&,
                                      float*);
template void launchMegakernel<double>(zt_ctx*, const DevScene&, const zt_render_options&, const zt_gpu_tuning&,
                                       double*);

void resetErr(zt_ctx* ctx) { ZT_HIP(hipMemsetAsync(ctx->errflag.p, 0, kStatusBytes, ctx->stream)); }

void checkErr(zt_ctx* ctx) {
  int w[kStatusBytes / sizeof(int)] = {};
  ZT_HIP(hipMemcpyAsync(w, ctx->errflag.p, kStatusBytes, hipMemcpyDeviceToHost, ctx->stream));
  ZT_HIP(hipStreamSynchronize(ctx->stream));
  uint64_t ext = 0, sh = 0;
  std::memcpy(&ext, w + 2, 8);
  std::memcpy(&sh, w + 4, 8);
  ctx->stats.extend_rays = ext;
  ctx->stats.shadow_rays = sh;
  if (w[0]) throw Error(ZT_ERR_INTERNAL, "device traversal stack overflow");
}

}  // namespace ztg

using namespace ztg;

extern "C" {

int zt_gpu_render(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, const zt_gpu_tuning* tuning,
                  void* out, int out_on_device) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    checkOptions(opts);
    if (!out) throw Error(ZT_ERR_INVALID_ARGUMENT, "null output");
    DevScene& s = sceneOf(ctx, scene_id);
    const zt_gpu_tuning tu = resolveTuning(ctx, s, *opts, tuning);
    setDevice(ctx);
    const bool f64 = tu.precision == ZT_PRECISION_FP64;
    const size_t elem = f64 ? sizeof(double) : sizeof(float);
    const size_t bytes = (size_t)opts->width * opts->height * 3 * elem;
    void* dout = out;
    if (!out_on_device) {
      ctx->out.ensure(bytes);
      dout = ctx->out.p;
    }
    ctx->stats.launches = 0;
    resetErr(ctx);
    ZT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (tu.method == ZT_METHOD_WAVEFRONT) {
      if (f64) launchWavefront<double>(ctx, s, *opts, tu, (double*)dout);
      else launchWavefront<float>(ctx, s, *opts, tu, (float*)dout);
    } else {
      if (f64) launchMegakernel<double>(ctx, s, *opts, tu, (double*)dout);
      else launchMegakernel<float>(ctx, s, *opts, tu, (float*)dout);
    }
    ZT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (!out_on_device) ZT_HIP(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    checkErr(ctx);
    float ms = 0;
    ZT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms = ms;
    ctx->stats.paths = (uint64_t)opts->width * opts->height * opts->spp;
  });
}

int zt_gpu_trace_rays(zt_ctx* ctx, uint32_t scene_id, const zt_ray* rays, uint64_t n, uint64_t* rng_states,
                      uint64_t max_depth, uint32_t light_sampling, uint32_t precision, double* out_rgb) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (n == 0) return;
    if (!rays || !rng_states || !out_rgb) throw Error(ZT_ERR_INVALID_ARGUMENT, "null buffer");
    if (precision > ZT_PRECISION_FP64) throw Error(ZT_ERR_INVALID_ARGUMENT, "bad precision");
    DevScene& s = sceneOf(ctx, scene_id);
    setDevice(ctx);
    zt_render_options o{};
    o.max_depth = max_depth;
    o.light_sampling = light_sampling;
    const Params pr = makeParams(s, o);
    DevBuf drays, dstates, dout;
    drays.alloc(n * sizeof(zt_ray));
    dstates.alloc(n * 32);
    dout.alloc(n * 3 * sizeof(double));
    ZT_HIP(hipMemcpyAsync(drays.p, rays, n * sizeof(zt_ray), hipMemcpyHostToDevice, ctx->stream));
    ZT_HIP(hipMemcpyAsync(dstates.p, rng_states, n * 32, hipMemcpyHostToDevice, ctx->stream));
    resetErr(ctx);
    const uint32_t threads = 256;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + threads - 1) / threads, kRayGridBlocks);
    ctx->spill.ensure(spillBytes((uint64_t)blocks * threads));
    if (precision == ZT_PRECISION_FP64)
      hipLaunchKernelGGL(traceKernel<double>, dim3(blocks), dim3(threads), ldsBytes(threads), ctx->stream,
                         viewOf<double>(s), pr, drays.as<zt_ray>(), n, dstates.as<uint64_t>(), dout.as<double>(),
                         ctx->spill.as<int>(), ctx->errflag.as<int>());
    else
      hipLaunchKernelGGL(traceKernel<float>, dim3(blocks), dim3(threads), ldsBytes(threads), ctx->stream,
                         viewOf<float>(s), pr, drays.as<zt_ray>(), n, dstates.as<uint64_t>(), dout.as<double>(),
                         ctx->spill.as<int>(), ctx->errflag.as<int>());
    ZT_HIP(hipGetLastError());
    ZT_HIP(hipMemcpyAsync(out_rgb, dout.p, n * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ZT_HIP(hipMemcpyAsync(rng_states, dstates.p, n * 32, hipMemcpyDeviceToHost, ctx->stream));
    checkErr(ctx);
  });
}

int zt_gpu_intersect(zt_ctx* ctx, uint32_t scene_id, const zt_ray* rays, uint64_t n, double t_min, double t_max,
                     int mode, uint32_t precision, zt_hit* out) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (n == 0) return;
    if (!rays || !out) throw Error(ZT_ERR_INVALID_ARGUMENT, "null buffer");
    if (precision > ZT_PRECISION_FP64 || mode < 0 || mode > 3) throw Error(ZT_ERR_INVALID_ARGUMENT, "bad enum");
    DevScene& s = sceneOf(ctx, scene_id);
    setDevice(ctx);
    DevBuf drays, dout;
    drays.alloc(n * sizeof(zt_ray));
    dout.alloc(n * sizeof(zt_hit));
    ZT_HIP(hipMemcpyAsync(drays.p, rays, n * sizeof(zt_ray), hipMemcpyHostToDevice, ctx->stream));
    resetErr(ctx);
    const uint32_t threads = 256;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + threads - 1) / threads, kRayGridBlocks);
    ctx->spill.ensure(spillBytes((uint64_t)blocks * threads));
    SceneView<float> viewQ = viewOf<float>(s);
    if (mode == 3 && viewQ.nodes4 && s.nodes4q.p) viewQ.nodes4q = s.nodes4q.as<DNode4Q>();  // 64-byte nodes
    if (precision == ZT_PRECISION_FP64)
      hipLaunchKernelGGL(intersectKernel<double>, dim3(blocks), dim3(threads), ldsBytes(threads), ctx->stream,
                         viewOf<double>(s), drays.as<zt_ray>(), n, t_min, t_max, mode, dout.as<zt_hit>(),
                         ctx->spill.as<int>(), ctx->errflag.as<int>());
    else
      hipLaunchKernelGGL(intersectKernel<float>, dim3(blocks), dim3(threads), ldsBytes(threads), ctx->stream,
                         viewQ, drays.as<zt_ray>(), n, (float)t_min, (float)t_max, mode,
                         dout.as<zt_hit>(), ctx->spill.as<int>(), ctx->errflag.as<int>());
    ZT_HIP(hipGetLastError());
    ZT_HIP(hipMemcpyAsync(out, dout.p, n * sizeof(zt_hit), hipMemcpyDeviceToHost, ctx->stream));
    checkErr(ctx);
  });
}

int zt_gpu_rng(zt_ctx* ctx, uint64_t seed, const uint64_t* streams, uint64_t n, uint32_t draws, int kind,
               uint64_t* out) {
  if (!ctx) return ZT_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (n == 0 || draws == 0) return;
    if (!streams || !out || kind < 0 || kind > 1) throw Error(ZT_ERR_INVALID_ARGUMENT, "bad argument");
    setDevice(ctx);
    DevBuf ds, dout;
    ds.alloc(n * 8);
    dout.alloc(n * draws * 8);
    ZT_HIP(hipMemcpyAsync(ds.p, streams, n * 8, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t threads = 256;
    hipLaunchKernelGGL(rngKernel, dim3((uint32_t)((n + threads - 1) / threads)), dim3(threads), 0, ctx->stream, seed,
                       ds.as<uint64_t>(), n, draws, kind, dout.as<uint64_t>());
    ZT_HIP(hipGetLastError());
    ZT_HIP(hipMemcpyAsync(out, dout.p, n * draws * 8, hipMemcpyDeviceToHost, ctx->stream));
    ZT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
