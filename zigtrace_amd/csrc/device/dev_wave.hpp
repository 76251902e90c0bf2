// Wave-level pieces of the kernels that walk a BVH with one ray per lane
// under wave-uniform passes (the wave-scheduled megakernel). `Src`
// gives the scene view of a lane's scene index, `src.view(si)`, and says
// whether every lane renders the same scene (`Src::kSingle`). The
// workgroup must be exactly one wave: the leaf pass orders its LDS traffic
// with __syncthreads, which is then a wave-local barrier that every pass
// reaches together.
#pragma once

#include "dev_trace.hpp"

namespace ztd {

// LDS of the wave-parallel leaf pass, after the traversal stacks: one 64-bit
// best key per lane, then the owner lane of each dealt-out primitive (a held
// leaf has at most 15 primitives).
constexpr size_t kLeafKeyBytes = 64 * sizeof(unsigned long long);
constexpr size_t kLeafOwnerBytes = 64 * 16;
constexpr size_t kLeafPassBytes = kLeafKeyBytes + kLeafOwnerBytes;

// One leaf pass over the whole wave. Every lane holding a leaf deals its
// primitives out: item w of the lanes' concatenated primitive lists (in lane
// order) is tested by lane w % 64 against its owner's ray and the interval
// (tmin, closest) the owner had when the pass began; the owner then keeps the
// least (t, primitive index) of its hits. That is exactly what the
// sequential loop of leafHits returns (a hit must beat the running closest
// strictly, so of equal t the earlier primitive wins, and a later hit
// beyond an earlier one is rejected either way). An any-hit query ends on
// any hit. Returns true for a lane whose any-hit query ended. Called by all
// 64 lanes together (the pass is wave-uniform).
template <typename Src>
ZT_D bool leafPassWave(const Src& src, uint32_t si, Trav4& tr, const Query4& q, bool at_leaf,
                       unsigned long long* keys, uint8_t* owner_of) {
  const uint32_t lane = threadIdx.x;
  int first = 0, count = 0;
  if (at_leaf) {
    const int v = -1 - tr.cur;
    first = v >> 4;
    count = v & 15;
  }
  // Exclusive prefix sum of the counts over lanes, and their total (counts
  // are below 16: one ballot per bit).
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t prefix = 0, total = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint64_t m = __ballot((count >> b) & 1);
    prefix += (uint32_t)__popcll(m & below) << b;
    total += (uint32_t)__popcll(m) << b;
  }
  for (int j = 0; j < count; ++j) owner_of[prefix + j] = (uint8_t)lane;
  keys[lane] = ~0ull;
  __syncthreads();
  const int base = first - (int)prefix;  // primitive of item w = owner's base + w
  const int self2 = q.self * 2 + (q.anyhit ? 1 : 0);
  for (uint32_t r = 0; r < total; r += 64) {
    const uint32_t w = r + lane;
    const bool live = w < total;
    const int own = live ? (int)owner_of[w] : (int)lane;
    const V3<float> o = mk(__shfl(q.o.x, own), __shfl(q.o.y, own), __shfl(q.o.z, own));
    const V3<float> d = mk(__shfl(q.d.x, own), __shfl(q.d.y, own), __shfl(q.d.z, own));
    const float tmax = __shfl(tr.closest, own);
    const int self = __shfl(self2, own) >> 1;
    const int excl = __shfl(q.excl, own);
    const int k = __shfl(base, own) + (int)w;
    const uint32_t osi = Src::kSingle ? 0u : __shfl(si, own);
    float t;
    if (live && primHit<float>(src.view(osi), k, o, d, q.tmin, tmax, self, excl, &t))
      atomicMin(&keys[own], ((unsigned long long)__float_as_uint(t) << 32) | (uint32_t)k);
  }
  __syncthreads();
  if (!at_leaf) return false;
  tr.cur = kNoChild;
  const unsigned long long key = keys[lane];
  if (key == ~0ull) return false;
  tr.closest = __uint_as_float((uint32_t)(key >> 32));  // t > tmin > 0: bit order is value order
  tr.best = (int)(uint32_t)key;
  if (!q.anyhit) return false;
  tr.found = true;
  return true;
}

// One wave-uniform traversal pass over the lanes' walkers (`tracing` lanes
// hold a query): a leaf pass once the lanes holding an untested leaf
// (n_leaf) are at least half as many as those walking interior nodes
// (n_node) -- the leaf pass deals the primitives out to every lane, so it
// pays to run it early -- otherwise a node pass (pops, then one node per
// lane), so no pass pays for the other's code. Returns true for a lane whose
// query ended in this pass (then `tr.finish` gives the hit).
template <bool Q, typename Src>
ZT_D bool wavePass(const Src& src, uint32_t si, Trav4& tr, const Query4& q, bool tracing, bool at_leaf,
                   uint32_t n_leaf, uint32_t n_node, const LaneStack& stk, int* overflow, unsigned long long* keys,
                   uint8_t* owner_of) {
  if (2 * n_leaf >= n_node) return leafPassWave(src, si, tr, q, at_leaf, keys, owner_of);
  if (tracing && !at_leaf) return tr.template inner<Q>(src.view(si), stk, q, overflow);
  return false;
}

// The same passes over two-child walkers (Trav2), with the same rule for a
// leaf pass; each held leaf is tested in its own lane (the tree's leaves are
// small, at most 4 primitives below the depth limit, and tested in the
// reference's order).
template <bool Q, typename R, typename Src>
ZT_D bool wavePass(const Src& src, uint32_t si, Trav2<R>& tr, const QueryR<R>& q, bool tracing, bool at_leaf,
                   uint32_t n_leaf, uint32_t n_node, const LaneStack& stk, int* overflow, unsigned long long*,
                   uint8_t*) {
  if (2 * n_leaf >= n_node) return at_leaf && tr.leaf(src.view(si), q);
  if (tracing && !at_leaf) return tr.inner(src.view(si), stk, q, overflow);
  return false;
}

}  // namespace ztd
