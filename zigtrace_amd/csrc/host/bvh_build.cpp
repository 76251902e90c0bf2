// Host BVH build: binned SAH over 12 bins, leaves of at most 4 primitives,
// depth cap 64, median fallback when binning cannot separate centroids. It
// must emit the reference's node array bit for bit (bvh.zig:41-262), so it
// repeats the reference's arithmetic exactly:
//   * NaN-ignoring min/max (fmin/fmax) in box merges;
//   * no FMA contraction (-ffp-contract=off) and the same operation order in
//     the SAH cost 1 + (A_L*N_L + A_R*N_R) / A_P;
//   * bin = trunc(clamp((c - lo) / extent * 12, 0, 11));
//   * the in-place swap partition, a stable median sort, and children
//     appended as a pair with the left subtree finished before the right.
// The recursion of the reference is replaced by an explicit stack that pops
// the left child first, which produces the same node numbering.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "host_scene.hpp"

namespace zth {

namespace {

constexpr int kBins = 12;
constexpr uint32_t kMaxLeaf = 4;
constexpr int kMaxLevel = 64;
const double kInf = std::numeric_limits<double>::infinity();

struct Box {
  double lo[3], hi[3];
};
inline Box emptyBox() { return Box{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}}; }
inline void grow(Box& a, const Box& b) {
  for (int i = 0; i < 3; ++i) {
    a.lo[i] = std::fmin(a.lo[i], b.lo[i]);
    a.hi[i] = std::fmax(a.hi[i], b.hi[i]);
  }
}
inline void growPoint(Box& a, const double* p) {
  for (int i = 0; i < 3; ++i) {
    a.lo[i] = std::fmin(a.lo[i], p[i]);
    a.hi[i] = std::fmax(a.hi[i], p[i]);
  }
}
inline double area(const Box& b) {  // aabb.zig:53-57
  if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) return 0;
  const double ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
  return 2.0 * (ex * ey + ey * ez + ez * ex);
}
inline int longest(const Box& b) {  // aabb.zig:59-63
  const double ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
  if (ex >= ey && ex >= ez) return 0;
  return ey >= ez ? 1 : 2;
}
inline int binIndex(double c, double lo, double extent) {
  const double f = (c - lo) / extent * (double)kBins;
  return (int)std::fmax(0.0, std::fmin(f, (double)(kBins - 1)));
}

// Shape bounds (geometry.zig:54-57, 101-103; aabb.zig:17-39).
bool shapeBox(const zt_shape& s, Box* out) {
  if (s.kind == ZT_SPHERE) {
    const double r = std::fabs(s.v[3]);
    for (int i = 0; i < 3; ++i) {
      out->lo[i] = s.v[i] - r;
      out->hi[i] = s.v[i] + r;
    }
    return true;
  }
  if (s.kind == ZT_TRIANGLE) {
    Box b;
    for (int i = 0; i < 3; ++i) {
      b.lo[i] = std::fmin(s.v[i], s.v[3 + i]);
      b.hi[i] = std::fmax(s.v[i], s.v[3 + i]);
    }
    growPoint(b, s.v + 6);
    const double delta = 1e-6;
    for (int i = 0; i < 3; ++i) {
      if (b.hi[i] - b.lo[i] < delta) {
        b.lo[i] -= delta / 2;
        b.hi[i] += delta / 2;
      }
    }
    *out = b;
    return true;
  }
  return false;
}

struct Builder {
  std::vector<Box> boxes;
  std::vector<double> cent[3];  // centroids, one array per axis
  std::vector<zt_bvh_node>* nodes;
  uint32_t* idx;

  struct Split {
    int axis = -1, bin = 0;
    double cost = 0;
  };

  bool bestSplit(uint32_t first, uint32_t count, const Box& bounds, const Box& cb, Split* best) const {
    const double parent = area(bounds);
    if (parent <= 0) return false;
    bool found = false;
    for (int axis = 0; axis < 3; ++axis) {
      const double lo = cb.lo[axis];
      const double extent = cb.hi[axis] - lo;
      if (!(extent > 0)) continue;
      Box bb[kBins];
      uint64_t bc[kBins] = {};
      for (auto& b : bb) b = emptyBox();
      const double* c = cent[axis].data();
      for (uint32_t j = first; j < first + count; ++j) {
        const uint32_t p = idx[j];
        const int b = binIndex(c[p], lo, extent);
        bc[b] += 1;
        grow(bb[b], boxes[p]);
      }
      double rarea[kBins];
      uint64_t rcount[kBins];
      Box acc = emptyBox();
      uint64_t n = 0;
      for (int b = kBins - 1; b >= 0; --b) {
        grow(acc, bb[b]);
        n += bc[b];
        rarea[b] = area(acc);
        rcount[b] = n;
      }
      acc = emptyBox();
      n = 0;
      for (int b = 0; b + 1 < kBins; ++b) {
        grow(acc, bb[b]);
        n += bc[b];
        const uint64_t rn = rcount[b + 1];
        if (n == 0 || rn == 0) continue;
        const double cost = 1.0 + (area(acc) * (double)n + rarea[b + 1] * (double)rn) / parent;
        if (!found || cost < best->cost) {
          best->axis = axis;
          best->bin = b;
          best->cost = cost;
          found = true;
        }
      }
    }
    return found;
  }

  void run(uint32_t n) {
    struct Work {
      uint32_t node;
      int level;
    };
    std::vector<Work> stack;
    nodes->push_back(zt_bvh_node{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}, 0, n});
    stack.push_back({0, 0});
    while (!stack.empty()) {
      const Work w = stack.back();
      stack.pop_back();
      const uint32_t first = (*nodes)[w.node].first;
      const uint32_t count = (*nodes)[w.node].count;
      Box bounds = emptyBox(), cb = emptyBox();
      for (uint32_t j = first; j < first + count; ++j) {
        const uint32_t p = idx[j];
        grow(bounds, boxes[p]);
        const double c3[3] = {cent[0][p], cent[1][p], cent[2][p]};
        growPoint(cb, c3);
      }
      zt_bvh_node& nd = (*nodes)[w.node];
      for (int i = 0; i < 3; ++i) {
        nd.min[i] = bounds.lo[i];
        nd.max[i] = bounds.hi[i];
      }
      if (count <= 1 || w.level >= kMaxLevel) continue;
      Split sp;
      uint32_t left_count;
      if (!bestSplit(first, count, bounds, cb, &sp)) {
        if (count <= kMaxLeaf) continue;
        // All centroids coincide: stable sort on the longest axis, halve.
        const double* c = cent[longest(cb)].data();
        std::stable_sort(idx + first, idx + first + count, [c](uint32_t a, uint32_t b) { return c[a] < c[b]; });
        left_count = count / 2;
      } else {
        if (count <= kMaxLeaf && sp.cost >= (double)count) continue;
        const double lo = cb.lo[sp.axis];
        const double extent = cb.hi[sp.axis] - lo;
        const double* c = cent[sp.axis].data();
        uint32_t mid = 0;
        for (uint32_t j = 0; j < count; ++j) {
          const uint32_t p = idx[first + j];
          if (binIndex(c[p], lo, extent) <= sp.bin) {
            std::swap(idx[first + mid], idx[first + j]);
            ++mid;
          }
        }
        left_count = mid;
      }
      const uint32_t left = (uint32_t)nodes->size();
      nodes->push_back(zt_bvh_node{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}, first, left_count});
      nodes->push_back(zt_bvh_node{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}, first + left_count, count - left_count});
      (*nodes)[w.node].first = left;
      (*nodes)[w.node].count = 0;
      stack.push_back({left + 1, w.level + 1});
      stack.push_back({left, w.level + 1});
    }
  }
};

}  // namespace

int buildBvh(const zt_shape* shapes, uint64_t n, std::vector<zt_bvh_node>* nodes, std::vector<uint32_t>* indices) {
  nodes->clear();
  indices->clear();
  if (n == 0) return ZT_OK;
  if (n > 0xFFFFFFFFull / 2) return ZT_ERR_TOO_MANY_PRIMITIVES;
  Builder b;
  b.boxes.resize(n);
  for (auto& c : b.cent) c.resize(n);
  indices->resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    (*indices)[i] = (uint32_t)i;
    if (!shapeBox(shapes[i], &b.boxes[i])) {
      indices->clear();
      return ZT_ERR_UNBOUNDED_SHAPE;
    }
    for (int a = 0; a < 3; ++a) b.cent[a][i] = (b.boxes[i].lo[a] + b.boxes[i].hi[a]) * 0.5;
  }
  nodes->reserve(2 * n - 1);
  b.nodes = nodes;
  b.idx = indices->data();
  b.run((uint32_t)n);
  return ZT_OK;
}

uint64_t bvhDepth(const std::vector<zt_bvh_node>& nodes) {  // bvh.zig:128-137, iteratively
  if (nodes.empty()) return 0;
  uint64_t best = 0;
  std::vector<std::pair<uint32_t, uint64_t>> st{{0u, 1ull}};
  while (!st.empty()) {
    const auto [i, d] = st.back();
    st.pop_back();
    if (nodes[i].count > 0) {
      best = std::max(best, d);
    } else {
      st.push_back({nodes[i].first, d + 1});
      st.push_back({nodes[i].first + 1, d + 1});
    }
  }
  return best;
}

}  // namespace zth
