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
//
// Parallel form (threads > 1): the split of a node depends only on the
// primitives in its index range, and the two children own disjoint ranges,
// so subtrees are built concurrently into node records; a final preorder
// walk (left first) then numbers the records exactly as the sequential build
// appends them (a node's children as a pair when the node is split), so the
// node array and index permutation are the same bits for any thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <limits>
#include <mutex>
#include <thread>
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

  unsigned pool = 1;  // threads a very large range may split its reductions over

  // Runs f(begin, end, part) over `parts` slices of [first, first + count) on
  // their own threads. Used only for reductions whose merge is exact in any
  // order (box min/max and counts), so the result does not depend on `parts`.
  template <typename F>
  void slices(uint32_t first, uint32_t count, unsigned parts, F&& f) const {
    std::vector<std::thread> th;
    for (unsigned k = 0; k < parts; ++k) {
      const uint32_t b = first + (uint32_t)((uint64_t)count * k / parts);
      const uint32_t e = first + (uint32_t)((uint64_t)count * (k + 1) / parts);
      if (k + 1 == parts) f(b, e, k);
      else th.emplace_back([&f, b, e, k] { f(b, e, k); });
    }
    for (auto& t : th) t.join();
  }
  unsigned partsFor(uint32_t count) const { return count >= (1u << 18) ? pool : 1; }

  bool bestSplit(uint32_t first, uint32_t count, const Box& bounds, const Box& cb, Split* best) const {
    const double parent = area(bounds);
    if (parent <= 0) return false;
    bool found = false;
    const unsigned parts = partsFor(count);
    for (int axis = 0; axis < 3; ++axis) {
      const double lo = cb.lo[axis];
      const double extent = cb.hi[axis] - lo;
      if (!(extent > 0)) continue;
      Box bb[kBins];
      uint64_t bc[kBins] = {};
      for (auto& b : bb) b = emptyBox();
      const double* c = cent[axis].data();
      auto binRange = [&](uint32_t b0, uint32_t e0, Box* lb, uint64_t* lc) {
        for (uint32_t j = b0; j < e0; ++j) {
          const uint32_t p = idx[j];
          const int b = binIndex(c[p], lo, extent);
          lc[b] += 1;
          grow(lb[b], boxes[p]);
        }
      };
      if (parts == 1) {
        binRange(first, first + count, bb, bc);
      } else {
        std::vector<Box> pbb((size_t)parts * kBins, emptyBox());
        std::vector<uint64_t> pbc((size_t)parts * kBins, 0);
        slices(first, count, parts, [&](uint32_t b0, uint32_t e0, unsigned k) {
          binRange(b0, e0, pbb.data() + (size_t)k * kBins, pbc.data() + (size_t)k * kBins);
        });
        for (unsigned k = 0; k < parts; ++k)
          for (int b = 0; b < kBins; ++b) {
            bc[b] += pbc[(size_t)k * kBins + b];
            grow(bb[b], pbb[(size_t)k * kBins + b]);
          }
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

  // Bounds of a node's range, and the split the reference would make:
  // returns false for a leaf, else the size of the left part (the range is
  // partitioned in place exactly as bvh.zig does).
  bool splitNode(uint32_t first, uint32_t count, int level, Box* bounds_out, uint32_t* left_count) {
    Box bounds = emptyBox(), cb = emptyBox();
    auto boundRange = [&](uint32_t b0, uint32_t e0, Box* bo, Box* co) {
      for (uint32_t j = b0; j < e0; ++j) {
        const uint32_t p = idx[j];
        grow(*bo, boxes[p]);
        const double c3[3] = {cent[0][p], cent[1][p], cent[2][p]};
        growPoint(*co, c3);
      }
    };
    const unsigned parts = partsFor(count);
    if (parts == 1) {
      boundRange(first, first + count, &bounds, &cb);
    } else {
      std::vector<Box> pb(parts, emptyBox()), pc(parts, emptyBox());
      slices(first, count, parts, [&](uint32_t b0, uint32_t e0, unsigned k) { boundRange(b0, e0, &pb[k], &pc[k]); });
      for (unsigned k = 0; k < parts; ++k) {
        grow(bounds, pb[k]);
        grow(cb, pc[k]);
      }
    }
    *bounds_out = bounds;
    if (count <= 1 || level >= kMaxLevel) return false;
    Split sp;
    if (!bestSplit(first, count, bounds, cb, &sp)) {
      if (count <= kMaxLeaf) return false;
      // All centroids coincide: stable sort on the longest axis, halve.
      const double* c = cent[longest(cb)].data();
      std::stable_sort(idx + first, idx + first + count, [c](uint32_t a, uint32_t b) { return c[a] < c[b]; });
      *left_count = count / 2;
      return true;
    }
    if (count <= kMaxLeaf && sp.cost >= (double)count) return false;
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
    *left_count = mid;
    return true;
  }

  // Parallel build. Records are allocated in pairs from an atomic counter;
  // rec.child = first record of the children pair, or -1 for a leaf.
  struct Rec {
    Box bounds;
    uint32_t first, count;
    int64_t child;
  };
  void runParallel(uint32_t n, unsigned threads) {
    std::vector<Rec> recs(2 * (size_t)n - 1);
    std::atomic<uint64_t> next{1};
    recs[0] = Rec{emptyBox(), 0, n, -1};
    struct Task {
      uint64_t rec;
      int level;
    };
    std::deque<Task> queue{{0, 0}};
    std::mutex mu;
    std::condition_variable cv;
    uint64_t outstanding = 1;  // tasks queued or running
    // Ranges at least this large hand their right child to the queue.
    const uint32_t kSpawn = std::max<uint32_t>(4096, n / (64 * threads));
    auto work = [&]() {
      std::vector<Task> local;
      for (;;) {
        Task t;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return !queue.empty() || outstanding == 0; });
          if (queue.empty()) return;
          t = queue.front();
          queue.pop_front();
        }
        local.push_back(t);
        while (!local.empty()) {
          const Task w = local.back();
          local.pop_back();
          Rec& r = recs[w.rec];
          uint32_t left_count = 0;
          if (!splitNode(r.first, r.count, w.level, &r.bounds, &left_count)) continue;
          const uint64_t c = next.fetch_add(2);
          recs[c] = Rec{emptyBox(), r.first, left_count, -1};
          recs[c + 1] = Rec{emptyBox(), r.first + left_count, r.count - left_count, -1};
          r.child = (int64_t)c;
          if (r.count - left_count >= kSpawn) {
            std::lock_guard<std::mutex> lk(mu);
            queue.push_back({c + 1, w.level + 1});
            ++outstanding;
            cv.notify_one();
          } else {
            local.push_back({c + 1, w.level + 1});
          }
          local.push_back({c, w.level + 1});
        }
        {
          std::lock_guard<std::mutex> lk(mu);
          if (--outstanding == 0) cv.notify_all();
        }
      }
    };
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < threads; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    // Number the records as the sequential build appends nodes.
    nodes->push_back(zt_bvh_node{});
    std::vector<std::pair<uint64_t, uint32_t>> st{{0, 0}};  // (record, node index)
    while (!st.empty()) {
      const auto [ri, ni] = st.back();
      st.pop_back();
      const Rec& r = recs[ri];
      zt_bvh_node nd{};
      for (int i = 0; i < 3; ++i) {
        nd.min[i] = r.bounds.lo[i];
        nd.max[i] = r.bounds.hi[i];
      }
      if (r.child < 0) {
        nd.first = r.first;
        nd.count = r.count;
        (*nodes)[ni] = nd;
        continue;
      }
      const uint32_t left = (uint32_t)nodes->size();
      nodes->push_back(zt_bvh_node{});
      nodes->push_back(zt_bvh_node{});
      nd.first = left;
      nd.count = 0;
      (*nodes)[ni] = nd;
      st.push_back({(uint64_t)r.child + 1, left + 1});
      st.push_back({(uint64_t)r.child, left});
    }
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
      Box bounds;
      uint32_t left_count = 0;
      const bool split = splitNode(first, count, w.level, &bounds, &left_count);
      zt_bvh_node& nd = (*nodes)[w.node];
      for (int i = 0; i < 3; ++i) {
        nd.min[i] = bounds.lo[i];
        nd.max[i] = bounds.hi[i];
      }
      if (!split) continue;
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

int buildBvh(const zt_shape* shapes, uint64_t n, std::vector<zt_bvh_node>* nodes, std::vector<uint32_t>* indices,
             unsigned threads) {
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
  if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  b.pool = threads;
  if (threads > 1 && n >= 65536) b.runParallel((uint32_t)n, threads);
  else b.run((uint32_t)n);
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
