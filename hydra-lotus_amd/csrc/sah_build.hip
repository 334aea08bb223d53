// Host construction of the binary tree's topology by the surface area
// heuristic, the slower, higher-quality alternative to the Morton order and
// radix tree of bvh_build.hip (the reference asks its driver for
// PREFER_FAST_TRACE builds, gpu_scene.cpp:1899-1950, whose algorithm is the
// driver's). Top-down binned SAH (Wald, "On fast Construction of SAH-based
// Bounding Volume Hierarchies", RT 2007): 32 bins per axis over the
// centroids, split cost = A(left) N(left) + A(right) N(right); a node of at
// most four triangles, or whose centroids coincide, splits at the median.
// The result has bvh_build.hip's binary format, so its boxes, sparse table
// and 4-wide collapse run on the GPU unchanged: a leaf order (`order`, the
// flattened triangle at each leaf position), internal nodes 0 .. n-2 (the
// root 0), children (`child`, a leaf at position i is n - 1 + i) and each
// internal node's leaf range (`range`, first and last position, every
// subtree a contiguous range of the order).
#include "internal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <thread>
#include <vector>

namespace lotus {
namespace {

constexpr int kBins = 32;
constexpr uint32_t kMedianBelow = 4;       // at most this many triangles: median split
constexpr uint32_t kParallelAbove = 16384;  // larger subtrees go back to the shared task list

struct Bounds {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const float* p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  void grow(const Bounds& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  double half_area() const {
    if (!(hi[0] >= lo[0])) return 0.0;
    const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    return x * y + y * z + z * x;
  }
};

struct Task {
  uint32_t node, begin, end;  // internal node id, leaf positions [begin, end)
};

struct Builder {
  const std::vector<Bounds>& box;
  const std::vector<float>& cen;  // 3 per triangle
  uint32_t n;
  uint32_t* order;
  uint32_t* child;
  uint32_t* range;  // 2 per internal node
  std::atomic<uint32_t> next_id{1};
  std::mutex lock;
  std::vector<Task> shared;
  std::atomic<uint32_t> busy{0};

  Builder(const std::vector<Bounds>& b, const std::vector<float>& c, uint32_t count, uint32_t* o, uint32_t* ch,
          uint32_t* r)
      : box(b), cen(c), n(count), order(o), child(ch), range(r) {}

  uint32_t child_id(uint32_t begin, uint32_t count) {
    return count == 1 ? n - 1 + begin : next_id.fetch_add(1);
  }

  // Splits [begin, end) into two non-empty halves; returns the split point.
  uint32_t split(uint32_t begin, uint32_t end) {
    const uint32_t count = end - begin;
    Bounds cb;
    for (uint32_t i = begin; i < end; ++i) cb.grow(&cen[3ull * order[i]]);
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
    auto median = [&](int ax) {
      const uint32_t mid = begin + count / 2;
      std::nth_element(order + begin, order + mid, order + end, [&](uint32_t x, uint32_t y) {
        const float cx = cen[3ull * x + ax], cy = cen[3ull * y + ax];
        return cx < cy || (cx == cy && x < y);
      });
      return mid;
    };
    if (count <= kMedianBelow || !(cb.hi[axis] > cb.lo[axis])) return median(axis);
    double best = INFINITY;
    int best_axis = -1, best_plane = 0;
    for (int a = 0; a < 3; ++a) {
      const float extent = cb.hi[a] - cb.lo[a];
      if (!(extent > 0.0f)) continue;
      const float scale = kBins / extent;
      Bounds bins[kBins];
      uint32_t counts[kBins] = {};
      for (uint32_t i = begin; i < end; ++i) {
        const uint32_t t = order[i];
        const int b = std::min(kBins - 1, static_cast<int>((cen[3ull * t + a] - cb.lo[a]) * scale));
        ++counts[b];
        bins[b].grow(box[t]);
      }
      double right_area[kBins];
      uint32_t right_count[kBins];
      Bounds acc;
      uint32_t c = 0;
      for (int b = kBins - 1; b > 0; --b) {
        acc.grow(bins[b]);
        c += counts[b];
        right_area[b] = acc.half_area();
        right_count[b] = c;
      }
      Bounds left;
      uint32_t lc = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        left.grow(bins[b]);
        lc += counts[b];
        const uint32_t rc = right_count[b + 1];
        if (lc == 0 || rc == 0) continue;
        const double cost = left.half_area() * lc + right_area[b + 1] * rc;
        if (cost < best) {
          best = cost;
          best_axis = a;
          best_plane = b + 1;
        }
      }
    }
    if (best_axis < 0) return median(axis);
    const float lo = cb.lo[best_axis], scale = kBins / (cb.hi[best_axis] - cb.lo[best_axis]);
    uint32_t* mid = std::partition(order + begin, order + end, [&](uint32_t t) {
      return std::min(kBins - 1, static_cast<int>((cen[3ull * t + best_axis] - lo) * scale)) < best_plane;
    });
    const uint32_t m = static_cast<uint32_t>(mid - order);
    return (m == begin || m == end) ? median(axis) : m;
  }

  void run(Task t, std::vector<Task>& local) {
    local.push_back(t);
    while (!local.empty()) {
      const Task k = local.back();
      local.pop_back();
      const uint32_t m = split(k.begin, k.end);
      const uint32_t lc = m - k.begin, rc = k.end - m;
      const uint32_t l = child_id(k.begin, lc), r = child_id(m, rc);
      child[2ull * k.node] = l;
      child[2ull * k.node + 1] = r;
      range[2ull * k.node] = k.begin;
      range[2ull * k.node + 1] = k.end - 1;
      for (const Task c : {Task{l, k.begin, m}, Task{r, m, k.end}}) {
        if (c.end - c.begin < 2) continue;
        if (c.end - c.begin > kParallelAbove) {
          std::lock_guard<std::mutex> g(lock);
          shared.push_back(c);
        } else {
          local.push_back(c);
        }
      }
    }
  }

  void work() {
    std::vector<Task> local;
    while (true) {
      Task t{};
      bool got = false;
      {
        std::lock_guard<std::mutex> g(lock);
        if (!shared.empty()) {
          t = shared.back();
          shared.pop_back();
          got = true;
          ++busy;
        } else if (busy == 0) {
          return;
        }
      }
      if (!got) {
        std::this_thread::yield();
        continue;
      }
      run(t, local);
      --busy;
    }
  }
};

}  // namespace

bool sah_topology(const Tri* world, uint32_t n, uint32_t* order, uint32_t* child, uint32_t* range,
                  std::string& error) {
  if (n < 2) {
    if (n == 1) order[0] = 0;
    return true;
  }
  std::vector<Bounds> box(n);
  std::vector<float> cen(3ull * n);
  for (uint32_t i = 0; i < n; ++i) {
    const Tri& t = world[i];
    const float p[3][3] = {{t.a.x, t.a.y, t.a.z}, {t.b.x, t.b.y, t.b.z}, {t.c.x, t.c.y, t.c.z}};
    for (int k = 0; k < 3; ++k) box[i].grow(p[k]);
    for (int a = 0; a < 3; ++a) cen[3ull * i + a] = (p[0][a] + p[1][a] + p[2][a]) * (1.0f / 3.0f);
    order[i] = i;
  }
  Builder b(box, cen, n, order, child, range);
  b.shared.push_back(Task{0u, 0u, n});
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned k = 1; k < threads; ++k) pool.emplace_back([&] { b.work(); });
  b.work();
  for (auto& th : pool) th.join();
  if (b.next_id.load() != n - 1) {
    error = "SAH build: " + std::to_string(b.next_id.load()) + " internal nodes for " + std::to_string(n) +
            " triangles";
    return false;
  }
  return true;
}

}  // namespace lotus
