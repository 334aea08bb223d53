// GPU construction of the wide BVH that replaces the driver's BLAS/TLAS
// (backend/vulkan/gpu_scene.cpp:1899-1950, 2250-2323, 2411-2435):
//   1. flatten instances into world-space triangles (the reference's
//      world_from_object applied as scene.slang:1412-1417 does);
//   2. 30-bit Morton codes of centroids, stable LSD radix sort (8-bit digits,
//      wave64 ballot ranking);
//   3. binary radix tree (Karras, "Maximizing Parallelism in the Construction
//      of BVHs, Octrees, and k-d Trees", HPG 2012); a node's box is the union
//      of the leaf boxes its key range covers, read from a sparse table of
//      power-of-two window unions (two lookups per node, no atomics or
//      device-wide fences);
//   4. top-down collapse into 4-wide nodes, opening the child of largest
//      surface area first, leaves of at most `leaf_size` triangles.
// With LOTUS_HIP_BVH_SAH, steps 2 and 3 are replaced by a binned-SAH
// topology built on the host (sah_build.hip); the boxes and the collapse
// stay here.
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>

namespace lotus {
namespace {

#define LOTUS_CHECK(expr)                                                  \
  do {                                                                     \
    const hipError_t e__ = (expr);                                         \
    if (e__ != hipSuccess) {                                               \
      error = std::string(#expr) + ": " + hipGetErrorString(e__);          \
      return false;                                                        \
    }                                                                      \
  } while (0)

// A start/stop pair of timing events, destroyed on every way out of a build.
struct EventPair {
  hipEvent_t start = nullptr, stop = nullptr;
  hipError_t create() {
    hipError_t e = hipEventCreate(&start);
    if (e == hipSuccess) e = hipEventCreate(&stop);
    return e;
  }
  ~EventPair() {
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
  }
};

constexpr int kBlock = 256;

inline uint32_t blocks_for(uint64_t n, uint32_t per_block) {
  return static_cast<uint32_t>((n + per_block - 1) / per_block);
}

// Float bits mapped so unsigned order is float order.
__device__ __forceinline__ uint32_t ordered(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Flattened triangle g in world space: its instance is the last one whose
// first flattened id is at most g.
__device__ __forceinline__ Tri world_triangle(const BuildInput& in, uint32_t g) {
  uint32_t lo = 0, hi = in.instance_count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (in.instance_base[mid] <= g) lo = mid; else hi = mid;
  }
  const uint32_t inst = lo, prim = g - in.instance_base[inst];
  const uint32_t geo = in.instance_geometry[inst];
  const uint32_t* tri = in.triangles + 3ull * (in.geometry_triangle_offset[geo] + prim);
  const float* pos = in.positions + 3ull * in.geometry_vertex_offset[geo];
  const float* m = in.instance_transform + 16ull * inst;
  float w[3][4];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float* p = pos + 3ull * tri[k];
    mul_point(m, p[0], p[1], p[2], 1.0f, w[k]);
  }
  Tri t;
  t.a = make_float4(w[0][0], w[0][1], w[0][2], __uint_as_float(g));
  t.b = make_float4(w[1][0], w[1][1], w[1][2], __uint_as_float(inst));
  const uint32_t covered = (in.instance_covered && in.instance_covered[inst]) ? kCoveredBit : 0u;
  t.c = make_float4(w[2][0], w[2][1], w[2][2], __uint_as_float(prim | covered));
  return t;
}

// Step 1: world-space triangles in flattened (instance, primitive) order, and
// the centroid bounds (ordered-int atomics: lo[3], hi[3]).
__global__ void __launch_bounds__(kBlock) k_flatten(BuildInput in, Tri* world, float4* centroid,
                                                    uint32_t* bounds) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {0.0f, 0.0f, 0.0f};
  const bool valid = g < in.tri_count;
  if (valid) {
    const Tri t = world_triangle(in, g);
    world[g] = t;
    c[0] = (t.a.x + t.b.x + t.c.x) * (1.0f / 3.0f);
    c[1] = (t.a.y + t.b.y + t.c.y) * (1.0f / 3.0f);
    c[2] = (t.a.z + t.b.z + t.c.z) * (1.0f / 3.0f);
    centroid[g] = make_float4(c[0], c[1], c[2], 0.0f);
  }
  __shared__ uint32_t s_lo[3], s_hi[3];
  if (threadIdx.x < 3) {
    s_lo[threadIdx.x] = 0xFFFFFFFFu;
    s_hi[threadIdx.x] = 0u;
  }
  __syncthreads();
  if (valid) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&s_lo[a], ordered(c[a]));
      atomicMax(&s_hi[a], ordered(c[a]));
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&bounds[threadIdx.x], s_lo[threadIdx.x]);
    atomicMax(&bounds[3 + threadIdx.x], s_hi[threadIdx.x]);
  }
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void __launch_bounds__(kBlock) k_morton(const float4* centroid, const uint32_t* bounds,
                                                   uint32_t n, uint32_t* keys, uint32_t* values) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float4 c = centroid[g];
  const float cc[3] = {c.x, c.y, c.z};
  // One quantum for every axis (the largest extent over 1024): a flat scene
  // keeps few bits along its thin axis, so the top of the tree splits it
  // across its width rather than into overlapping layers.
  float ext = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) ext = fmaxf(ext, unordered(bounds[3 + a]) - unordered(bounds[a]));
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = unordered(bounds[a]);
    const float t = ext > 0.0f ? (cc[a] - lo) / ext : 0.5f;
    q[a] = static_cast<uint32_t>(fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f));
  }
  keys[g] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
  values[g] = g;
}

// ------------------------------------------------------------ radix sort --
constexpr int kSortItems = 8;                      // keys per thread per tile
constexpr uint32_t kSortTile = kBlock * kSortItems;  // 2048

__global__ void __launch_bounds__(kBlock) k_sort_histogram(const uint32_t* keys, uint32_t n, int shift,
                                                           uint32_t tiles, uint32_t* counts) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t k = base + i * kBlock + threadIdx.x;
    if (k < n) atomicAdd(&hist[(keys[k] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of the digit-major counts in two steps: block d scans digit
// d's `tiles` counts in place and writes their total; then every count gains
// the totals of the smaller digits.
__global__ void __launch_bounds__(1024) k_scan_digit(uint32_t* counts, uint32_t tiles, uint32_t* totals) {
  __shared__ uint32_t part[1024];
  __shared__ uint32_t carry;
  uint32_t* seg = counts + size_t(blockIdx.x) * tiles;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  constexpr uint32_t kPer = 4;
  for (uint32_t start = 0; start < tiles; start += 1024 * kPer) {
    uint32_t v[kPer];
    uint32_t sum = 0;
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t k = start + threadIdx.x * kPer + i;
      v[i] = k < tiles ? seg[k] : 0u;
      sum += v[i];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
      const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    uint32_t run = carry + part[threadIdx.x] - sum;
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t k = start + threadIdx.x * kPer + i;
      if (k < tiles) seg[k] = run;
      run += v[i];
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t* counts, uint32_t tiles, const uint32_t* totals) {
  __shared__ uint32_t base;
  if (threadIdx.x == 0) {
    uint32_t b = 0;
    for (uint32_t d = 0; d < blockIdx.x; ++d) b += totals[d];
    base = b;
  }
  __syncthreads();
  uint32_t* seg = counts + size_t(blockIdx.x) * tiles;
  for (uint32_t k = threadIdx.x; k < tiles; k += blockDim.x) seg[k] += base;
}

// Stable scatter: items are ranked in tile order, 256 per round; within a
// wave by ballot matching of the digit bits, across the block's four waves
// by their per-digit counts in LDS.
__global__ void __launch_bounds__(kBlock) k_sort_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                         uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                                         int shift, uint32_t tiles, const uint32_t* offsets) {
  constexpr int kWaves = kBlock / 64;
  __shared__ uint32_t base[256];
  __shared__ uint32_t wave_count[kWaves][256];
  const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid / 64;
  base[tid] = offsets[tid * tiles + blockIdx.x];
  for (int w = 0; w < kWaves; ++w) wave_count[w][tid] = 0;
  __syncthreads();
  const uint32_t start = blockIdx.x * kSortTile;
  for (int round = 0; round < kSortItems; ++round) {
    const uint32_t k = start + round * kBlock + tid;
    const bool valid = k < n;
    const uint32_t key = valid ? keys_in[k] : 0u;
    const uint32_t val = valid ? vals_in[k] : 0u;
    const uint32_t digit = (key >> shift) & 255u;
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(same & ((1ull << lane) - 1ull)));
    const bool leader = valid && rank == 0;
    if (leader) wave_count[wave][digit] = static_cast<uint32_t>(__popcll(same));
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (uint32_t w = 0; w < wave; ++w) before += wave_count[w][digit];
      const uint32_t pos = base[digit] + before + rank;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t total = 0;
    for (int w = 0; w < kWaves; ++w) {
      total += wave_count[w][tid];
      wave_count[w][tid] = 0;
    }
    base[tid] += total;
    __syncthreads();
  }
}

// ------------------------------------------------------ binary radix tree --
__device__ __forceinline__ int delta(const uint32_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j));
  return __clz(a ^ b);
}

// Internal node i in [0, n-2]; leaves are n-1+k. child[2i], child[2i+1];
// range [first, last] of internal nodes.
__global__ void __launch_bounds__(kBlock) k_karras(const uint32_t* keys, int n, uint32_t* child, uint2* range) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) / 2;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int first = i < j ? i : j, last = i < j ? j : i;
  const uint32_t left = first == gamma ? static_cast<uint32_t>(n - 1 + gamma) : static_cast<uint32_t>(gamma);
  const uint32_t right = last == gamma + 1 ? static_cast<uint32_t>(n - 1 + gamma + 1)
                                           : static_cast<uint32_t>(gamma + 1);
  child[2 * i] = left;
  child[2 * i + 1] = right;
  range[i] = make_uint2(static_cast<uint32_t>(first), static_cast<uint32_t>(last));
}

struct Box {
  float lo[3], hi[3];
};

__device__ __forceinline__ Box unite(const Box& a, const Box& b) {
  Box u;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    u.lo[k] = fminf(a.lo[k], b.lo[k]);
    u.hi[k] = fmaxf(a.hi[k], b.hi[k]);
  }
  return u;
}

// Triangles in sorted order, and their boxes: level 0 of the sparse table.
__global__ void __launch_bounds__(kBlock) k_leaf_boxes(const Tri* world, const uint32_t* sorted, uint32_t n,
                                                       Box* table, Tri* tris, uint32_t* tri_of_global) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t g = sorted[k];
  const Tri t = world[g];
  tris[k] = t;
  tri_of_global[g] = k;
  Box b;
  b.lo[0] = fminf(fminf(t.a.x, t.b.x), t.c.x);
  b.lo[1] = fminf(fminf(t.a.y, t.b.y), t.c.y);
  b.lo[2] = fminf(fminf(t.a.z, t.b.z), t.c.z);
  b.hi[0] = fmaxf(fmaxf(t.a.x, t.b.x), t.c.x);
  b.hi[1] = fmaxf(fmaxf(t.a.y, t.b.y), t.c.y);
  b.hi[2] = fmaxf(fmaxf(t.a.z, t.b.z), t.c.z);
  table[k] = b;
}

// Sparse table level j from level j - 1: next[k] is the union of the 2^j
// leaf boxes from k, for every k whose window fits (k + 2^j <= n).
__global__ void __launch_bounds__(kBlock) k_table_level(const Box* prev, Box* next, uint32_t n, uint32_t half) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k + 2 * half > n) return;
  next[k] = unite(prev[k], prev[k + half]);
}

// Every binary node's box: a leaf's from level 0, an internal node's over its
// range [f, l] as the union of the two windows of 2^j, j = floor(log2(l-f+1)),
// that start at f and end at l. Min and max are exact, so this is the box a
// bottom-up union of its children would give.
__global__ void __launch_bounds__(kBlock) k_node_boxes(const uint2* range, const Box* table, uint32_t n,
                                                       uint32_t levels, Box* boxes, uint32_t* fault) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n - 1) return;
  if (i >= n - 1) {
    boxes[i] = table[i - (n - 1)];
    return;
  }
  const uint2 r = range[i];
  if (r.x >= r.y || r.y >= n) {  // a broken tree: reported, not followed
    atomicOr(fault, 1u);
    return;
  }
  const uint32_t j = 31u - static_cast<uint32_t>(__clz(r.y - r.x + 1u));
  if (j >= levels) {
    atomicOr(fault, 1u);
    return;
  }
  const Box* level = table + size_t(j) * n;
  boxes[i] = unite(level[r.x], level[r.y + 1u - (1u << j)]);
}

// ---------------------------------------------------------------- collapse --
struct Task {
  uint32_t binary;  // binary node to expand
  uint32_t wide;    // wide node to write
};

__device__ __forceinline__ float half_area(const Box& b) {
  const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
  return x * y + y * z + z * x;
}

// Boxes grow by a relative margin so no rounding of their corners can cull a
// triangle the watertight test accepts.
__device__ __forceinline__ void inflate(float& lo, float& hi) {
  const float m = (fabsf(lo) + fabsf(hi) + (hi - lo)) * 1.0e-6f + 1.0e-30f;
  lo -= m;
  hi += m;
}

// One level of tasks: each expands its binary node into at most four
// children, opening the child of largest surface area first. Inner children
// take the next level's task slots (one reservation per wave), and a task's
// wide node index is the next level's first index plus its slot, so every
// level's nodes follow the previous level's.
__global__ void __launch_bounds__(kBlock) k_collapse(const Task* tasks, const uint32_t* task_count, Task* next,
                                                     uint32_t* next_count, uint32_t next_base, uint32_t* leaf_total,
                                                     const uint32_t* child, const uint2* range, const Box* boxes,
                                                     int n, uint32_t leaf_size, WideNode* nodes, uint32_t* fault) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < *task_count;  // every lane reaches the reservations below
  const Task task = valid ? tasks[t] : Task{0u, 0u};
  uint32_t list[4];
  int count = 0;
  auto size_of = [&](uint32_t b) -> uint32_t {
    if (b >= static_cast<uint32_t>(n - 1)) return 1u;
    const uint2 r = range[b];
    return r.y - r.x + 1u;
  };
  if (!valid) {
  } else if (n == 1) {
    list[count++] = 0;  // the single leaf is binary node 0 (n - 1 + 0)
  } else if (size_of(task.binary) <= leaf_size) {
    list[count++] = task.binary;  // the whole subtree is one leaf child
  } else {
    list[count++] = child[2 * task.binary];
    list[count++] = child[2 * task.binary + 1];
    while (count < 4) {
      int pick = -1;
      float best = -1.0f;
      for (int c = 0; c < count; ++c) {
        const uint32_t b = list[c];
        if (size_of(b) > leaf_size) {
          const float a = half_area(boxes[b]);
          if (a > best) {
            best = a;
            pick = c;
          }
        }
      }
      if (pick < 0) break;
      const uint32_t b = list[pick];
      list[pick] = child[2 * b];
      list[count++] = child[2 * b + 1];
    }
  }
  // 0: empty or broken, 1: leaf, 2: inner
  uint32_t kind[4] = {0u, 0u, 0u, 0u}, inner = 0, leaves = 0;
  for (int c = 0; c < count; ++c) {
    const uint32_t b = list[c];
    const uint32_t first_b = b >= static_cast<uint32_t>(n - 1) ? b - static_cast<uint32_t>(n - 1) : range[b].x;
    if (b >= static_cast<uint32_t>(2 * n - 1) || first_b + size_of(b) > static_cast<uint32_t>(n)) {
      atomicOr(fault, 2u);  // a broken tree: the child stays empty
      continue;
    }
    kind[c] = size_of(b) <= leaf_size ? 1u : 2u;
    inner += kind[c] == 2u;
    leaves += kind[c] == 1u;
  }
  uint32_t slot = wave_reserve(next_count, inner);
  (void)wave_reserve(leaf_total, leaves);
  if (!valid) return;
  WideNode out;
  for (int c = 0; c < 4; ++c) {
    if (kind[c] == 0u) {
      out.lox[c] = out.loy[c] = out.loz[c] = INFINITY;
      out.hix[c] = out.hiy[c] = out.hiz[c] = -INFINITY;
      out.child[c] = kEmptyChild;
      continue;
    }
    const uint32_t b = list[c];
    Box box = boxes[b];
#pragma unroll
    for (int a = 0; a < 3; ++a) inflate(box.lo[a], box.hi[a]);
    out.lox[c] = box.lo[0];
    out.loy[c] = box.lo[1];
    out.loz[c] = box.lo[2];
    out.hix[c] = box.hi[0];
    out.hiy[c] = box.hi[1];
    out.hiz[c] = box.hi[2];
    if (kind[c] == 1u) {
      const uint32_t first_b = b >= static_cast<uint32_t>(n - 1) ? b - static_cast<uint32_t>(n - 1) : range[b].x;
      out.child[c] = leaf_word(first_b, size_of(b));
    } else {
      const uint32_t w = next_base + slot;
      out.child[c] = w;
      next[slot++] = Task{b, w};
    }
  }
  out.pad[0] = out.pad[1] = out.pad[2] = out.pad[3] = 0;
  nodes[task.wide] = out;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ refit --
// Triangles re-placed from new positions or transforms, each to its slot in
// leaf order.
__global__ void __launch_bounds__(kBlock) k_reflatten(BuildInput in, const uint32_t* tri_of_global, Tri* tris) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= in.tri_count) return;
  tris[tri_of_global[g]] = world_triangle(in, g);
}

// The covered bits of the triangles after the instances' materials changed.
__global__ void __launch_bounds__(kBlock) k_mark_covered(Tri* tris, uint32_t n, const uint32_t* instance_covered) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inst = __float_as_uint(tris[i].b.w);
  const uint32_t prim = __float_as_uint(tris[i].c.w) & ~kCoveredBit;
  tris[i].c.w = __uint_as_float(prim | (instance_covered && instance_covered[inst] ? kCoveredBit : 0u));
}

// One level's nodes [first, first + count): each child's box from its
// triangles (a leaf) or from the exact box its node wrote one launch earlier,
// with the build's margin; the node's exact box is their union.
__global__ void __launch_bounds__(kBlock) k_refit(WideNode* nodes, float4* exact, const Tri* tris, uint32_t first,
                                                  uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t w = first + i;
  WideNode node = nodes[w];
  Box all;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    all.lo[a] = INFINITY;
    all.hi[a] = -INFINITY;
  }
  for (int c = 0; c < 4; ++c) {
    const uint32_t word = node.child[c];
    if (word == kEmptyChild) continue;
    Box b;
    if (word & kLeafBit) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        b.lo[a] = INFINITY;
        b.hi[a] = -INFINITY;
      }
      const uint32_t f = leaf_first(word), n = leaf_count(word);
      for (uint32_t k = 0; k < n; ++k) {
        const Tri t = tris[f + k];
        b.lo[0] = fminf(b.lo[0], fminf(fminf(t.a.x, t.b.x), t.c.x));
        b.lo[1] = fminf(b.lo[1], fminf(fminf(t.a.y, t.b.y), t.c.y));
        b.lo[2] = fminf(b.lo[2], fminf(fminf(t.a.z, t.b.z), t.c.z));
        b.hi[0] = fmaxf(b.hi[0], fmaxf(fmaxf(t.a.x, t.b.x), t.c.x));
        b.hi[1] = fmaxf(b.hi[1], fmaxf(fmaxf(t.a.y, t.b.y), t.c.y));
        b.hi[2] = fmaxf(b.hi[2], fmaxf(fmaxf(t.a.z, t.b.z), t.c.z));
      }
    } else {
      const float4 lo = exact[2ull * word], hi = exact[2ull * word + 1];
      b.lo[0] = lo.x;
      b.lo[1] = lo.y;
      b.lo[2] = lo.z;
      b.hi[0] = hi.x;
      b.hi[1] = hi.y;
      b.hi[2] = hi.z;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      all.lo[a] = fminf(all.lo[a], b.lo[a]);
      all.hi[a] = fmaxf(all.hi[a], b.hi[a]);
      inflate(b.lo[a], b.hi[a]);
    }
    node.lox[c] = b.lo[0];
    node.loy[c] = b.lo[1];
    node.loz[c] = b.lo[2];
    node.hix[c] = b.hi[0];
    node.hiy[c] = b.hi[1];
    node.hiz[c] = b.hi[2];
  }
  nodes[w] = node;
  exact[2ull * w] = make_float4(all.lo[0], all.lo[1], all.lo[2], 0.0f);
  exact[2ull * w + 1] = make_float4(all.hi[0], all.hi[1], all.hi[2], 0.0f);
}

__global__ void k_empty_root(WideNode* nodes) {
  WideNode out;
  for (int c = 0; c < 4; ++c) {
    out.lox[c] = out.loy[c] = out.loz[c] = INFINITY;
    out.hix[c] = out.hiy[c] = out.hiz[c] = -INFINITY;
    out.child[c] = kEmptyChild;
    out.pad[c] = 0;
  }
  nodes[0] = out;
}

template <typename T>
bool alloc(T*& p, size_t count, std::string& error) {
  p = nullptr;
  if (count == 0) count = 1;
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
  if (e != hipSuccess) {
    error = std::string("hipMalloc: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

}  // namespace

bool radix_sort_pairs(uint32_t* keys, uint32_t* values, uint32_t n, uint32_t key_bits,
                      hipStream_t stream, std::string& error) {
  if (n <= 1) return true;
  const uint32_t tiles = blocks_for(n, kSortTile);
  uint32_t *k2 = nullptr, *v2 = nullptr, *counts = nullptr;
  if (!alloc(k2, n, error) || !alloc(v2, n, error) || !alloc(counts, 256ull * tiles + 256, error)) {
    (void)hipFree(k2);
    (void)hipFree(v2);
    return false;
  }
  uint32_t *ka = keys, *va = values, *kb = k2, *vb = v2;
  const int passes = static_cast<int>((key_bits + 7) / 8);
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * 8;
    k_sort_histogram<<<tiles, kBlock, 0, stream>>>(ka, n, shift, tiles, counts);
    k_scan_digit<<<256, 1024, 0, stream>>>(counts, tiles, counts + 256ull * tiles);
    k_scan_add<<<256, 256, 0, stream>>>(counts, tiles, counts + 256ull * tiles);
    k_sort_scatter<<<tiles, kBlock, 0, stream>>>(ka, va, kb, vb, n, shift, tiles, counts);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && ka != keys) {
    e = hipMemcpyAsync(keys, ka, n * 4ull, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(values, va, n * 4ull, hipMemcpyDeviceToDevice, stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(k2);
  (void)hipFree(v2);
  (void)hipFree(counts);
  if (e != hipSuccess) {
    error = std::string("radix sort: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

bool build_bvh(const BuildInput& in, SceneBuffers& out, hipStream_t stream, double& gpu_ms,
               std::string& error) {
  const uint32_t n = in.tri_count;
  EventPair events;
  LOTUS_CHECK(events.create());
  const hipEvent_t start = events.start, stop = events.stop;
  LOTUS_CHECK(hipEventRecord(start, stream));
  if (!alloc(out.tris, n, error) || !alloc(out.tri_of_global, n, error)) return false;
  out.bytes += uint64_t(n) * (sizeof(Tri) + 4);
  if (n == 0) {
    if (!alloc(out.nodes, 1, error)) return false;
    k_empty_root<<<1, 1, 0, stream>>>(out.nodes);
    out.node_count = 1;
    out.depth = 1;
    out.level_sizes.assign(1, 1u);
    out.leaf_count = 0;
    out.bytes += sizeof(WideNode);
    LOTUS_CHECK(hipStreamSynchronize(stream));
    gpu_ms = 0.0;
    return true;
  }
  Tri* world = nullptr;
  float4* centroid = nullptr;
  uint32_t *bounds = nullptr, *keys = nullptr, *vals = nullptr;
  uint32_t* child = nullptr;
  uint2* range = nullptr;
  Box *boxes = nullptr, *table = nullptr;
  Task *tasks_a = nullptr, *tasks_b = nullptr;
  uint32_t* counters = nullptr;  // 0: task count a, 1: task count b, 2: unused, 3: leaves, 4: fault
  const uint32_t internal = n > 1 ? n - 1 : 1;
  // sparse table levels 0 .. floor(log2 n): windows of 1, 2, 4, ... leaves
  const uint32_t levels = 32u - static_cast<uint32_t>(__builtin_clz(n));
  bool ok = alloc(world, n, error) && alloc(centroid, n, error) && alloc(bounds, 6, error) &&
            alloc(keys, n, error) && alloc(vals, n, error) && alloc(child, 2ull * internal, error) &&
            alloc(range, internal, error) && alloc(boxes, 2ull * n, error) &&
            alloc(table, uint64_t(n) * levels, error) && alloc(tasks_a, n, error) && alloc(tasks_b, n, error) &&
            alloc(counters, 5, error);
  auto release = [&]() {
    (void)hipFree(world);
    (void)hipFree(centroid);
    (void)hipFree(bounds);
    (void)hipFree(keys);
    (void)hipFree(vals);
    (void)hipFree(child);
    (void)hipFree(table);
    (void)hipFree(range);
    (void)hipFree(boxes);
    (void)hipFree(tasks_a);
    (void)hipFree(tasks_b);
    (void)hipFree(counters);
  };
  if (!ok) {
    release();
    return false;
  }
  // upper bound of wide nodes: every internal binary node
  const uint32_t max_nodes = internal + 1;
  if (!alloc(out.nodes, max_nodes, error)) {
    release();
    return false;
  }
  const uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  hipError_t e = hipMemcpyAsync(bounds, init, sizeof init, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) {
    k_flatten<<<blocks_for(n, kBlock), kBlock, 0, stream>>>(in, world, centroid, bounds);
    if (in.builder != LOTUS_HIP_BVH_SAH) k_morton<<<blocks_for(n, kBlock), kBlock, 0, stream>>>(centroid, bounds, n, keys, vals);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    error = std::string("flatten: ") + hipGetErrorString(e);
    release();
    return false;
  }
  const bool sah = in.builder == LOTUS_HIP_BVH_SAH && n > 1;
  if (sah) {
    // the topology on the host from the world triangles: leaf order, children, ranges
    std::vector<Tri> host(n);
    std::vector<uint32_t> order(n), kids(2ull * (n - 1)), ranges(2ull * (n - 1));
    e = hipMemcpyAsync(host.data(), world, n * sizeof(Tri), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && !sah_topology(host.data(), n, order.data(), kids.data(), ranges.data(), error)) {
      release();
      return false;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(vals, order.data(), 4ull * n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(child, kids.data(), 4ull * kids.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(range, ranges.data(), 4ull * ranges.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);  // the host vectors go out of scope
    if (e != hipSuccess) {
      error = std::string("SAH topology upload: ") + hipGetErrorString(e);
      release();
      return false;
    }
  } else if (!radix_sort_pairs(keys, vals, n, 30, stream, error)) {
    release();
    return false;
  }
  // the collapse's counters and the fault word start before the box kernels
  const Task root{0u, 0u};  // the root task expands binary node 0
  const uint32_t init_counts[5] = {1u, 0u, 1u, 0u, 0u};
  e = hipMemcpyAsync(tasks_a, &root, sizeof root, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(counters, init_counts, sizeof init_counts, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) {
    if (n > 1 && !sah) k_karras<<<blocks_for(n - 1, kBlock), kBlock, 0, stream>>>(keys, static_cast<int>(n), child, range);
    k_leaf_boxes<<<blocks_for(n, kBlock), kBlock, 0, stream>>>(world, vals, n, table, out.tris, out.tri_of_global);
    for (uint32_t j = 1; j < levels; ++j) {
      const uint32_t half = 1u << (j - 1);
      k_table_level<<<blocks_for(n - 2 * half + 1, kBlock), kBlock, 0, stream>>>(
          table + uint64_t(j - 1) * n, table + uint64_t(j) * n, n, half);
    }
    k_node_boxes<<<blocks_for(2ull * n - 1, kBlock), kBlock, 0, stream>>>(range, table, n, levels, boxes,
                                                                          counters + 4);
    e = hipGetLastError();
  }
  // collapse, level by level
  uint32_t pending = 1, depth = 0, allocated = 1;  // the root is node 0
  Task *cur = tasks_a, *nxt = tasks_b;
  int cur_slot = 0;
  out.level_sizes.clear();
  while (e == hipSuccess && pending > 0 && depth < 256) {
    ++depth;
    out.level_sizes.push_back(pending);
    const int nxt_slot = 1 - cur_slot;
    k_fill_u32<<<1, 1, 0, stream>>>(counters + nxt_slot, 0u, 1u);
    k_collapse<<<blocks_for(pending, kBlock), kBlock, 0, stream>>>(
        cur, counters + cur_slot, nxt, counters + nxt_slot, allocated, counters + 3, child, range, boxes,
        static_cast<int>(n), in.leaf_size, out.nodes, counters + 4);
    e = hipMemcpyAsync(&pending, counters + nxt_slot, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    allocated += pending;
    if (allocated > max_nodes) break;  // reported below
    std::swap(cur, nxt);
    cur_slot = nxt_slot;
  }
  uint32_t final_counts[5] = {0, 0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(final_counts, counters, sizeof final_counts, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipEventRecord(stop, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    error = std::string("BVH build: ") + hipGetErrorString(e);
    release();
    return false;
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, start, stop);
  gpu_ms = ms;
  if (final_counts[4] != 0 || pending > 0 || allocated > max_nodes) {
    error = "BVH build produced an invalid tree (code " + std::to_string(final_counts[4]) + ")";
    release();
    return false;
  }
  out.node_count = allocated;
  out.leaf_count = final_counts[3];
  out.depth = depth;
  out.bytes += uint64_t(max_nodes) * sizeof(WideNode);
  release();
  return true;
}

hipError_t mark_covered(SceneBuffers& s, uint32_t tri_count, const uint32_t* instance_covered, hipStream_t stream) {
  if (tri_count == 0) return hipSuccess;
  k_mark_covered<<<blocks_for(tri_count, kBlock), kBlock, 0, stream>>>(s.tris, tri_count, instance_covered);
  return hipGetLastError();
}

bool refit_bvh(const BuildInput& in, SceneBuffers& s, hipStream_t stream, double& gpu_ms, std::string& error) {
  gpu_ms = 0.0;
  const uint32_t n = in.tri_count;
  if (n == 0) return true;  // the empty root has no box to move
  uint64_t listed = 0;
  for (uint32_t size : s.level_sizes) listed += size;
  if (listed != s.node_count) {
    error = "BVH refit: the level sizes do not cover the nodes";
    return false;
  }
  if (!s.exact) {
    if (!alloc(s.exact, 2ull * s.node_count, error)) return false;
    s.bytes += 2ull * s.node_count * sizeof(float4);
  }
  EventPair events;
  LOTUS_CHECK(events.create());
  const hipEvent_t start = events.start, stop = events.stop;
  hipError_t e = hipEventRecord(start, stream);
  if (e == hipSuccess) {
    k_reflatten<<<blocks_for(n, kBlock), kBlock, 0, stream>>>(in, s.tri_of_global, s.tris);
    // Deepest level first: a node's inner children sit one level below it.
    uint32_t first = s.node_count;
    for (size_t d = s.level_sizes.size(); d-- > 0;) {
      const uint32_t count = s.level_sizes[d];
      first -= count;
      k_refit<<<blocks_for(count, kBlock), kBlock, 0, stream>>>(s.nodes, s.exact, s.tris, first, count);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(stop, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  float ms = 0.0f;
  if (e == hipSuccess) (void)hipEventElapsedTime(&ms, start, stop);
  if (e != hipSuccess) {
    error = std::string("BVH refit: ") + hipGetErrorString(e);
    return false;
  }
  gpu_ms = ms;
  return true;
}

}  // namespace lotus
