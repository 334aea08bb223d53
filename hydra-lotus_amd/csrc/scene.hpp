// The GPU scene as the kernels read it, and software ray traversal of its wide
// BVH. Replaces the hardware ray queries of backend/vulkan/shaders/common/
// scene.slang:424-460 (Trace, ShadowTransmittance) over the driver-built
// BLAS/TLAS of backend/vulkan/gpu_scene.cpp:1899-1950, 2250-2323.
//
// Layout in HBM (all instances flattened into one world-space level):
//   nodes    : WideNode[node_count], 128 B, the root at 0. Child boxes are
//              stored axis-major (lox[4], hix[4], ...) so one node is eight
//              16-byte loads and four slab tests share every load.
//   tris     : Tri[tri_count], 48 B, in BVH leaf order: the three world-space
//              corners, with the flattened triangle id, instance and
//              primitive in the w lanes; the primitive's bit 31 marks a
//              triangle whose material's coverage is tested.
// A child word is a node index, or (bit 31) a leaf: first triangle in bits
// 0-26, count - 1 in bits 27-30. An empty child has an inverted box.
#pragma once

#include "device_common.hpp"

namespace lotus {

constexpr int kWidth = 4;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
// Traversal steps after which a ray gives up: far beyond any valid tree's
// need (each node is visited at most once), it bounds a broken one.
constexpr uint32_t kMaxSteps = 1u << 22;

struct alignas(16) WideNode {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(WideNode) == 128, "wide node is 128 bytes");

struct alignas(16) Tri {
  float4 a;  // v0, w = flattened id (bits)
  float4 b;  // v1, w = instance (bits)
  float4 c;  // v2, w = primitive (bits)
};
static_assert(sizeof(Tri) == 48, "triangle record is 48 bytes");

__host__ __device__ inline uint32_t leaf_word(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1u) << 27) | first;
}
__device__ __forceinline__ uint32_t leaf_first(uint32_t w) { return w & 0x07FFFFFFu; }
__device__ __forceinline__ uint32_t leaf_count(uint32_t w) { return ((w >> 27) & 15u) + 1u; }

struct Ray {
  V3 o;
  float tmin;
  V3 d;
  float tmax;
};

// Bit 31 of a Tri's primitive word: the triangle's instance has a material
// whose coverage (scene.slang:388-406) may be below one, so a ray tests it.
constexpr uint32_t kCoveredBit = 0x80000000u;

// Coverage policies of a walk. NoCoverage: every triangle is opaque (no
// code for the test is compiled). A policy with kEnabled answers the
// coverage of candidate (instance, primitive, u, v) of a covered triangle.
struct NoCoverage {
  static constexpr bool kEnabled = false;
  __device__ __forceinline__ float operator()(uint32_t, uint32_t, float, float) const { return 1.0f; }
};

// AcceptSurface (scene.slang:408-413) draws a candidate's uniform number
// from the path's stream in the order the hardware reports candidates, an
// order no other traversal reproduces. Here the number is a hash of the
// ray (its origin and direction bits) and the candidate's flattened id: the
// same law (each candidate accepted with probability equal to its coverage,
// independently), decided the same way whatever order candidates are met
// in, and by the oracle's brute-force trace alike. The path's stream is not
// advanced.
__device__ __forceinline__ uint32_t coverage_ray_key(V3 o, V3 d) {
  uint32_t h = pcg_hash(__float_as_uint(d.z));
  h = pcg_hash(__float_as_uint(d.y) ^ h);
  h = pcg_hash(__float_as_uint(d.x) ^ h);
  h = pcg_hash(__float_as_uint(o.z) ^ h);
  h = pcg_hash(__float_as_uint(o.y) ^ h);
  return pcg_hash(__float_as_uint(o.x) ^ h);
}
__device__ __forceinline__ float coverage_uniform(uint32_t key, uint32_t g) {
  return static_cast<float>(pcg_hash(key ^ pcg_hash(g)) >> 8u) * (1.0f / 16777216.0f);
}

struct HitRecord {
  uint32_t instance, primitive, tri;  // tri: index in BVH order
  float u, v, t;
};

// ----------------------------------------------- watertight intersection --
// Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1),
// 2013, with its double-precision fallback for zero edge functions.
// The axis permutation is packed in one word (kx | ky << 2 | kz << 4) to
// keep a walk's registers few.
struct RayPre {
  uint32_t k;
  float sx, sy, sz;
  __device__ __forceinline__ int kx() const { return static_cast<int>(k & 3u); }
  __device__ __forceinline__ int ky() const { return static_cast<int>((k >> 2) & 3u); }
  __device__ __forceinline__ int kz() const { return static_cast<int>(k >> 4); }
};
__device__ __forceinline__ RayPre precompute(const Ray& r) {
  RayPre p;
  const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
  const int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  const float dz = comp(r.d, kz);
  if (dz < 0.0f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  p.k = static_cast<uint32_t>(kx) | (static_cast<uint32_t>(ky) << 2) | (static_cast<uint32_t>(kz) << 4);
  p.sx = comp(r.d, kx) / dz;
  p.sy = comp(r.d, ky) / dz;
  p.sz = 1.0f / dz;
  return p;
}

__device__ __forceinline__ bool intersect_triangle(const Ray& r, const RayPre& pk, V3 a0, V3 a1, V3 a2,
                                                   float& t, float& bu, float& bv) {
  const struct {
    int kx, ky, kz;
    float sx, sy, sz;
  } p = {pk.kx(), pk.ky(), pk.kz(), pk.sx, pk.sy, pk.sz};
  const V3 A = a0 - r.o, B = a1 - r.o, C = a2 - r.o;
  const float Az0 = comp(A, p.kz), Bz0 = comp(B, p.kz), Cz0 = comp(C, p.kz);
  const float Ax = comp(A, p.kx) - p.sx * Az0;
  const float Ay = comp(A, p.ky) - p.sy * Az0;
  const float Bx = comp(B, p.kx) - p.sx * Bz0;
  const float By = comp(B, p.ky) - p.sy * Bz0;
  const float Cx = comp(C, p.kx) - p.sx * Cz0;
  const float Cy = comp(C, p.ky) - p.sy * Cz0;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = static_cast<float>(static_cast<double>(Cx) * By - static_cast<double>(Cy) * Bx);
    V = static_cast<float>(static_cast<double>(Ax) * Cy - static_cast<double>(Ay) * Cx);
    W = static_cast<float>(static_cast<double>(Bx) * Ay - static_cast<double>(By) * Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = p.sz * Az0, Bz = p.sz * Bz0, Cz = p.sz * Cz0;
  const float T = (U * Az + V * Bz) + W * Cz;
  const float inv = 1.0f / det;
  t = T * inv;
  if (!(t >= r.tmin) || !(t <= r.tmax)) return false;
  bu = V * inv;
  bv = W * inv;
  return true;
}

// Conservative slab distances (Ize, "Robust BVH Ray Traversal", JCGT 2013):
// the far distance grows by 1 + 2 gamma(3); inverse directions never hold
// infinities, so no 0 * inf arises.
__device__ __forceinline__ float safe_inverse(float d) {
  return 1.0f / (fabsf(d) > 1.0e-30f ? d : copysignf(1.0e-30f, d));
}

// Traversal stack of one lane: a ring of S (node, distance) entries in LDS,
// laid out entry-major ([S][W lanes of the workgroup]) so a wave's pushes and
// pops touch consecutive words, spilling its oldest entries to a
// lane-interleaved global area (entry k of lane l at spill[k * lanes + l])
// when the ring is full. W is the workgroup size, a constant, so an entry's
// address is a shift and an add.
template <int S, int W>
struct LdsStack {
  // An entry is one 64-bit word, node in the low half, distance bits in the
  // high half. Pops read the ring top with a volatile LDS load before they
  // choose between it and the spill area: a load merged from an LDS and a
  // global address would be a flat access that also waits for the lane's
  // outstanding global loads.
  typedef __attribute__((address_space(3))) unsigned long long LdsEntry;
  LdsEntry* mine;   // this lane's column of the workgroup's ring
  unsigned long long* spill;
  uint32_t lanes, lane;
  int capacity;     // spill entries per lane
  int base, count, spilled;

  __device__ __forceinline__ void init(uint2* lds, uint2* spill_area, int cap) {
    mine = (LdsEntry*)(lds) + threadIdx.x;  // an address space cast: lds is the workgroup's ring
    spill = reinterpret_cast<unsigned long long*>(spill_area);
    lanes = gridDim.x * W;
    lane = blockIdx.x * W + threadIdx.x;
    capacity = cap;
    clear();
  }
  __device__ __forceinline__ void clear() {
    base = 0;
    count = 0;
    spilled = 0;
  }
  static __device__ __forceinline__ unsigned long long entry(uint32_t n, float d) {
    return static_cast<unsigned long long>(n) | (static_cast<unsigned long long>(__float_as_uint(d)) << 32);
  }
  // Entry i of the ring (i < 2S).
  __device__ __forceinline__ LdsEntry& at(int i) { return mine[(i >= S ? i - S : i) * W]; }
  // Frees room for k more entries by spilling the oldest. The area holds
  // 3 x (tree depth) entries, the most a nearest-first walk of a 4-wide tree
  // keeps; the bound keeps every store in it.
  __device__ __forceinline__ void reserve(int k) {
    while (count + k > S) {
      if (spilled < capacity) spill[static_cast<size_t>(spilled++) * lanes + lane] = at(base);
      base = base + 1 == S ? 0 : base + 1;
      --count;
    }
  }
  __device__ __forceinline__ void push(uint32_t n, float d) {
    reserve(1);
    at(base + count) = entry(n, d);
    ++count;
  }
  // Pushes children 1 .. k of a node's sorted children (k <= 3), farthest
  // first, so child 1 is popped next.
  __device__ __forceinline__ void push_sorted(int k, const uint32_t ch[4], const float tn[4]) {
    reserve(k);
    const int top = base + count + k;
#pragma unroll
    for (int m = 1; m <= 3; ++m)
      if (m <= k) at(top - m) = entry(ch[m], tn[m]);
    count += k;
  }
  __device__ __forceinline__ bool pop(uint32_t& n, float& d) {
    unsigned long long e = *static_cast<volatile LdsEntry*>(&at(base + count + (count > 0 ? -1 : S - 1)));
    if (count > 0) {
      --count;
    } else if (spilled > 0) {
      e = spill[static_cast<size_t>(--spilled) * lanes + lane];
    } else {
      return false;
    }
    n = static_cast<uint32_t>(e);
    d = __uint_as_float(static_cast<uint32_t>(e >> 32));
    return true;
  }
};

// A wide node's words as loaded: the six box planes (axis-major) and the
// four child words.
struct NodeWords {
  float4 lox, hix, loy, hiy, loz, hiz;
  uint4 ch;
};
__device__ __forceinline__ NodeWords load_node(const WideNode* __restrict__ nodes, uint32_t index) {
  const float4* p = reinterpret_cast<const float4*>(nodes + index);
  NodeWords n;
  n.lox = p[0];
  n.hix = p[1];
  n.loy = p[2];
  n.hiy = p[3];
  n.loz = p[4];
  n.hiz = p[5];
  n.ch = reinterpret_cast<const uint4*>(p)[6];
  return n;
}

// The slab test of one node's four children against [tmin, tmax].
__device__ __forceinline__ void test_children(const NodeWords& n, V3 o, V3 inv, float tmin, float tmax,
                                              float tn[4], uint32_t child[4]) {
  const float lx[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w}, hx[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
  const float ly[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w}, hy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w};
  const float lz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w}, hz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
  child[0] = n.ch.x;
  child[1] = n.ch.y;
  child[2] = n.ch.z;
  child[3] = n.ch.w;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float x0 = (lx[c] - o.x) * inv.x, x1 = (hx[c] - o.x) * inv.x;
    float y0 = (ly[c] - o.y) * inv.y, y1 = (hy[c] - o.y) * inv.y;
    float z0 = (lz[c] - o.z) * inv.z, z1 = (hz[c] - o.z) * inv.z;
    const float near = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    const float far = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), kFloatMax)) *
                      1.0000003576278687f;
    tn[c] = child[c] != kEmptyChild && near <= fminf(far, tmax) ? near : INFINITY;
  }
}
__device__ __forceinline__ void test_children(const WideNode* __restrict__ nodes, uint32_t index,
                                              V3 o, V3 inv, float tmin, float tmax, float tn[4],
                                              uint32_t child[4]) {
  test_children(load_node(nodes, index), o, inv, tmin, tmax, tn, child);
}

// Orders up to four (distance, child) pairs ascending; misses are INFINITY.
__device__ __forceinline__ void sort4(float t[4], uint32_t c[4]) {
#define LOTUS_CSWAP(i, j)                          \
  if (t[j] < t[i]) {                               \
    const float tt = t[i];                         \
    t[i] = t[j];                                   \
    t[j] = tt;                                     \
    const uint32_t cc = c[i];                      \
    c[i] = c[j];                                   \
    c[j] = cc;                                     \
  }
  LOTUS_CSWAP(0, 1) LOTUS_CSWAP(2, 3) LOTUS_CSWAP(0, 2) LOTUS_CSWAP(1, 3) LOTUS_CSWAP(1, 2)
#undef LOTUS_CSWAP
}

constexpr uint32_t kWalkDone = 0xFFFFFFFEu;  // no node left (never a node or leaf word)

// One ray's resumable walk of the wide BVH: each step tests one node's four
// children (nearest-first descent) or one leaf's triangles. Closest hit:
// ties in t go to the lowest flattened triangle id, so the result does not
// depend on the order nodes are visited. Any hit (ShadowTransmittance): the
// walk ends at the first opaque triangle in [tmin, tmax].
//
// Under a coverage policy (scene.slang:388-460), a covered triangle's
// candidate counts only if accepted: a closest-hit walk accepts it with
// probability equal to its coverage (decided by coverage_uniform), and an
// any-hit walk ends at it only at coverage one, otherwise multiplying its
// transmittance by 1 - coverage. The product is kept in double, where the
// factors of up to two surfaces multiply exactly and more round far below
// float's precision, so its float value does not depend on the order
// candidates are met in (the reference multiplies in float in the
// hardware's order).
template <class Cov>
struct WalkT {
  Ray r;
  RayPre pre;
  V3 inv;
  uint32_t cur;
  float best_t;
  uint32_t best_g, best_tri;
  bool any;
  uint32_t steps;
  uint32_t visits;  // wide nodes tested
  uint32_t tests;   // triangles tested
  Cov cov;
  uint32_t key;     // the ray's coverage key (closest hit)
  double trans;     // transmittance so far (any hit)

  __device__ __forceinline__ void start(const Ray& ray, bool any_hit) {
    r = ray;
    pre = precompute(ray);
    inv = V3{safe_inverse(ray.d.x), safe_inverse(ray.d.y), safe_inverse(ray.d.z)};
    cur = 0;
    best_t = ray.tmax;
    best_g = kMiss;
    best_tri = kMiss;
    any = any_hit;
    steps = 0;
    visits = 0;
    tests = 0;
    if constexpr (Cov::kEnabled) {
      key = any_hit ? 0u : coverage_ray_key(ray.o, ray.d);
      trans = 1.0;
    }
  }
  __device__ __forceinline__ bool done() const { return cur == kWalkDone; }
  __device__ __forceinline__ bool hit() const { return best_g != kMiss; }
  // An any-hit walk's result: 0 at an opaque hit, else the product of the
  // covered surfaces' transparencies (1 without a coverage policy).
  __device__ __forceinline__ float transmittance() const {
    if (hit()) return 0.0f;
    if constexpr (Cov::kEnabled) return static_cast<float>(trans);
    return 1.0f;
  }

  // The next entry no farther than the best hit (any entry for any hit).
  template <int S, int W>
  __device__ __forceinline__ void next(LdsStack<S, W>& st) {
    uint32_t n;
    float d;
    while (st.pop(n, d)) {
      if (d <= best_t) {
        cur = n;
        return;
      }
    }
    cur = kWalkDone;
  }

  // One inner node: its children nearest first, the nearest taken now and
  // the others pushed; with none hit, the next entry.
  template <int S, int W>
  __device__ __forceinline__ void visit(const WideNode* __restrict__ nodes, LdsStack<S, W>& st) {
    float tn[4];
    uint32_t ch[4];
    test_children(nodes, cur, r.o, inv, r.tmin, best_t, tn, ch);
    ++visits;
    sort4(tn, ch);
    if (tn[0] != INFINITY) {
      st.push_sorted(static_cast<int>(tn[1] != INFINITY) + static_cast<int>(tn[2] != INFINITY) +
                         static_cast<int>(tn[3] != INFINITY),
                     ch, tn);
      cur = ch[0];
      return;
    }
    next(st);
  }

  // The triangles of leaf `word`; true when an any-hit walk has found its hit.
  __device__ __forceinline__ bool test_leaf(const Tri* __restrict__ tris, uint32_t word) {
    const uint32_t first = leaf_first(word), n = leaf_count(word);
    for (uint32_t k = 0; k < n; ++k) {
      const Tri t4 = tris[first + k];
      float t, u, v;
      ++tests;
      if (intersect_triangle(r, pre, xyz(t4.a), xyz(t4.b), xyz(t4.c), t, u, v)) {
        const uint32_t g = __float_as_uint(t4.a.w);
        if constexpr (Cov::kEnabled) {
          const uint32_t pw = __float_as_uint(t4.c.w);
          if (pw & kCoveredBit) {
            // only a candidate that could become the closest hit is tested
            if (!any && !(t < best_t || (t == best_t && g < best_g))) continue;
            const float c = cov(__float_as_uint(t4.b.w), pw & ~kCoveredBit, u, v);
            if (any) {
              if (!(c >= 1.0f)) {
                if (c > 0.0f) trans *= static_cast<double>(1.0f - c);
                continue;
              }
            } else if (!(c > 0.0f) || (c < 1.0f && !(coverage_uniform(key, g) < c))) {
              continue;
            }
          }
        }
        if (any) {
          best_g = g;
          best_tri = first + k;
          best_t = t;
          return true;
        }
        if (t < best_t || (t == best_t && g < best_g)) {
          best_t = t;
          best_g = g;
          best_tri = first + k;
        }
      }
    }
    return false;
  }

  // One node or one leaf (resumable: the drivers that hand lanes new rays
  // step each walk separately).
  template <int S, int W>
  __device__ __forceinline__ void step(const WideNode* __restrict__ nodes, const Tri* __restrict__ tris,
                                       LdsStack<S, W>& st) {
    if (cur == kWalkDone) return;
    if (++steps >= kMaxSteps) {
      st.clear();
      cur = kWalkDone;
      return;
    }
    if ((cur & kLeafBit) == 0) {
      visit(nodes, st);
      return;
    }
    if (test_leaf(tris, cur)) {
      st.clear();
      cur = kWalkDone;
      return;
    }
    next(st);
  }

  // The walk to its end, the wave's lanes together. A step of the wave
  // visits inner nodes or tests leaves, not both: lanes that reach a leaf
  // wait until as many lanes wait at leaves as walk inner nodes, then the
  // waiting lanes test their leaves together, so the leaf code runs for many
  // lanes at once instead of in every step (one of the two paths of a step
  // otherwise runs for a few lanes while the rest of the wave idles). Results
  // do not depend on the order of steps.
  template <int S, int W>
  __device__ __forceinline__ void run(const WideNode* __restrict__ nodes, const Tri* __restrict__ tris,
                                      LdsStack<S, W>& st) {
    while (true) {
      const bool live = cur != kWalkDone;
      const bool at_leaf = live && (cur & kLeafBit) != 0;
      const uint64_t leaves = __ballot(at_leaf), inners = __ballot(live && !at_leaf);
      if ((leaves | inners) == 0) break;
      if (++steps >= kMaxSteps) {
        st.clear();
        cur = kWalkDone;
        continue;
      }
      if (inners != 0 && __popcll(leaves) < __popcll(inners)) {
        if (live && !at_leaf) visit(nodes, st);
      } else if (at_leaf) {
        if (test_leaf(tris, cur)) {
          st.clear();
          cur = kWalkDone;
        } else {
          next(st);
        }
      }
    }
  }

  // The closest hit's record (call when done() and hit()); its barycentrics
  // come from the same test again, which gives the same values.
  __device__ __forceinline__ HitRecord record(const Tri* __restrict__ tris) const {
    const Tri t4 = tris[best_tri];
    HitRecord h;
    h.instance = __float_as_uint(t4.b.w);
    h.primitive = __float_as_uint(t4.c.w) & ~kCoveredBit;
    h.tri = best_tri;
    float t = best_t;
    h.u = h.v = 0.0f;
    (void)intersect_triangle(r, pre, xyz(t4.a), xyz(t4.b), xyz(t4.c), t, h.u, h.v);
    h.t = best_t;
    return h;
  }
};

using Walk = WalkT<NoCoverage>;

// A walk followed to its end by one lane.
template <class Cov = NoCoverage, int S, int W>
__device__ __forceinline__ bool walk_closest(const WideNode* nodes, const Tri* tris, LdsStack<S, W>& st,
                                             const Ray& r, HitRecord& h, uint32_t& visits, uint32_t& tests,
                                             const Cov& cov = Cov{}) {
  WalkT<Cov> w;
  w.cov = cov;
  w.start(r, false);
  st.clear();
  w.run(nodes, tris, st);
  visits += w.visits;
  tests += w.tests;
  if (!w.hit()) return false;
  h = w.record(tris);
  return true;
}
// ShadowTransmittance (scene.slang:445-460) followed to its end by one lane.
template <class Cov = NoCoverage, int S, int W>
__device__ __forceinline__ float walk_transmittance(const WideNode* nodes, const Tri* tris, LdsStack<S, W>& st,
                                                   const Ray& r, uint32_t& visits, uint32_t& tests,
                                                   const Cov& cov = Cov{}) {
  WalkT<Cov> w;
  w.cov = cov;
  w.start(r, true);
  st.clear();
  w.run(nodes, tris, st);
  visits += w.visits;
  tests += w.tests;
  return w.transmittance();
}

}  // namespace lotus
