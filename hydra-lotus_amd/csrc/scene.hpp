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
//              primitive in the w lanes.
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

struct HitRecord {
  uint32_t instance, primitive, tri;  // tri: index in BVH order
  float u, v, t;
};

// ----------------------------------------------- watertight intersection --
// Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1),
// 2013, with its double-precision fallback for zero edge functions.
struct RayPre {
  int kx, ky, kz;
  float sx, sy, sz;
};
__device__ __forceinline__ RayPre precompute(const Ray& r) {
  RayPre p;
  const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
  p.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  p.kx = p.kz + 1 == 3 ? 0 : p.kz + 1;
  p.ky = p.kx + 1 == 3 ? 0 : p.kx + 1;
  const float dz = comp(r.d, p.kz);
  if (dz < 0.0f) {
    const int t = p.kx;
    p.kx = p.ky;
    p.ky = t;
  }
  p.sx = comp(r.d, p.kx) / dz;
  p.sy = comp(r.d, p.ky) / dz;
  p.sz = 1.0f / dz;
  return p;
}

__device__ __forceinline__ bool intersect_triangle(const Ray& r, const RayPre& p, V3 a0, V3 a1, V3 a2,
                                                   float& t, float& bu, float& bv) {
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

// Register stack of N (node, distance) entries with spill to a lane-
// interleaved global area when full (entry k of lane l at spill[k * lanes + l]).
template <int N>
struct TraversalStack {
  uint32_t node[N];
  float dist[N];
  int count;
  int spilled;
  uint2* spill;
  uint32_t lanes;
  uint32_t lane;
  uint32_t visits;  // wide nodes tested
  uint32_t tests;   // triangles tested
  int capacity;     // spill entries per lane

  __device__ __forceinline__ void init(uint2* spill_area, uint32_t lane_count, uint32_t lane_id, int cap) {
    count = 0;
    spilled = 0;
    capacity = cap;
    visits = 0;
    tests = 0;
    spill = spill_area;
    lanes = lane_count;
    lane = lane_id;
  }
  __device__ __forceinline__ void push(uint32_t n, float d) {
    if (count == N) {
      // The area holds 3 x (tree depth) entries, the most a nearest-first
      // walk of a 4-wide tree keeps; the bound keeps every store in it.
      if (spilled < capacity) {
        spill[static_cast<size_t>(spilled) * lanes + lane] = make_uint2(node[N - 1], __float_as_uint(dist[N - 1]));
        ++spilled;
      }
    } else {
      ++count;
    }
#pragma unroll
    for (int i = N - 1; i > 0; --i) {
      node[i] = node[i - 1];
      dist[i] = dist[i - 1];
    }
    node[0] = n;
    dist[0] = d;
  }
  __device__ __forceinline__ bool pop(uint32_t& n, float& d) {
    if (count > 0) {
      n = node[0];
      d = dist[0];
#pragma unroll
      for (int i = 0; i < N - 1; ++i) {
        node[i] = node[i + 1];
        dist[i] = dist[i + 1];
      }
      --count;
      return true;
    }
    if (spilled > 0) {
      --spilled;
      const uint2 e = spill[static_cast<size_t>(spilled) * lanes + lane];
      n = e.x;
      d = __uint_as_float(e.y);
      return true;
    }
    return false;
  }
};

// The slab test of one node's four children against [tmin, tmax].
__device__ __forceinline__ void test_children(const WideNode* __restrict__ nodes, uint32_t index,
                                              V3 o, V3 inv, float tmin, float tmax, float tn[4],
                                              uint32_t child[4]) {
  const float4* p = reinterpret_cast<const float4*>(nodes + index);
  const float4 lox = p[0], hix = p[1], loy = p[2], hiy = p[3], loz = p[4], hiz = p[5];
  const uint4 ch = reinterpret_cast<const uint4*>(p)[6];
  const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
  const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
  const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
  child[0] = ch.x;
  child[1] = ch.y;
  child[2] = ch.z;
  child[3] = ch.w;
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

template <int N>
struct Traverser {
  const WideNode* __restrict__ nodes;
  const Tri* __restrict__ tris;

  // Closest hit; ties in t go to the lowest flattened triangle id, so the
  // result does not depend on the order in which nodes are visited.
  __device__ __forceinline__ bool closest(const Ray& r, TraversalStack<N>& st, HitRecord& hit) const {
    const RayPre pre = precompute(r);
    const V3 inv = V3{safe_inverse(r.d.x), safe_inverse(r.d.y), safe_inverse(r.d.z)};
    float best_t = r.tmax;
    uint32_t best_g = kMiss, best_tri = kMiss;
    float best_u = 0.0f, best_v = 0.0f;
    uint32_t cur = 0;
    float cur_t = r.tmin;
    bool alive = true;
    uint32_t guard = 0;
    st.count = 0;
    st.spilled = 0;
    while (alive && ++guard < kMaxSteps) {
      if ((cur & kLeafBit) == 0) {
        float tn[4];
        uint32_t ch[4];
        test_children(nodes, cur, r.o, inv, r.tmin, best_t, tn, ch);
        ++st.visits;
        sort4(tn, ch);
        if (tn[0] == INFINITY) {
          alive = false;
        } else {
          if (tn[3] != INFINITY) st.push(ch[3], tn[3]);
          if (tn[2] != INFINITY) st.push(ch[2], tn[2]);
          if (tn[1] != INFINITY) st.push(ch[1], tn[1]);
          cur = ch[0];
          cur_t = tn[0];
          continue;
        }
      } else {
        const uint32_t first = leaf_first(cur), n = leaf_count(cur);
        st.tests += n;
        for (uint32_t k = 0; k < n; ++k) {
          const Tri tr = tris[first + k];
          float t, u, v;
          if (intersect_triangle(r, pre, xyz(tr.a), xyz(tr.b), xyz(tr.c), t, u, v)) {
            const uint32_t g = __float_as_uint(tr.a.w);
            if (t < best_t || (t == best_t && g < best_g)) {
              best_t = t;
              best_g = g;
              best_tri = first + k;
              best_u = u;
              best_v = v;
            }
          }
        }
        alive = false;
      }
      // next entry no farther than the best hit
      uint32_t n;
      float d;
      while (st.pop(n, d)) {
        if (d <= best_t) {
          cur = n;
          cur_t = d;
          alive = true;
          break;
        }
      }
    }
    (void)cur_t;
    if (best_g == kMiss) return false;
    const Tri tr = tris[best_tri];
    hit.instance = __float_as_uint(tr.b.w);
    hit.primitive = __float_as_uint(tr.c.w);
    hit.tri = best_tri;
    hit.u = best_u;
    hit.v = best_v;
    hit.t = best_t;
    return true;
  }

  // Any hit in [tmin, tmax] (ShadowTransmittance for opaque surfaces).
  __device__ __forceinline__ bool occluded(const Ray& r, TraversalStack<N>& st) const {
    const RayPre pre = precompute(r);
    const V3 inv = V3{safe_inverse(r.d.x), safe_inverse(r.d.y), safe_inverse(r.d.z)};
    uint32_t cur = 0;
    bool alive = true;
    uint32_t guard = 0;
    st.count = 0;
    st.spilled = 0;
    while (alive && ++guard < kMaxSteps) {
      if ((cur & kLeafBit) == 0) {
        float tn[4];
        uint32_t ch[4];
        test_children(nodes, cur, r.o, inv, r.tmin, r.tmax, tn, ch);
        ++st.visits;
        sort4(tn, ch);
        if (tn[0] != INFINITY) {
          if (tn[3] != INFINITY) st.push(ch[3], tn[3]);
          if (tn[2] != INFINITY) st.push(ch[2], tn[2]);
          if (tn[1] != INFINITY) st.push(ch[1], tn[1]);
          cur = ch[0];
          continue;
        }
      } else {
        const uint32_t first = leaf_first(cur), n = leaf_count(cur);
        for (uint32_t k = 0; k < n; ++k) {
          const Tri tr = tris[first + k];
          float t, u, v;
          ++st.tests;
          if (intersect_triangle(r, pre, xyz(tr.a), xyz(tr.b), xyz(tr.c), t, u, v)) {
            // drain the stack so the spill bookkeeping ends clean
            uint32_t dn;
            float dd;
            while (st.pop(dn, dd)) {
            }
            return true;
          }
        }
      }
      float d;
      alive = st.pop(cur, d);
    }
    return false;
  }
};

}  // namespace lotus
