// Device LBVH build and the executor's triangle queries (first hit, inside, bounded distance).
//
// Build (replaces parry's TriMesh BVH behind RayScene::new / of_mesh / of_subset,
// crates/sherd-core/src/spatial/bvh.rs:43-82):
//   1. one thread per face: the triangle record, its centroid; two-level bound reductions;
//   2. 30-bit Morton codes of the centroids in the centroid box;
//   3. a stable LSD radix sort (sort.hpp), so equal codes keep face order;
//   4. leaves in sorted order, then the hierarchy bottom-up in one pass (agglomerative LBVH: each
//      node joins the neighbour with the longer common key prefix, the second child to arrive
//      unions the boxes and climbs), min/max unions being order-free, so the tree is deterministic;
//   5. escape links: the node after subtree [l, r] in pre-order is the right child of split r.
#include <hip/hip_runtime.h>

#include <vector>

#include "bvh_query.hpp"
#include "kernels.hpp"
#include "sort.hpp"

namespace sherd {
namespace {

constexpr int BT = 256;

__global__ void k_tri_setup(const float *V, const uint32_t *F, int m, Tri *tris, float *cent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  Tri t;
  for (int k = 0; k < 3; ++k) {
    uint32_t vi = F[3 * i + k];
    t.v[3 * k] = V[3 * vi];
    t.v[3 * k + 1] = V[3 * vi + 1];
    t.v[3 * k + 2] = V[3 * vi + 2];
  }
  t.face = (uint32_t)i;
  t.pad[0] = t.pad[1] = 0;
  tris[i] = t;
  for (int a = 0; a < 3; ++a) cent[3 * i + a] = (t.v[a] + t.v[3 + a] + t.v[6 + a]) / 3.0f;
}

// Per-block min/max of the centroids (slots 0-5) and the vertices (6-11), and max |coordinate|.
__global__ void k_bounds_partial(const Tri *tris, const float *cent, int m, float *part) {
  __shared__ float sh[13][BT];
  int t = threadIdx.x;
  float v[13];
  for (int k = 0; k < 3; ++k) {
    v[k] = __builtin_inff();
    v[3 + k] = -__builtin_inff();
    v[6 + k] = __builtin_inff();
    v[9 + k] = -__builtin_inff();
  }
  v[12] = 0.0f;
  for (int i = blockIdx.x * BT + t; i < m; i += gridDim.x * BT) {
    for (int k = 0; k < 3; ++k) {
      float c = cent[3 * i + k];
      v[k] = fminf(v[k], c);
      v[3 + k] = fmaxf(v[3 + k], c);
      for (int q = 0; q < 3; ++q) {
        float x = tris[i].v[3 * q + k];
        v[6 + k] = fminf(v[6 + k], x);
        v[9 + k] = fmaxf(v[9 + k], x);
        v[12] = fmaxf(v[12], fabsf(x));
      }
    }
  }
  for (int k = 0; k < 13; ++k) sh[k][t] = v[k];
  __syncthreads();
  for (int s = BT / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 13; ++k) {
        bool mx = (k >= 3 && k < 6) || (k >= 9);
        sh[k][t] = mx ? fmaxf(sh[k][t], sh[k][t + s]) : fminf(sh[k][t], sh[k][t + s]);
      }
    __syncthreads();
  }
  if (t < 13) part[blockIdx.x * 13 + t] = sh[t][0];
}

__global__ void k_bounds_final(const float *part, int blocks, float *out) {
  int k = threadIdx.x;
  if (k >= 13) return;
  bool mx = (k >= 3 && k < 6) || (k >= 9);
  float v = mx ? -__builtin_inff() : __builtin_inff();
  if (k == 12) v = 0.0f;
  for (int b = 0; b < blocks; ++b) v = mx ? fmaxf(v, part[b * 13 + k]) : fminf(v, part[b * 13 + k]);
  out[k] = v;
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ void k_morton(const float *cent, int m, const float *bounds, uint32_t *code, uint32_t *val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    float lo = bounds[k], hi = bounds[3 + k];
    float ext = hi - lo;
    float u = ext > 0.0f ? (cent[3 * i + k] - lo) / ext : 0.0f;
    float s = u * 1024.0f;
    s = s < 0.0f ? 0.0f : (s > 1023.0f ? 1023.0f : s);
    q[k] = (uint32_t)s;
  }
  code[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  val[i] = (uint32_t)i;
}

__global__ void k_leaves(const Tri *in, const uint32_t *order, int m, Tri *out, Node *nodes, int *other) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  Tri t = in[order[j]];
  out[j] = t;
  Node nd;
  for (int k = 0; k < 3; ++k) {
    float a = t.v[k], b = t.v[3 + k], c = t.v[6 + k];
    nd.lo[k] = fminf(a, fminf(b, c));
    nd.hi[k] = fmaxf(a, fmaxf(b, c));
  }
  nd.a = j;
  nd.esc = -1;
  nodes[m - 1 + j] = nd;
  if (j < m - 1) other[j] = -1;
}

// Common-prefix distance between sorted keys i and i+1, made unique by the position.
__device__ __forceinline__ uint64_t delta(const uint32_t *code, int i) {
  uint64_t a = ((uint64_t)code[i] << 32) | (uint32_t)i;
  uint64_t b = ((uint64_t)code[i + 1] << 32) | (uint32_t)(i + 1);
  return a ^ b;
}

__global__ void k_hierarchy(const uint32_t *code, int m, Node *nodes, int *other, int *childb, int *range_r,
                            int *root) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int cur = m - 1 + j;
  int l = j, r = j;
  while (true) {
    bool left = (l == 0) || (r != m - 1 && delta(code, r) < delta(code, l - 1));
    int parent = left ? r : l - 1;
    if (left) {
      __hip_atomic_store(&nodes[parent].a, cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(&childb[parent], cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    int prev = __hip_atomic_exchange(&other[parent], left ? l : r, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev < 0) return;
    __threadfence();
    if (left) r = prev; else l = prev;
    int ca = __hip_atomic_load(&nodes[parent].a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int cb = __hip_atomic_load(&childb[parent], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      float la = __hip_atomic_load(&nodes[ca].lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      float lb = __hip_atomic_load(&nodes[cb].lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      float ha = __hip_atomic_load(&nodes[ca].hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      float hb = __hip_atomic_load(&nodes[cb].hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      lo[k] = fminf(la, lb);
      hi[k] = fmaxf(ha, hb);
    }
    for (int k = 0; k < 3; ++k) {
      __hip_atomic_store(&nodes[parent].lo[k], lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&nodes[parent].hi[k], hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    range_r[parent] = r;
    cur = parent;
    if (l == 0 && r == m - 1) {
      *root = parent;
      return;
    }
  }
}

__global__ void k_ropes(int m, Node *nodes, const int *childb, const int *range_r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * m - 1) return;
  int r = i < m - 1 ? range_r[i] : i - (m - 1);
  nodes[i].esc = (r >= m - 1) ? -1 : childb[r];
}

#define CK(x)                                 \
  do {                                        \
    hipError_t e_ = (x);                      \
    if (e_ != hipSuccess) return e_;          \
  } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

inline int blocks_for(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

hipError_t scene_build(const float *dV, const uint32_t *dF, int m, DevScene *out, hipStream_t s) {
  out->nodes = nullptr;
  out->tris = nullptr;
  out->n = m;
  if (m <= 0) return hipErrorInvalidValue;
  float *cent = nullptr, *part = nullptr, *bounds = nullptr;
  Tri *raw = nullptr;
  uint32_t *code = nullptr, *val = nullptr, *tk = nullptr, *tv = nullptr, *hist = nullptr;
  int *other = nullptr, *childb = nullptr, *range_r = nullptr, *root = nullptr;
  hipError_t err = hipSuccess;
  int pblocks = blocks_for(m, BT);
  if (pblocks > 1024) pblocks = 1024;
  do {
    if ((err = dalloc(&out->nodes, 2 * (size_t)m - 1)) != hipSuccess) break;
    if ((err = dalloc(&out->tris, m)) != hipSuccess) break;
    if ((err = dalloc(&raw, m)) != hipSuccess) break;
    if ((err = dalloc(&cent, 3 * (size_t)m)) != hipSuccess) break;
    if ((err = dalloc(&part, 13 * (size_t)pblocks)) != hipSuccess) break;
    if ((err = dalloc(&bounds, 13)) != hipSuccess) break;
    if ((err = dalloc(&code, m)) != hipSuccess) break;
    if ((err = dalloc(&val, m)) != hipSuccess) break;
    if ((err = dalloc(&tk, m)) != hipSuccess) break;
    if ((err = dalloc(&tv, m)) != hipSuccess) break;
    if ((err = dalloc(&hist, rsort::hist_words(m))) != hipSuccess) break;
    if ((err = dalloc(&other, m)) != hipSuccess) break;
    if ((err = dalloc(&childb, m)) != hipSuccess) break;
    if ((err = dalloc(&range_r, m)) != hipSuccess) break;
    if ((err = dalloc(&root, 1)) != hipSuccess) break;
    hipLaunchKernelGGL(k_tri_setup, dim3(blocks_for(m, BT)), dim3(BT), 0, s, dV, dF, m, raw, cent);
    hipLaunchKernelGGL(k_bounds_partial, dim3(pblocks), dim3(BT), 0, s, raw, cent, m, part);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, part, pblocks, bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(m, BT)), dim3(BT), 0, s, cent, m, bounds, code, val);
    if ((err = hipGetLastError()) != hipSuccess) break;
    if ((err = rsort::sort_pairs<uint32_t>(code, val, tk, tv, hist, m, 0, 30, s)) != hipSuccess) break;
    hipLaunchKernelGGL(k_leaves, dim3(blocks_for(m, BT)), dim3(BT), 0, s, raw, val, m, out->tris, out->nodes, other);
    int zero = m == 1 ? 0 : -1;
    if ((err = hipMemcpyAsync(root, &zero, sizeof(int), hipMemcpyHostToDevice, s)) != hipSuccess) break;
    if (m > 1) {
      hipLaunchKernelGGL(k_hierarchy, dim3(blocks_for(m, BT)), dim3(BT), 0, s, code, m, out->nodes, other, childb,
                         range_r, root);
    }
    hipLaunchKernelGGL(k_ropes, dim3(blocks_for(2L * m - 1, BT)), dim3(BT), 0, s, m, out->nodes, childb, range_r);
    if ((err = hipGetLastError()) != hipSuccess) break;
    float hb[13];
    if ((err = hipMemcpyAsync(hb, bounds, sizeof hb, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    int r = -1;
    if ((err = hipMemcpyAsync(&r, root, sizeof r, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
    out->root = r;
    for (int k = 0; k < 3; ++k) {
      out->lo[k] = hb[6 + k];
      out->hi[k] = hb[9 + k];
    }
    out->absmax = hb[12];
    if (r < 0) err = hipErrorUnknown;
  } while (0);
  (void)hipFree(raw);
  (void)hipFree(cent);
  (void)hipFree(part);
  (void)hipFree(bounds);
  (void)hipFree(code);
  (void)hipFree(val);
  (void)hipFree(tk);
  (void)hipFree(tv);
  (void)hipFree(hist);
  (void)hipFree(other);
  (void)hipFree(childb);
  (void)hipFree(range_r);
  (void)hipFree(root);
  if (err != hipSuccess) scene_free(out);
  return err;
}

void scene_free(DevScene *s) {
  if (s->nodes) (void)hipFree(s->nodes);
  if (s->tris) (void)hipFree(s->tris);
  s->nodes = nullptr;
  s->tris = nullptr;
}

SceneView view_of(const DevScene &d) {
  SceneView v;
  v.nodes = d.nodes;
  v.tris = d.tris;
  v.n = d.n;
  v.root = d.root;
  for (int k = 0; k < 3; ++k) {
    v.lo[k] = d.lo[k];
    v.hi[k] = d.hi[k];
  }
  v.absmax = d.absmax;
  return v;
}

// ---- the executor's per-point queries ----------------------------------------------------------

namespace {

__global__ void __launch_bounds__(256) k_first_hit(SceneView sv, const float *orig, const float *dir, int n,
                                                   int retest, float *t_out, uint32_t *prim_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 o = V3{orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]};
  V3 d = V3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
  float t;
  uint32_t f;
  first_hit(sv, o, d, retest, &t, &f);
  t_out[i] = t;
  prim_out[i] = f;
}

__global__ void __launch_bounds__(256) k_inside(SceneView sv, const double *pts, int n, Pose T, int retest,
                                                float *depth, uint32_t *inside) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = moved_f32(T.m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d = 0.0f;
  bool in = inside_depth(sv, p, retest, &d);
  depth[i] = in ? d : 0.0f;
  inside[i] = in ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_bounded(SceneView sv, const double *pts, int n, Pose T, float window,
                                                 float *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = moved_f32(T.m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d;
  out[i] = bounded_distance(sv, p, window, &d) ? d : __builtin_inff();
}

// DistReduce::Min (cpu.rs:100-111): each point searches within the best distance found so far;
// the bits of a non-negative float order as an unsigned integer, so atomicMin is the running min.
__global__ void __launch_bounds__(256) k_bounded_min(SceneView sv, const double *pts, int n, Pose T,
                                                     unsigned int *best) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = moved_f32(T.m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float window = __uint_as_float(__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  float d;
  if (bounded_distance(sv, p, window, &d)) atomicMin(best, __float_as_uint(d));
}

}  // namespace

hipError_t launch_first_hit(const SceneView &sv, const float *orig, const float *dir, int n, int retest, float *t,
                            uint32_t *prim, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_first_hit, dim3(blocks_for(n, 256)), dim3(256), 0, s, sv, orig, dir, n, retest, t, prim);
  return hipGetLastError();
}

hipError_t launch_inside(const SceneView &sv, const double *pts, int n, const Pose &T, int retest, float *depth,
                         uint32_t *inside, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_inside, dim3(blocks_for(n, 256)), dim3(256), 0, s, sv, pts, n, T, retest, depth, inside);
  return hipGetLastError();
}

hipError_t launch_bounded(const SceneView &sv, const double *pts, int n, const Pose &T, float window, float *out,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounded, dim3(blocks_for(n, 256)), dim3(256), 0, s, sv, pts, n, T, window, out);
  return hipGetLastError();
}

hipError_t launch_bounded_min(const SceneView &sv, const double *pts, int n, const Pose &T, unsigned int *best,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounded_min, dim3(blocks_for(n, 256)), dim3(256), 0, s, sv, pts, n, T, best);
  return hipGetLastError();
}

}  // namespace sherd
