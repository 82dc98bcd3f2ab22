// Device LBVH build and the executor's triangle queries (first hit, inside, bounded distance).
//
// Build (replaces parry's TriMesh BVH behind RayScene::new / of_mesh / of_subset,
// crates/sherd-core/src/spatial/bvh.rs:43-82):
//   1. one thread per face: the triangle record, its centroid; two-level bound reductions;
//   2. 30-bit Morton codes of the centroids in the centroid box;
//   3. a stable LSD radix sort (sort.hpp), so equal codes keep face order;
//   4. leaves in sorted order, then the topology bottom-up in one pass (agglomerative LBVH: each
//      node joins the neighbour with the longer common key prefix, the second child to arrive
//      climbs on), then every internal node's box as the union of its leaf range from a doubling
//      table; min/max unions being order-free, the tree is deterministic;
//   5. escape links: the node after subtree [l, r] in pre-order is the right child of split r;
//   6. the nodes moved to pre-order (a node's left child is the next node).
// Parity grids (scene_parity_grids), built on first use by an inside query: per ray axis, a uniform
// grid over the other two coordinates listing the triangles each cell's rays can cross.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

// One wave per bound (13 workgroups of 64): lanes reduce the block partials in stride, then a
// shuffle tree. min and max are order-free, so the result is exact.
__global__ void k_bounds_final(const float *part, int blocks, float *out) {
  int k = blockIdx.x, lane = threadIdx.x;
  bool mx = (k >= 3 && k < 6) || (k >= 9);
  float v = mx ? -__builtin_inff() : __builtin_inff();
  if (k == 12) v = 0.0f;
  for (int b = lane; b < blocks; b += 64) v = mx ? fmaxf(v, part[b * 13 + k]) : fminf(v, part[b * 13 + k]);
  for (int o = 32; o > 0; o >>= 1) {
    float w = __shfl_xor(v, o);
    v = mx ? fmaxf(v, w) : fminf(v, w);
  }
  if (lane == 0) out[k] = v;
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

__global__ void k_leaves(const Tri *in, const uint32_t *order, int m, Tri *out, Node *nodes, int *other, int *root) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  if (j == 0) *root = m == 1 ? 0 : -1;  // a single leaf is the root; otherwise the climb finds it
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

// The agglomerative climb builds the topology only: each leaf's thread climbs while it is the
// second child to reach a node. The atomic exchange on `other` hands the first child's range end
// to the second, which is all the second needs, so no other memory is shared during the climb
// (links and ranges are read by later kernels). The boxes come afterwards from range unions.
__global__ void k_hierarchy(const uint32_t *code, int m, Node *nodes, int *other, int *childb, int *range_l,
                            int *range_r, int *parent_of, int *root) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int cur = m - 1 + j;
  int l = j, r = j;
  while (true) {
    bool left = (l == 0) || (r != m - 1 && delta(code, r) < delta(code, l - 1));
    int parent = left ? r : l - 1;
    if (left)
      nodes[parent].a = cur;
    else
      childb[parent] = cur;
    parent_of[cur] = parent;
    int prev = __hip_atomic_exchange(&other[parent], left ? l : r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev < 0) return;
    if (left) r = prev; else l = prev;
    range_l[parent] = l;
    range_r[parent] = r;
    cur = parent;
    if (l == 0 && r == m - 1) {
      *root = parent;
      return;
    }
  }
}

// Boxes of leaf ranges by doubling (a sparse table): level k holds the union of leaves
// [i, i + 2^k). Level 0 is the leaves' own boxes.
__global__ void k_st_leaves(const Node *nodes, int m, float *st0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Node &nd = nodes[m - 1 + i];
  for (int k = 0; k < 3; ++k) {
    st0[6 * (size_t)i + k] = nd.lo[k];
    st0[6 * (size_t)i + 3 + k] = nd.hi[k];
  }
}

__global__ void k_st_level(const float *prev, int m, int half, float *next) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 2 * half > m) return;
  const float *a = prev + 6 * (size_t)i, *b = prev + 6 * (size_t)(i + half);
  float *o = next + 6 * (size_t)i;
  for (int k = 0; k < 3; ++k) {
    o[k] = fminf(a[k], b[k]);
    o[3 + k] = fmaxf(a[3 + k], b[3 + k]);
  }
}

// An internal node's box: the union of its leaves [L, R], two overlapping table entries (min and
// max are idempotent, so the overlap changes nothing).
__global__ void k_node_boxes(int m, const int *range_l, const int *range_r, const float *st, Node *nodes) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= m - 1) return;
  int L = range_l[x], R = range_r[x];
  int len = R - L + 1;
  int k = 31 - __clz(len);
  const float *lv = st + 6 * (size_t)m * k;
  const float *a = lv + 6 * (size_t)L, *b = lv + 6 * (size_t)(R - (1 << k) + 1);
  Node &nd = nodes[x];
  for (int q = 0; q < 3; ++q) {
    nd.lo[q] = fminf(a[q], b[q]);
    nd.hi[q] = fmaxf(a[3 + q], b[3 + q]);
  }
}

__global__ void k_ropes(int m, Node *nodes, const int *childb, const int *range_r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * m - 1) return;
  int r = i < m - 1 ? range_r[i] : i - (m - 1);
  nodes[i].esc = (r >= m - 1) ? -1 : childb[r];
}

// Pre-order position of every node. A node whose leaves start at L is preceded in pre-order by its
// ancestors and by every node of the subtrees wholly left of it: L leaves and L minus (the number
// of ancestors it lies right of) internal nodes. That totals 2L + (the number of ancestors whose
// left subtree holds it), counted here by walking up the parent links.
__global__ void k_preorder(int m, const Node *nodes, const int *parent_of, const int *range_l, const int *root_at,
                           int *pre) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 2 * m - 1) return;
  int root = *root_at;
  int L = x >= m - 1 ? x - (m - 1) : range_l[x];
  int lt = 0, c = x;
  for (int steps = 0; c != root && c >= 0 && steps < 2 * m; ++steps) {
    int p = parent_of[c];
    lt += nodes[p].a == c ? 1 : 0;
    c = p;
  }
  pre[x] = 2 * L + lt;
}

// Moves every node to its pre-order position: an internal node's left child is then the next
// node, and its `a` becomes the right child's position, so a descent has both children's
// addresses at hand and loads them together. Leaves get a = ~(triangle slot).
__global__ void k_relayout(int m, const Node *in, const int *childb, const int *pre, Node *out) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 2 * m - 1) return;
  Node nd = in[x];
  nd.a = x < m - 1 ? pre[childb[x]] : ~nd.a;
  nd.esc = nd.esc < 0 ? -1 : pre[nd.esc];
  out[pre[x]] = nd;
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
  int pblocks = blocks_for(m, BT);
  if (pblocks > 1024) pblocks = 1024;
  const size_t nn = 2 * (size_t)m - 1, mm = (size_t)m;
  // One workspace for every temporary, carved in 256-byte aligned pieces; the bounds and the root
  // sit together so that one copy brings both back.
  size_t at = 0;
  auto take = [&](size_t bytes) {
    size_t o = at;
    at = (at + bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t o_tree = take(sizeof(Node) * nn), o_parent = take(sizeof(int) * nn), o_pre = take(sizeof(int) * nn);
  size_t o_rl = take(sizeof(int) * mm), o_rr = take(sizeof(int) * mm), o_other = take(sizeof(int) * mm);
  size_t o_childb = take(sizeof(int) * mm), o_raw = take(sizeof(Tri) * mm), o_cent = take(sizeof(float) * 3 * mm);
  size_t o_part = take(sizeof(float) * 13 * pblocks), o_code = take(sizeof(uint32_t) * mm);
  size_t o_val = take(sizeof(uint32_t) * mm), o_tk = take(sizeof(uint32_t) * mm), o_tv = take(sizeof(uint32_t) * mm);
  size_t o_hist = take(sizeof(uint32_t) * rsort::hist_words(m)), o_res = take(sizeof(float) * 16);
  int levels = 1;
  while ((1 << levels) <= m) ++levels;  // table levels 0 .. levels - 1
  size_t o_st = take(sizeof(float) * 6 * mm * (size_t)levels);
  char *ws = nullptr;
  hipError_t err = hipSuccess;
  do {
    if ((err = dalloc(&out->nodes, nn)) != hipSuccess) break;
    if ((err = dalloc(&out->tris, mm)) != hipSuccess) break;
    if ((err = hipMalloc((void **)&ws, at)) != hipSuccess) break;
    Node *tree = (Node *)(ws + o_tree);
    int *parent_of = (int *)(ws + o_parent), *pre = (int *)(ws + o_pre), *range_l = (int *)(ws + o_rl);
    int *range_r = (int *)(ws + o_rr), *other = (int *)(ws + o_other), *childb = (int *)(ws + o_childb);
    Tri *raw = (Tri *)(ws + o_raw);
    float *cent = (float *)(ws + o_cent), *part = (float *)(ws + o_part);
    uint32_t *code = (uint32_t *)(ws + o_code), *val = (uint32_t *)(ws + o_val), *tk = (uint32_t *)(ws + o_tk);
    uint32_t *tv = (uint32_t *)(ws + o_tv), *hist = (uint32_t *)(ws + o_hist);
    float *bounds = (float *)(ws + o_res);  // 13 bounds, then the root at [13]
    int *root = (int *)(bounds + 13);
    hipLaunchKernelGGL(k_tri_setup, dim3(blocks_for(m, BT)), dim3(BT), 0, s, dV, dF, m, raw, cent);
    hipLaunchKernelGGL(k_bounds_partial, dim3(pblocks), dim3(BT), 0, s, raw, cent, m, part);
    hipLaunchKernelGGL(k_bounds_final, dim3(13), dim3(64), 0, s, part, pblocks, bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(m, BT)), dim3(BT), 0, s, cent, m, bounds, code, val);
    if ((err = hipGetLastError()) != hipSuccess) break;
    if ((err = rsort::sort_pairs<uint32_t>(code, val, tk, tv, hist, m, 0, 30, s)) != hipSuccess) break;
    hipLaunchKernelGGL(k_leaves, dim3(blocks_for(m, BT)), dim3(BT), 0, s, raw, val, m, out->tris, tree, other, root);
    if (m > 1) {
      hipLaunchKernelGGL(k_hierarchy, dim3(blocks_for(m, BT)), dim3(BT), 0, s, code, m, tree, other, childb,
                         range_l, range_r, parent_of, root);
    }
    float *st = (float *)(ws + o_st);
    hipLaunchKernelGGL(k_st_leaves, dim3(blocks_for(m, BT)), dim3(BT), 0, s, tree, m, st);
    for (int k = 1; k < levels; ++k)
      hipLaunchKernelGGL(k_st_level, dim3(blocks_for(m, BT)), dim3(BT), 0, s, st + 6 * mm * (size_t)(k - 1), m,
                         1 << (k - 1), st + 6 * mm * (size_t)k);
    if (m > 1)
      hipLaunchKernelGGL(k_node_boxes, dim3(blocks_for(m - 1, BT)), dim3(BT), 0, s, m, range_l, range_r, st, tree);
    hipLaunchKernelGGL(k_ropes, dim3(blocks_for(2L * m - 1, BT)), dim3(BT), 0, s, m, tree, childb, range_r);
    hipLaunchKernelGGL(k_preorder, dim3(blocks_for(2L * m - 1, BT)), dim3(BT), 0, s, m, tree, parent_of, range_l,
                       root, pre);
    hipLaunchKernelGGL(k_relayout, dim3(blocks_for(2L * m - 1, BT)), dim3(BT), 0, s, m, tree, childb, pre,
                       out->nodes);
    if ((err = hipGetLastError()) != hipSuccess) break;
    float hb[14];
    if ((err = hipMemcpyAsync(hb, bounds, sizeof hb, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
    int r;
    std::memcpy(&r, &hb[13], sizeof r);
    if (r < 0 || r >= 2 * m - 1) {
      err = hipErrorUnknown;
      break;
    }
    out->root = 0;
    for (int k = 0; k < 3; ++k) {
      out->lo[k] = hb[6 + k];
      out->hi[k] = hb[9 + k];
    }
    out->absmax = hb[12];
  } while (0);
  if (ws) (void)hipFree(ws);
  if (err != hipSuccess) scene_free(out);
  return err;
}

void scene_free(DevScene *s) {
  if (s->nodes) (void)hipFree(s->nodes);
  if (s->tris) (void)hipFree(s->tris);
  s->nodes = nullptr;
  s->tris = nullptr;
  for (AxisGrid &g : s->pg) {
    if (g.start) (void)hipFree(g.start);
    if (g.ix) (void)hipFree(g.ix);
    if (g.iy) (void)hipFree(g.iy);
    if (g.iz) (void)hipFree(g.iz);
    g = AxisGrid{};
  }
  if (s->db.lb) (void)hipFree(s->db.lb);
  if (s->db.cut) (void)hipFree(s->db.cut);
  s->db = DistBound{};
}

// ---- distance bounds ---------------------------------------------------------------------------
//
// Cells a triangle's (slack-grown) box overlaps are marked; an exact Euclidean distance transform
// (Felzenszwalb and Huttenlocher's lower envelope of parabolas, one axis at a time) gives every
// cell the squared distance d, in cells, from its centre to the nearest marked centre. A point p of
// the cell and a triangle point q in marked cell c' are each within sqrt(3)/2 cells of their
// centres, so |p - q| >= (sqrt(d) - sqrt(3)) h: the stored bound, rounded down.

namespace {

constexpr int DB_MAX_CELLS = 1 << 22;
constexpr float DB_FAR = 1e30f;

__global__ void k_db_fill(float *d, size_t n, float v) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

__global__ void k_db_mark(const Tri *tris, int m, DistBound g, float pad, float *d, uint32_t *diag) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const Tri &t = tris[j];
  int c0[3], c1[3];
  float inv = 1.0f / g.h;
  float e2 = 0.0f;
  for (int k = 0; k < 3; ++k) {
    float lo = fminf(t.v[k], fminf(t.v[3 + k], t.v[6 + k])) - pad;
    float hi = fmaxf(t.v[k], fmaxf(t.v[3 + k], t.v[6 + k])) + pad;
    e2 += (hi - lo) * (hi - lo);
    c0[k] = grid_cell(lo, g.org[k], inv, g.n[k]);
    c1[k] = grid_cell(hi, g.org[k], inv, g.n[k]);
  }
  // The largest padded triangle box diagonal (non-negative f32 bits order as the values).
  atomicMax(diag, __float_as_uint(sqrtf(e2)));
  for (int z = c0[2]; z <= c1[2]; ++z)
    for (int y = c0[1]; y <= c1[1]; ++y)
      for (int x = c0[0]; x <= c1[0]; ++x) d[((size_t)z * g.n[1] + y) * g.n[0] + x] = 0.0f;
}

// One row along `axis` per thread: d(q) <- min_r (f(r) + (q - r)^2), f being the row's values.
// v, fv and zb are this row's work arrays (n, n and n + 1 entries): the lower envelope's
// parabolas (position, value) and the boundaries between them.
__global__ void k_db_edt(float *d, int nx, int ny, int nz, int axis, int *vwork, float *fwork, double *zwork) {
  int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  int rows = axis == 0 ? ny * nz : (axis == 1 ? nx * nz : nx * ny);
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  size_t base, stride;
  if (axis == 0) {
    base = (size_t)r * nx;
    stride = 1;
  } else if (axis == 1) {
    int x = r % nx, z = r / nx;
    base = (size_t)z * ny * nx + x;
    stride = nx;
  } else {
    base = r;
    stride = (size_t)nx * ny;
  }
  int *v = vwork + (size_t)r * n;
  float *fv = fwork + (size_t)r * n;
  double *zb = zwork + (size_t)r * (n + 1);
  int k = -1;
  for (int q = 0; q < n; ++q) {
    float fq = d[base + q * stride];
    if (!(fq < DB_FAR)) continue;
    // Values are integers below 2^24 (exact in f32); the intersections are formed in f64.
    double s = -1e300;
    while (k >= 0) {
      int p = v[k];
      s = (((double)fq + (double)q * q) - ((double)fv[k] + (double)p * p)) / (2.0 * (q - p));
      if (s > zb[k]) break;
      --k;
    }
    ++k;
    v[k] = q;
    fv[k] = fq;
    zb[k] = k == 0 ? -1e300 : s;
    zb[k + 1] = 1e300;
  }
  if (k < 0) return;  // nothing finite in the row: it stays far
  int j = 0;
  for (int q = 0; q < n; ++q) {
    while (zb[j + 1] < (double)q) ++j;
    float dq = (float)(q - v[j]);
    d[base + q * stride] = dq * dq + fv[j];
  }
}

// The tree's cut (DistBound::cut) by one wave: from the root, the frontier node holding the most
// triangles (the lowest lane on a tie) is replaced by its two children until CUT_BOXES nodes are
// reached or only leaves remain. A pre-order subtree from node i to its escape e (or the end)
// holds (e - i + 1) / 2 triangles.
__global__ void k_db_cut(const Node *nodes, int m, int root, float *cut, int *n_cut) {
  const int lane = threadIdx.x;
  const int total = 2 * m - 1;
  int node = lane == 0 ? root : -1;
  int n = 1;
  while (n < CUT_BOXES) {
    int tri = 0;
    if (node >= 0) {
      const Node nd = nodes[node];
      if (nd.a >= 0) tri = ((nd.esc < 0 ? total : nd.esc) - node + 1) / 2;
    }
    int best = tri, who = lane;
    for (int o = 32; o > 0; o >>= 1) {
      int bt = __shfl_xor(best, o), bw = __shfl_xor(who, o);
      if (bt > best || (bt == best && bw < who)) {
        best = bt;
        who = bw;
      }
    }
    if (best < 2) break;  // (uniform) only leaves left
    const int split = __shfl(node, who);
    if (lane == who) node = split + 1;              // the left child is the next node
    if (lane == n) node = nodes[split].a;           // the right child
    ++n;
  }
  if (node >= 0) {
    const Node nd = nodes[node];
    float *o = cut + 8 * lane;
    o[0] = nd.lo[0];
    o[1] = nd.lo[1];
    o[2] = nd.lo[2];
    o[3] = __int_as_float(node);
    o[4] = nd.hi[0];
    o[5] = nd.hi[1];
    o[6] = nd.hi[2];
    o[7] = 0.0f;
  }
  if (lane == 0) *n_cut = n;
}

// Squared cell distances to the stored bounds.
__global__ void k_db_final(float *d, size_t n, float h) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = d[i];
  if (!(v < DB_FAR)) {
    d[i] = DB_FAR;
    return;
  }
  float b = (sqrtf(v) - 1.7320509f) * h * 0.999f;
  d[i] = b > 0.0f ? b : 0.0f;
}

}  // namespace

// ---- parity grids ------------------------------------------------------------------------------

namespace {

constexpr double PG_CELLS_PER_TRI = 2.0;  // cells per triangle of each parity grid
constexpr int PG_MAX_CELLS = 1 << 26;

__device__ __forceinline__ void pg_range(const Tri &t, int axis, const AxisGrid &g, float pad, int *u0, int *u1,
                                         int *v0, int *v1) {
  int U = axis == 0 ? 1 : 0, W = axis == 2 ? 1 : 2;
  float lu = fminf(t.v[U], fminf(t.v[3 + U], t.v[6 + U])) - pad;
  float hu = fmaxf(t.v[U], fmaxf(t.v[3 + U], t.v[6 + U])) + pad;
  float lv = fminf(t.v[W], fminf(t.v[3 + W], t.v[6 + W])) - pad;
  float hv = fmaxf(t.v[W], fmaxf(t.v[3 + W], t.v[6 + W])) + pad;
  *u0 = grid_cell(lu, g.org_u, g.inv, g.nu);
  *u1 = grid_cell(hu, g.org_u, g.inv, g.nu);
  *v0 = grid_cell(lv, g.org_v, g.inv, g.nv);
  *v1 = grid_cell(hv, g.org_v, g.inv, g.nv);
}

__global__ void k_pg_count(const Tri *tris, int m, AxisGrid g, int axis, float pad, uint32_t *count) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int u0, u1, v0, v1;
  pg_range(tris[j], axis, g, pad, &u0, &u1, &v0, &v1);
  for (int cu = u0; cu <= u1; ++cu)
    for (int cv = v0; cv <= v1; ++cv) atomicAdd(&count[(size_t)cu * g.nv + cv], 1u);
}

__global__ void k_pg_fill(const Tri *tris, int m, AxisGrid g, int axis, float pad, uint32_t *cursor) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int u0, u1, v0, v1;
  const Tri t = tris[j];
  pg_range(t, axis, g, pad, &u0, &u1, &v0, &v1);
  const float4 x = make_float4(t.v[0], t.v[1], t.v[2], t.v[3]), y = make_float4(t.v[4], t.v[5], t.v[6], t.v[7]);
  for (int cu = u0; cu <= u1; ++cu)
    for (int cv = v0; cv <= v1; ++cv) {
      uint32_t k = atomicAdd(&cursor[(size_t)cu * g.nv + cv], 1u);
      g.ix[k] = x;
      g.iy[k] = y;
      g.iz[k] = t.v[8];
    }
}

}  // namespace

hipError_t scene_distance_bounds(DevScene *sc, hipStream_t s) {
  if (sc->db.lb) return hipSuccess;
  DistBound g{};
  double ext[3], vol = 1.0;
  for (int k = 0; k < 3; ++k) {
    ext[k] = std::max((double)sc->hi[k] - sc->lo[k], 1e-6 * std::max(1.0, (double)sc->absmax));
    vol *= ext[k];
  }
  double h = std::cbrt(vol / DB_MAX_CELLS);
  for (int k = 0; k < 3; ++k) h = std::max(h, ext[k] / 1000.0);
  for (int tries = 0; tries < 8; ++tries) {
    double cells = 1.0;
    for (int k = 0; k < 3; ++k) cells *= std::ceil(ext[k] / h) + 2.0;
    if (cells <= 1.2 * DB_MAX_CELLS) break;
    h *= 1.05;
  }
  g.h = (float)h;
  size_t ncell = 1;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = (int)std::ceil(ext[k] / h) + 2;
    g.org[k] = (float)((double)sc->lo[k] - h);
    ncell *= (size_t)g.n[k];
  }
  float pad = 1e-5f * sc->absmax + 1e-30f;
  int *vwork = nullptr;
  float *fwork = nullptr;
  double *zwork = nullptr;
  float *diag = nullptr;
  hipError_t err = hipSuccess;
  do {
    if ((err = dalloc(&g.lb, ncell)) != hipSuccess) break;
    if ((err = dalloc(&diag, 2)) != hipSuccess) break;
    hipLaunchKernelGGL(k_db_fill, dim3(1), dim3(BT), 0, s, diag, (size_t)2, 0.0f);
    // Row r of a pass uses entries [r n, r n + n) and [r (n + 1), r (n + 1) + n + 1).
    if ((err = dalloc(&vwork, ncell)) != hipSuccess) break;
    if ((err = dalloc(&fwork, ncell)) != hipSuccess) break;
    if ((err = dalloc(&zwork, 2 * ncell + 1)) != hipSuccess) break;
    hipLaunchKernelGGL(k_db_fill, dim3(blocks_for((long)ncell, BT)), dim3(BT), 0, s, g.lb, ncell, DB_FAR);
    hipLaunchKernelGGL(k_db_mark, dim3(blocks_for(sc->n, BT)), dim3(BT), 0, s, sc->tris, sc->n, g, pad, g.lb,
                       (uint32_t *)diag);
    for (int axis = 0; axis < 3; ++axis) {
      long nrows = (long)(ncell / (size_t)g.n[axis]);
      hipLaunchKernelGGL(k_db_edt, dim3(blocks_for(nrows, 64)), dim3(64), 0, s, g.lb, g.n[0], g.n[1], g.n[2], axis,
                         vwork, fwork, zwork);
    }
    hipLaunchKernelGGL(k_db_final, dim3(blocks_for((long)ncell, BT)), dim3(BT), 0, s, g.lb, ncell, g.h);
    if ((err = dalloc(&g.cut, 8 * CUT_BOXES)) != hipSuccess) break;
    // diag's second word takes the cut's size.
    if (sc->n > 0 && sc->root >= 0)
      hipLaunchKernelGGL(k_db_cut, dim3(1), dim3(64), 0, s, sc->nodes, sc->n, sc->root, g.cut, (int *)diag + 1);
    if ((err = hipGetLastError()) != hipSuccess) break;
    uint32_t hb[2] = {0u, 0u};
    if ((err = hipMemcpyAsync(hb, diag, sizeof hb, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
    float hd;
    std::memcpy(&hd, &hb[0], sizeof hd);
    g.n_cut = (int)hb[1];
    g.ub_add = (float)((2.0 * 1.7320508075688772 * h + (double)hd + 2.0 * (double)pad) * 1.001);
  } while (0);
  if (diag) (void)hipFree(diag);
  if (vwork) (void)hipFree(vwork);
  if (fwork) (void)hipFree(fwork);
  if (zwork) (void)hipFree(zwork);
  if (err != hipSuccess) {
    if (g.lb) (void)hipFree(g.lb);
    if (g.cut) (void)hipFree(g.cut);
    return err;
  }
  sc->db = g;
  return hipSuccess;
}

hipError_t scene_parity_grids(DevScene *sc, hipStream_t s) {
  if (sc->pg[0].start) return hipSuccess;
  int m = sc->n;
  // The slack every parity ray prunes with: its origin is inside the scene's box (the AABB reject
  // comes first), so slack_of gives 1e-5 * absmax.
  float pad = 1e-5f * sc->absmax + 1e-30f;
  hipError_t err = hipSuccess;
  uint32_t *cursor = nullptr;
  for (int axis = 0; axis < 3 && err == hipSuccess; ++axis) {
    int U = axis == 0 ? 1 : 0, W = axis == 2 ? 1 : 2;
    double eu = (double)sc->hi[U] - sc->lo[U], ev = (double)sc->hi[W] - sc->lo[W];
    double ext = std::max(eu, ev);
    double fl = ext > 0.0 ? ext * 1e-6 : 1.0;
    eu = std::max(eu, fl);
    ev = std::max(ev, fl);
    double cell = std::sqrt(eu * ev / (PG_CELLS_PER_TRI * m));
    double nu = std::ceil(eu / cell), nv = std::ceil(ev / cell);
    if (nu * nv > PG_MAX_CELLS) {
      cell *= std::sqrt(nu * nv / PG_MAX_CELLS) * 1.01;
      nu = std::ceil(eu / cell);
      nv = std::ceil(ev / cell);
    }
    AxisGrid g{};
    g.nu = (int)std::max(1.0, nu);
    g.nv = (int)std::max(1.0, nv);
    g.org_u = sc->lo[U];
    g.org_v = sc->lo[W];
    g.inv = (float)(1.0 / cell);
    size_t ncell = (size_t)g.nu * g.nv;
    do {
      if ((err = dalloc(&g.start, ncell + 1)) != hipSuccess) break;
      if ((err = hipMemsetAsync(g.start, 0, sizeof(uint32_t) * (ncell + 1), s)) != hipSuccess) break;
      hipLaunchKernelGGL(k_pg_count, dim3(blocks_for(m, BT)), dim3(BT), 0, s, sc->tris, m, g, axis, pad, g.start);
      if ((err = dalloc(&cursor, std::max(ncell, rsort::scan_parts(ncell + 1)))) != hipSuccess) break;
      if ((err = rsort::scan_exclusive(g.start, ncell + 1, cursor, s)) != hipSuccess) break;
      uint32_t total = 0;
      if ((err = hipMemcpyAsync(&total, g.start + ncell, sizeof total, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
      if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
      if (total > (1u << 31)) {
        err = hipErrorInvalidValue;
        break;
      }
      if ((err = dalloc(&g.ix, total)) != hipSuccess) break;
      if ((err = dalloc(&g.iy, total)) != hipSuccess) break;
      if ((err = dalloc(&g.iz, total)) != hipSuccess) break;
      if ((err = hipMemcpyAsync(cursor, g.start, sizeof(uint32_t) * ncell, hipMemcpyDeviceToDevice, s)) != hipSuccess)
        break;
      hipLaunchKernelGGL(k_pg_fill, dim3(blocks_for(m, BT)), dim3(BT), 0, s, sc->tris, m, g, axis, pad, cursor);
      if ((err = hipGetLastError()) != hipSuccess) break;
      if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
    } while (0);
    if (cursor) (void)hipFree(cursor);
    cursor = nullptr;
    sc->pg[axis] = g;
  }
  if (err != hipSuccess) {
    for (AxisGrid &g : sc->pg) {
      if (g.start) (void)hipFree(g.start);
      if (g.ix) (void)hipFree(g.ix);
      if (g.iy) (void)hipFree(g.iy);
      if (g.iz) (void)hipFree(g.iz);
      g = AxisGrid{};
    }
  }
  return err;
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
  for (int k = 0; k < 3; ++k) v.pg[k] = d.pg[k];
  v.db = d.db;
  return v;
}

// ---- the executor's per-point queries ----------------------------------------------------------

namespace {

__global__ void __launch_bounds__(256) k_first_hit(SceneView sv, const float *orig, const float *dir, int n,
                                                   int retest, float *t_out, uint32_t *prim_out) {
  __shared__ uint2 s_stk[QSTACK][256];
  LdsStack<256, QSTACK> st{&s_stk[0][threadIdx.x]};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 o = V3{orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]};
  V3 d = V3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
  float t;
  uint32_t f;
  first_hit_with(st, sv, o, d, retest, &t, &f);
  t_out[i] = t;
  prim_out[i] = f;
}

__global__ void __launch_bounds__(256) k_inside(SceneView sv, const double *pts, int n, Pose T, int retest,
                                                float *depth, uint32_t *inside) {
  __shared__ uint2 s_stk[QSTACK][256];
  LdsStack<256, QSTACK> st{&s_stk[0][threadIdx.x]};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = moved_f32(T.m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d = 0.0f;
  bool in = inside_depth_with(st, sv, p, retest, &d);
  depth[i] = in ? d : 0.0f;
  inside[i] = in ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_bounded(SceneView sv, const double *pts, int n, Pose T, float window,
                                                 float *out) {
  __shared__ uint2 s_stk[QSTACK][256];
  LdsStack<256, QSTACK> st{&s_stk[0][threadIdx.x]};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = moved_f32(T.m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d;
  out[i] = bounded_distance_with(st, sv, p, window, &d) ? d : __builtin_inff();
}

// DistReduce::Min (cpu.rs:100-111): each point searches within the best distance found so far;
// the bits of a non-negative float order as an unsigned integer, so atomicMin is the running min.
__global__ void __launch_bounds__(256) k_bounded_min(SceneView sv, const double *pts, int n, Pose T,
                                                     unsigned int *best) {
  __shared__ uint2 s_stk[QSTACK][256];
  LdsStack<256, QSTACK> st{&s_stk[0][threadIdx.x]};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = moved_f32(T.m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float window = __uint_as_float(__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  float d;
  if (bounded_distance_with(st, sv, p, window, &d)) atomicMin(best, __float_as_uint(d));
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
