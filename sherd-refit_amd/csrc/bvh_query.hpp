// Stackless LBVH traversal: every query walks the tree in pre-order through the escape links, so
// a thread carries no stack and a wave's lanes stay on one instruction stream per step.
//
// Every query answers exactly what a scan over all triangles would: the boxes are tested grown by
// the pruning slack (slack_of), which exceeds the f32 rounding of the per-triangle arithmetic, so
// no box can hide the triangle that attains the answer.
#pragma once

#include "common.hpp"

namespace sherd {

constexpr int DSTACK = 24;  // deferred nodes of an ordered traversal
constexpr int MSTACK = 16;  // the same for min_d2, whose overflow walk starts from the best found
constexpr int LSTACK = 8;   // min_d2's deferred nodes when they live in LDS (see LdsStack)
constexpr int QSTACK = 16;  // the same for the one-query-per-thread kernels of the executor's calls

// Where min_d2 keeps its deferred nodes. A per-lane array indexed by the depth compiles to
// private (scratch) memory, whose spills go out to HBM; LdsStack is the lane's column of a
// workgroup's LDS array [LSTACK][STRIDE] (STRIDE = the workgroup size, lane-minor, so the lanes of
// a wave touch consecutive words; N entries), and RegStack the per-lane array.
template <int STRIDE, int N = LSTACK>
struct LdsStack {
  static constexpr int CAP = N;
  uint2 *col;  // &lds[0][threadIdx.x]
  __device__ __forceinline__ void put(int k, int node, float d) { col[k * STRIDE] = make_uint2((uint32_t)node, __float_as_uint(d)); }
  __device__ __forceinline__ int node(int k) const { return (int)col[k * STRIDE].x; }
  __device__ __forceinline__ float dist(int k) const { return __uint_as_float(col[k * STRIDE].y); }
};
struct RegStack {
  static constexpr int CAP = MSTACK;
  int n[MSTACK];
  float d[MSTACK];
  __device__ __forceinline__ void put(int k, int node, float dd) {
    n[k] = node;
    d[k] = dd;
  }
  __device__ __forceinline__ int node(int k) const { return n[k]; }
  __device__ __forceinline__ float dist(int k) const { return d[k]; }
};

__device__ __forceinline__ Node load_node(const Node *nodes, int i) {
  const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
  float4 a = gld4(p), b = gld4(p + 1);
  Node n;
  n.lo[0] = a.x; n.lo[1] = a.y; n.lo[2] = a.z; n.a = __float_as_int(a.w);
  n.hi[0] = b.x; n.hi[1] = b.y; n.hi[2] = b.z; n.esc = __float_as_int(b.w);
  return n;
}

__device__ __forceinline__ void load_tri(const Tri *tris, int j, V3 *a, V3 *b, V3 *c, uint32_t *face) {
  const float4 *p = reinterpret_cast<const float4 *>(tris + j);
  float4 x = gld4(p), y = gld4(p + 1), z = gld4(p + 2);
  *a = V3{x.x, x.y, x.z};
  *b = V3{x.w, y.x, y.y};
  *c = V3{y.z, y.w, z.x};
  *face = __float_as_uint(z.y);
}

// Slab test of a grown box against origin + t dir, t in [0, tmax]; exact zero components handled
// by containment. Returns the entry parameter, or a negative value on a miss.
__device__ __forceinline__ float ray_enter(const Node &nd, V3 o, V3 d, V3 inv, float pad, float tmax) {
  float t0 = 0.0f, t1 = tmax;
  float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, ii[3] = {inv.x, inv.y, inv.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float lo = nd.lo[k] - pad, hi = nd.hi[k] + pad;
    if (dd[k] == 0.0f) {
      if (oo[k] < lo || oo[k] > hi) return -1.0f;
    } else {
      float a = (lo - oo[k]) * ii[k], b = (hi - oo[k]) * ii[k];
      float mn = a < b ? a : b, mx = a < b ? b : a;
      t0 = t0 < mn ? mn : t0;
      t1 = t1 < mx ? t1 : mx;
    }
  }
  return t0 <= t1 ? t0 : -1.0f;
}

// First hit: smallest t, ties to the lowest scene-local face (bvh.rs:84-91 contract).
__device__ inline void first_hit_stackless(const SceneView &sv, V3 o, V3 d, int retest, float *t_out,
                                            uint32_t *face_out) {
  float best = __builtin_inff();
  uint32_t bface = 0xFFFFFFFFu;
  float pad = slack_of(sv.absmax, o);
  float dm = absf(d.x);
  dm = dm < absf(d.y) ? absf(d.y) : dm;
  dm = dm < absf(d.z) ? absf(d.z) : dm;
  float tpad = pad / (dm > 0.0f ? dm : 1.0f);
  V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  int nodes = 2 * sv.n - 1, steps = 0;
  int i = sv.root;
  // Pre-order visits each node at most once; the cap only bounds a corrupted tree.
  while (i >= 0 && i < nodes && steps++ < nodes) {
    Node nd = load_node(sv.nodes, i);
    if (nd.a >= 0) {
      float te = ray_enter(nd, o, d, inv, pad, best + tpad);
      i = te >= 0.0f ? i + 1 : nd.esc;
    } else {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, ~nd.a, &a, &b, &c, &face);
      float t;
      if (mt_hit(o, d, a, b, c, retest, &t)) {
        if (t < best || (t == best && face < bface)) {
          best = t;
          bface = face;
        }
      }
      i = nd.esc;
    }
  }
  *t_out = best;
  *face_out = bface;
}

// The same first hit, front to back: both children are entered by their entry parameter, the
// nearer first, the farther deferred and dropped once a hit closer than its entry is known. A full
// short stack hands the ray to the exhaustive walk above, so the answer is the same.
template <class Stack>
__device__ inline void first_hit_with(Stack &st, const SceneView &sv, V3 o, V3 d, int retest, float *t_out,
                                      uint32_t *face_out) {
  float best = __builtin_inff();
  uint32_t bface = 0xFFFFFFFFu;
  float pad = slack_of(sv.absmax, o);
  float dm = absf(d.x);
  dm = dm < absf(d.y) ? absf(d.y) : dm;
  dm = dm < absf(d.z) ? absf(d.z) : dm;
  float tpad = pad / (dm > 0.0f ? dm : 1.0f);
  V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  *t_out = best;
  *face_out = bface;
  if (cur < 0 || cur >= nodes) return;
  Node nd = load_node(sv.nodes, cur);
  if (ray_enter(nd, o, d, inv, pad, best) < 0.0f) return;
  int sp = 0;
  bool overflow = false;
  for (int iter = 0; iter <= nodes; ++iter) {
    int next = -1;
    if (nd.a < 0) {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, ~nd.a, &a, &b, &c, &face);
      float t;
      if (mt_hit(o, d, a, b, c, retest, &t)) {
        if (t < best || (t == best && face < bface)) {
          best = t;
          bface = face;
        }
      }
    } else if (nd.a > cur && nd.a < nodes) {
      int l = cur + 1, r = nd.a;
      Node ln = load_node(sv.nodes, l), rn = load_node(sv.nodes, r);
      float tmax = best + tpad;
      float tl = ray_enter(ln, o, d, inv, pad, tmax);
      float tr = ray_enter(rn, o, d, inv, pad, tmax);
      bool hl = tl >= 0.0f, hr = tr >= 0.0f;
      if (hl && hr) {
        bool lf = tl <= tr;
        if (sp < Stack::CAP) {
          st.put(sp, lf ? r : l, lf ? tr : tl);
          ++sp;
        } else {
          overflow = true;
        }
        next = lf ? l : r;
        nd = lf ? ln : rn;
      } else if (hl) {
        next = l;
        nd = ln;
      } else if (hr) {
        next = r;
        nd = rn;
      }
    }
    if (next < 0) {
      while (next < 0 && sp > 0) {
        --sp;
        if (st.dist(sp) <= best + tpad) next = st.node(sp);
      }
      if (next < 0) break;
      nd = load_node(sv.nodes, next);
    }
    cur = next;
  }
  if (overflow) {
    first_hit_stackless(sv, o, d, retest, &best, &bface);
  }
  *t_out = best;
  *face_out = bface;
}

// A first hit's deferred nodes in a per-lane array (DSTACK entries).
struct RayRegStack {
  static constexpr int CAP = DSTACK;
  int n[DSTACK];
  float d[DSTACK];
  __device__ __forceinline__ void put(int k, int node, float dd) {
    n[k] = node;
    d[k] = dd;
  }
  __device__ __forceinline__ int node(int k) const { return n[k]; }
  __device__ __forceinline__ float dist(int k) const { return d[k]; }
};
__device__ inline void first_hit(const SceneView &sv, V3 o, V3 d, int retest, float *t_out, uint32_t *face_out) {
  RayRegStack st;
  first_hit_with(st, sv, o, d, retest, t_out, face_out);
}

// Crossings of the ray from o along +AXIS (RayScene::crossings, bvh.rs:160-178): every hit, from
// the triangles of the ray's cell in the scene's parity grid for that axis. The cell lists every
// triangle whose slack-grown box the ray's line meets, so this counts what an exhaustive scan
// would. Triangles wholly behind the origin (by more than the slack) cannot be hit and are skipped
// before the intersection test. *nearest gets the smallest hit parameter (+inf without a hit).
//
// A cell's items are independent of one another, so they are taken XCH at a time with every load
// of the group issued before any test: a long cell (a wall seen edge-on lists hundreds) then costs
// a few memory round trips per group. The grid keeps each item's coordinates in its cell's run, so
// a ray reads its cell as three contiguous arrays with no index to follow (scattered 48-byte
// records behind an index list moved 3.6 GB per collection step). The count and the minimum do
// not depend on the order.
constexpr int XCH = 4;

template <int AXIS>
__device__ __forceinline__ void crossing_test(V3 o, int retest, float pad, const float4 &x, const float4 &y, float zx,
                                              uint32_t *hits, float *tmin) {
  V3 v0 = V3{x.x, x.y, x.z}, v1 = V3{x.w, y.x, y.y}, v2 = V3{y.z, y.w, zx};
  float oa = AXIS == 0 ? o.x : (AXIS == 1 ? o.y : o.z);
  float x0 = AXIS == 0 ? v0.x : (AXIS == 1 ? v0.y : v0.z);
  float x1 = AXIS == 0 ? v1.x : (AXIS == 1 ? v1.y : v1.z);
  float x2 = AXIS == 0 ? v2.x : (AXIS == 1 ? v2.y : v2.z);
  if (fmaxf(x0, fmaxf(x1, x2)) + pad < oa) return;
  V3 d = V3{AXIS == 0 ? 1.0f : 0.0f, AXIS == 1 ? 1.0f : 0.0f, AXIS == 2 ? 1.0f : 0.0f};
  float t;
  if (mt_hit(o, d, v0, v1, v2, retest, &t)) {
    ++*hits;
    *tmin = t < *tmin ? t : *tmin;
  }
}

template <int AXIS>
__device__ inline uint32_t crossings_axis(const SceneView &sv, V3 o, int retest, float *nearest) {
  constexpr int U = AXIS == 0 ? 1 : 0, W = AXIS == 2 ? 1 : 2;
  const AxisGrid &g = sv.pg[AXIS];
  float oo[3] = {o.x, o.y, o.z};
  float pad = slack_of(sv.absmax, o);
  int cell = grid_cell(oo[U], g.org_u, g.inv, g.nu) * g.nv + grid_cell(oo[W], g.org_v, g.inv, g.nv);
  uint32_t b = gldu(g.start + cell), e = gldu(g.start + cell + 1);
  uint32_t hits = 0;
  float tmin = __builtin_inff();
  uint32_t k = b;
  for (; k + XCH <= e; k += XCH) {
    float4 x[XCH], y[XCH];
    float zx[XCH];
#pragma unroll
    for (int u = 0; u < XCH; ++u) {
      x[u] = gld4(g.ix + k + u);
      y[u] = gld4(g.iy + k + u);
      zx[u] = gldf(g.iz + k + u);
    }
#pragma unroll
    for (int u = 0; u < XCH; ++u) crossing_test<AXIS>(o, retest, pad, x[u], y[u], zx[u], &hits, &tmin);
  }
  for (; k < e; ++k) crossing_test<AXIS>(o, retest, pad, gld4(g.ix + k), gld4(g.iy + k), gldf(g.iz + k), &hits, &tmin);
  *nearest = tmin;
  return hits;
}

// How many triangles the parity grid lists in the cell of the ray from o along +AXIS.
template <int AXIS>
__device__ __forceinline__ uint32_t cell_items(const SceneView &sv, V3 o) {
  constexpr int U = AXIS == 0 ? 1 : 0, W = AXIS == 2 ? 1 : 2;
  const AxisGrid &g = sv.pg[AXIS];
  float oo[3] = {o.x, o.y, o.z};
  int cell = grid_cell(oo[U], g.org_u, g.inv, g.nu) * g.nv + grid_cell(oo[W], g.org_v, g.inv, g.nv);
  return gldu(g.start + cell + 1) - gldu(g.start + cell);
}

__device__ __forceinline__ uint32_t crossings_on(int axis, const SceneView &sv, V3 o, int retest, float *nearest) {
  if (axis == 0) return crossings_axis<0>(sv, o, retest, nearest);
  if (axis == 1) return crossings_axis<1>(sv, o, retest, nearest);
  return crossings_axis<2>(sv, o, retest, nearest);
}

// Smallest squared distance to the scene among triangles within `window` (all when window is
// +inf); +inf when none is. Pruned against min(best, window) grown by the slack.
__device__ inline float min_d2_stackless(const SceneView &sv, V3 p, float window) {
  float best = __builtin_inff();
  float slack = slack_of(sv.absmax, p);
  float lim = window + slack;
  float lim2 = lim * lim;
  int nodes = 2 * sv.n - 1, steps = 0;
  int i = sv.root;
  // Pre-order visits each node at most once; the cap only bounds a corrupted tree.
  while (i >= 0 && i < nodes && steps++ < nodes) {
    Node nd = load_node(sv.nodes, i);
    if (nd.a >= 0) {
      i = box_d2(nd, p, slack) <= lim2 ? i + 1 : nd.esc;
    } else {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, ~nd.a, &a, &b, &c, &face);
      float d2 = tri_d2(p, a, b, c);
      if (d2 < best) {
        best = d2;
        float l = mnf(sqrtf(best), window) + slack;
        lim2 = l * l;
      }
      i = nd.esc;
    }
  }
  return best;
}

// The same minimum, nearest child first: both children's boxes are measured, the nearer is
// entered and the farther deferred on a short stack, and a deferred node is dropped when the
// minimum found since has moved past it. Both children are loaded together (the left is the next
// node, the right is the parent's `a`) and the entered child's node is carried into the next step,
// so a level of the descent costs one memory round trip. Should the stack ever fill, the
// exhaustive walk above finishes the query, so the answer is the same.
//
// With a step budget (steps > 0) the walk gives up after that many nodes: *complete is then false
// and the result is only the smallest distance met so far, one the point really attains.
template <class Stack>
__device__ inline float min_d2_with(Stack &st, const SceneView &sv, V3 p, float window, int steps, bool *complete) {
  if (complete) *complete = true;
  float best = __builtin_inff();
  float slack = slack_of(sv.absmax, p);
  float lim = window + slack;
  float lim2 = lim * lim;
  int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  if (cur < 0 || cur >= nodes) return best;
  Node nd = load_node(sv.nodes, cur);
  if (box_d2(nd, p, slack) > lim2) return best;
  int sp = 0;
  bool overflow = false;
  for (int iter = 0; iter <= nodes; ++iter) {
    if (steps > 0 && iter >= steps) {
      *complete = false;
      return best;
    }
    int next = -1;
    if (nd.a < 0) {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, ~nd.a, &a, &b, &c, &face);
      float d2 = tri_d2(p, a, b, c);
      if (d2 < best) {
        best = d2;
        float l = mnf(sqrtf(best), window) + slack;
        lim2 = l * l;
      }
    } else if (nd.a > cur && nd.a < nodes) {
      int l = cur + 1, r = nd.a;
      Node ln = load_node(sv.nodes, l), rn = load_node(sv.nodes, r);
      float dl = box_d2(ln, p, slack), dr = box_d2(rn, p, slack);
      bool hl = dl <= lim2, hr = dr <= lim2;
      if (hl && hr) {
        bool lf = dl <= dr;
        if (sp < Stack::CAP) {
          st.put(sp, lf ? r : l, lf ? dr : dl);
          ++sp;
        } else {
          overflow = true;
        }
        next = lf ? l : r;
        nd = lf ? ln : rn;
      } else if (hl) {
        next = l;
        nd = ln;
      } else if (hr) {
        next = r;
        nd = rn;
      }
    }
    if (next < 0) {
      while (next < 0 && sp > 0) {
        --sp;
        if (st.dist(sp) <= lim2) next = st.node(sp);
      }
      if (next < 0) break;
      nd = load_node(sv.nodes, next);
    }
    cur = next;
  }
  if (overflow) {
    // The nearest triangle is no farther than the best met so far, so the exhaustive walk needs
    // only that bound.
    float e = min_d2_stackless(sv, p, mnf(sqrtf(best), window));
    best = e < best ? e : best;
  }
  return best;
}

__device__ inline float min_d2(const SceneView &sv, V3 p, float window, int steps = 0, bool *complete = nullptr) {
  RegStack st;
  return min_d2_with(st, sv, p, window, steps, complete);
}

// The minimum over the triangles of node c's subtree within window (+inf when none is): the
// stackless pre-order walk from c to the end of its subtree, which is where c's escape link points
// (or the end of the tree): every escape taken inside the subtree lands at most there.
__device__ inline float min_d2_subtree(const SceneView &sv, int c, V3 p, float window) {
  float best = __builtin_inff();
  float slack = slack_of(sv.absmax, p);
  float lim = window + slack;
  float lim2 = lim * lim;
  int nodes = 2 * sv.n - 1, steps = 0;
  if (c < 0 || c >= nodes) return best;
  const int end = load_node(sv.nodes, c).esc;
  for (int i = c; i >= 0 && i < nodes && i != end && steps++ < nodes;) {
    Node nd = load_node(sv.nodes, i);
    if (nd.a >= 0) {
      i = box_d2(nd, p, slack) <= lim2 ? i + 1 : nd.esc;
    } else {
      V3 a, b, cc;
      uint32_t face;
      load_tri(sv.tris, ~nd.a, &a, &b, &cc, &face);
      float d2 = tri_d2(p, a, b, cc);
      if (d2 < best) {
        best = d2;
        float l = mnf(sqrtf(best), window) + slack;
        lim2 = l * l;
      }
      i = nd.esc;
    }
  }
  return best;
}

// A squared distance from p to one triangle of the scene, found by descending into the child of
// nearer box at every node (+inf for an empty scene): an upper bound of the minimum, usually close
// to it, for one tree depth of node fetches.
__device__ inline float greedy_d2(const SceneView &sv, V3 p) {
  const int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  if (cur < 0 || cur >= nodes) return __builtin_inff();
  Node nd = load_node(sv.nodes, cur);
  for (int steps = 0; nd.a >= 0 && steps < nodes; ++steps) {
    const int l = cur + 1, r = nd.a;
    if (!(r > cur && r < nodes)) return __builtin_inff();
    const Node ln = load_node(sv.nodes, l), rn = load_node(sv.nodes, r);
    const bool left = box_d2(ln, p, 0.0f) <= box_d2(rn, p, 0.0f);
    cur = left ? l : r;
    nd = left ? ln : rn;
  }
  if (nd.a >= 0) return __builtin_inff();
  V3 a, b, c;
  uint32_t face;
  load_tri(sv.tris, ~nd.a, &a, &b, &c, &face);
  return tri_d2(p, a, b, c);
}

// min_d2 for one point by a whole wave: p and window are wave-uniform, and the wave keeps its
// pending nodes in a stack in LDS (stk, cap entries, this wave's own), expanding up to 64 of them
// per step, one per lane. Children are pushed when their box is within the current bound, leaves
// are measured, and the wave's minimum tightens the bound. The result is the same minimum; a query
// whose frontier would overflow the stack has the overflowing lanes finish their subtrees alone.
// Every lane of the wave calls it.
//
// max_steps bounds the steps (each up to 64 nodes); a query that reaches it stops with *complete
// false and the smallest distance met so far (one the point attains).
__device__ inline float min_d2_wave(const SceneView &sv, V3 p, float window, int *stk, int cap, int max_steps,
                                    bool *complete) {
  *complete = true;
  int lane = threadIdx.x & 63;
  float slack = slack_of(sv.absmax, p);
  float best = __builtin_inff();
  int nodes = 2 * sv.n - 1;
  if (sv.root < 0 || sv.root >= nodes) return best;
  {
    float lim = window + slack;
    if (box_d2(load_node(sv.nodes, sv.root), p, slack) > lim * lim) return best;
  }
  if (lane == 0) stk[0] = sv.root;
  int sp = 1;
  for (int guard = 0; sp > 0 && guard <= nodes; ++guard) {
    if (guard >= max_steps) {
      *complete = false;
      break;
    }
    float lim = mnf(sqrtf(best), window) + slack;
    float lim2 = lim * lim;
    int take = sp < 64 ? sp : 64;
    int node = lane < take ? stk[sp - take + lane] : -1;
    sp -= take;
    __builtin_amdgcn_wave_barrier();
    float mine = __builtin_inff();
    int c0 = -1, c1 = -1;
    if (node >= 0 && node < nodes) {
      Node nd = load_node(sv.nodes, node);
      if (nd.a < 0) {
        V3 a, b, c;
        uint32_t face;
        load_tri(sv.tris, ~nd.a, &a, &b, &c, &face);
        mine = tri_d2(p, a, b, c);
      } else if (nd.a > node && nd.a < nodes) {
        int l = node + 1, r = nd.a;
        Node ln = load_node(sv.nodes, l), rn = load_node(sv.nodes, r);
        if (box_d2(ln, p, slack) <= lim2) c0 = l;
        if (box_d2(rn, p, slack) <= lim2) c1 = r;
      }
    }
    uint32_t wb = wave_min(__float_as_uint(mine));
    float wm = __uint_as_float(wb);
    best = wm < best ? wm : best;
    uint64_t m0 = __ballot(c0 >= 0), m1 = __ballot(c1 >= 0);
    int n0 = (int)__popcll(m0), n1 = (int)__popcll(m1);
    if (sp + n0 + n1 <= cap) {
      if (c0 >= 0) stk[sp + (int)lanes_below(m0)] = c0;
      if (c1 >= 0) stk[sp + n0 + (int)lanes_below(m1)] = c1;
      sp += n0 + n1;
    } else {
      // No room: each lane walks its own children's subtrees (the same bound).
      float w = mnf(sqrtf(best), window);
      float own = __builtin_inff();
      if (c0 >= 0) own = min_d2_subtree(sv, c0, p, w);
      if (c1 >= 0) {
        float e = min_d2_subtree(sv, c1, p, w);
        own = e < own ? e : own;
      }
      uint32_t ob = wave_min(__float_as_uint(own));
      float om = __uint_as_float(ob);
      best = om < best ? om : best;
    }
    __builtin_amdgcn_wave_barrier();
  }
  return best;
}

// RayScene::bounded_distance (bvh.rs:123-135): the distance when it is at most max_dist.
template <class Stack>
__device__ __forceinline__ bool bounded_distance_with(Stack &st, const SceneView &sv, V3 p, float max_dist, float *d) {
  if (!(max_dist > 0.0f)) return false;
  float dist = sqrtf(min_d2_with(st, sv, p, max_dist, 0, nullptr));
  if (!(dist <= max_dist)) return false;
  *d = dist;
  return true;
}

// CpuExecutor::inside's per-point body (cpu.rs:182-188, bvh.rs:146-158) in two halves: the AABB
// reject and three axis rays with a majority of odd counts (inside_parity), then the unbounded
// distance of an inside point (depth_within).
//
// The parity half: whether p is inside and, when it is, the nearest ray crossing's distance in *t
// (+inf if no ray crossed anything).
__device__ __forceinline__ bool inside_parity(const SceneView &sv, V3 p, int retest, float *t) {
  if (p.x < sv.lo[0] || p.x > sv.hi[0] || p.y < sv.lo[1] || p.y > sv.hi[1] || p.z < sv.lo[2] ||
      p.z > sv.hi[2])
    return false;
  // Majority of three odd counts: any two agreeing rays decide it, and the third is cast only when
  // the first two disagree. The two rays whose cells list the fewest triangles go first (a wall
  // seen edge-on fills one axis's cell with hundreds); the vote does not depend on the order.
  uint32_t n[3] = {cell_items<0>(sv, p), cell_items<1>(sv, p), cell_items<2>(sv, p)};
  int hi = n[2] >= n[1] ? (n[2] >= n[0] ? 2 : 0) : (n[1] >= n[0] ? 1 : 0);
  int a0 = hi == 0 ? 1 : 0, a1 = hi == 2 ? 1 : 2;
  float t0, t1, t2 = __builtin_inff();
  int v0 = (int)(crossings_on(a0, sv, p, retest, &t0) & 1u);
  int v1 = (int)(crossings_on(a1, sv, p, retest, &t1) & 1u);
  int votes = v0 + v1;
  if (v0 != v1) votes += (int)(crossings_on(hi, sv, p, retest, &t2) & 1u);
  if (votes < 2) return false;
  *t = fminf(t0, fminf(t1, t2));
  return true;
}

// The search bound of an inside point's depth from its nearest crossing t: a ray's nearest crossing
// is a point of the surface t away, so the nearest triangle is no farther than t, and t grown by
// a relative 1e-4 and the slack (which exceed the f32 rounding of the crossing and of the
// distance) bounds its distance as computed. +inf when t is.
__device__ __forceinline__ float depth_bound(const SceneView &sv, V3 p, float t) {
  return t < __builtin_inff() ? t * 1.0001f + slack_of(sv.absmax, p) : __builtin_inff();
}

// The depth of an inside point, the distance to its nearest triangle, searched within the bound
// wb = depth_bound(t) (which finds the same minimum: every triangle within it is measured); only
// if that bound somehow missed does the unbounded search run.
template <class Stack>
__device__ __forceinline__ float depth_within(Stack &st, const SceneView &sv, V3 p, float wb) {
  float d2 = __builtin_inff();
  if (wb < __builtin_inff()) {
    d2 = min_d2_with(st, sv, p, wb, 0, nullptr);
    if (!(d2 <= wb * wb)) d2 = __builtin_inff();
  }
  if (d2 == __builtin_inff()) d2 = min_d2_with(st, sv, p, __builtin_inff(), 0, nullptr);
  return sqrtf(d2);
}

// Both halves: inside and, for an inside point, its depth.
template <class Stack>
__device__ __forceinline__ bool inside_depth_with(Stack &st, const SceneView &sv, V3 p, int retest, float *depth) {
  float t;
  if (!inside_parity(sv, p, retest, &t)) return false;
  *depth = depth_within(st, sv, p, depth_bound(sv, p, t));
  return true;
}

__device__ __forceinline__ bool bounded_distance(const SceneView &sv, V3 p, float max_dist, float *d) {
  RegStack st;
  return bounded_distance_with(st, sv, p, max_dist, d);
}
__device__ __forceinline__ bool inside_depth(const SceneView &sv, V3 p, int retest, float *depth) {
  RegStack st;
  return inside_depth_with(st, sv, p, retest, depth);
}

}  // namespace sherd
