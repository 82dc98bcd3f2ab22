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

__device__ __forceinline__ Node load_node(const Node *nodes, int i) {
  const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
  float4 a = p[0], b = p[1];
  Node n;
  n.lo[0] = a.x; n.lo[1] = a.y; n.lo[2] = a.z; n.a = __float_as_int(a.w);
  n.hi[0] = b.x; n.hi[1] = b.y; n.hi[2] = b.z; n.esc = __float_as_int(b.w);
  return n;
}

__device__ __forceinline__ void load_tri(const Tri *tris, int j, V3 *a, V3 *b, V3 *c, uint32_t *face) {
  const float4 *p = reinterpret_cast<const float4 *>(tris + j);
  float4 x = p[0], y = p[1], z = p[2];
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
  int leaf0 = sv.n - 1;
  int nodes = 2 * sv.n - 1, steps = 0;
  int i = sv.root;
  // Pre-order visits each node at most once; the cap only bounds a corrupted tree.
  while (i >= 0 && i < nodes && steps++ < nodes) {
    Node nd = load_node(sv.nodes, i);
    if (i < leaf0) {
      float te = ray_enter(nd, o, d, inv, pad, best + tpad);
      i = te >= 0.0f ? nd.a : nd.esc;
    } else {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, nd.a, &a, &b, &c, &face);
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
__device__ inline void first_hit(const SceneView &sv, V3 o, V3 d, int retest, float *t_out, uint32_t *face_out) {
  float best = __builtin_inff();
  uint32_t bface = 0xFFFFFFFFu;
  float pad = slack_of(sv.absmax, o);
  float dm = absf(d.x);
  dm = dm < absf(d.y) ? absf(d.y) : dm;
  dm = dm < absf(d.z) ? absf(d.z) : dm;
  float tpad = pad / (dm > 0.0f ? dm : 1.0f);
  V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  int leaf0 = sv.n - 1;
  int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  *t_out = best;
  *face_out = bface;
  if (cur < 0 || cur >= nodes) return;
  if (ray_enter(load_node(sv.nodes, cur), o, d, inv, pad, best) < 0.0f) return;
  int st_n[DSTACK];
  float st_t[DSTACK];
  int sp = 0;
  bool overflow = false;
  for (int iter = 0; iter <= nodes; ++iter) {
    Node nd = load_node(sv.nodes, cur);
    int next = -1;
    if (cur >= leaf0) {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, nd.a, &a, &b, &c, &face);
      float t;
      if (mt_hit(o, d, a, b, c, retest, &t)) {
        if (t < best || (t == best && face < bface)) {
          best = t;
          bface = face;
        }
      }
    } else {
      int l = nd.a;
      Node ln = load_node(sv.nodes, l);
      int r = ln.esc;
      float tmax = best + tpad;
      float tl = ray_enter(ln, o, d, inv, pad, tmax);
      float tr = -1.0f;
      if (r >= 0 && r < nodes) tr = ray_enter(load_node(sv.nodes, r), o, d, inv, pad, tmax);
      bool hl = tl >= 0.0f, hr = tr >= 0.0f;
      if (hl && hr) {
        int near = tl <= tr ? l : r, far = tl <= tr ? r : l;
        float tfar = tl <= tr ? tr : tl;
        if (sp < DSTACK) {
          st_n[sp] = far;
          st_t[sp] = tfar;
          ++sp;
        } else {
          overflow = true;
        }
        next = near;
      } else if (hl) {
        next = l;
      } else if (hr) {
        next = r;
      }
    }
    while (next < 0 && sp > 0) {
      --sp;
      if (st_t[sp] <= best + tpad) next = st_n[sp];
    }
    if (next < 0) break;
    cur = next;
  }
  if (overflow) {
    first_hit_stackless(sv, o, d, retest, &best, &bface);
  }
  *t_out = best;
  *face_out = bface;
}

// Crossings of the ray from o along +axis (RayScene::crossings, bvh.rs:160-178): every hit.
__device__ inline uint32_t crossings_axis(const SceneView &sv, V3 o, int axis, int retest) {
  float pad = slack_of(sv.absmax, o);
  float oo[3] = {o.x, o.y, o.z};
  int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  V3 d = V3{axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
  uint32_t hits = 0;
  int leaf0 = sv.n - 1;
  int nodes = 2 * sv.n - 1, steps = 0;
  int i = sv.root;
  // Pre-order visits each node at most once; the cap only bounds a corrupted tree.
  while (i >= 0 && i < nodes && steps++ < nodes) {
    Node nd = load_node(sv.nodes, i);
    if (i < leaf0) {
      bool in = oo[a1] >= nd.lo[a1] - pad && oo[a1] <= nd.hi[a1] + pad && oo[a2] >= nd.lo[a2] - pad &&
                oo[a2] <= nd.hi[a2] + pad && nd.hi[axis] + pad >= oo[axis];
      i = in ? nd.a : nd.esc;
    } else {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, nd.a, &a, &b, &c, &face);
      float t;
      hits += (uint32_t)mt_hit(o, d, a, b, c, retest, &t);
      i = nd.esc;
    }
  }
  return hits;
}

// Smallest squared distance to the scene among triangles within `window` (all when window is
// +inf); +inf when none is. Pruned against min(best, window) grown by the slack.
__device__ inline float min_d2_stackless(const SceneView &sv, V3 p, float window) {
  float best = __builtin_inff();
  float slack = slack_of(sv.absmax, p);
  float lim = window + slack;
  float lim2 = lim * lim;
  int leaf0 = sv.n - 1;
  int nodes = 2 * sv.n - 1, steps = 0;
  int i = sv.root;
  // Pre-order visits each node at most once; the cap only bounds a corrupted tree.
  while (i >= 0 && i < nodes && steps++ < nodes) {
    Node nd = load_node(sv.nodes, i);
    if (i < leaf0) {
      i = box_d2(nd, p, slack) <= lim2 ? nd.a : nd.esc;
    } else {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, nd.a, &a, &b, &c, &face);
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
// minimum found since has moved past it. In this layout the right child of an internal node is the
// escape link of its left child, so no child pointer beyond the tree's own is needed. Should the
// stack ever fill, the exhaustive walk above finishes the query, so the answer is the same.
__device__ inline float min_d2(const SceneView &sv, V3 p, float window) {
  float best = __builtin_inff();
  float slack = slack_of(sv.absmax, p);
  float lim = window + slack;
  float lim2 = lim * lim;
  int leaf0 = sv.n - 1;
  int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  if (cur < 0 || cur >= nodes) return best;
  if (box_d2(load_node(sv.nodes, cur), p, slack) > lim2) return best;
  int st_n[DSTACK];
  float st_d[DSTACK];
  int sp = 0;
  bool overflow = false;
  for (int iter = 0; iter <= nodes; ++iter) {
    Node nd = load_node(sv.nodes, cur);
    int next = -1;
    if (cur >= leaf0) {
      V3 a, b, c;
      uint32_t face;
      load_tri(sv.tris, nd.a, &a, &b, &c, &face);
      float d2 = tri_d2(p, a, b, c);
      if (d2 < best) {
        best = d2;
        float l = mnf(sqrtf(best), window) + slack;
        lim2 = l * l;
      }
    } else {
      int l = nd.a;
      Node ln = load_node(sv.nodes, l);
      int r = ln.esc;
      float dl = box_d2(ln, p, slack);
      float dr = __builtin_inff();
      if (r >= 0 && r < nodes) dr = box_d2(load_node(sv.nodes, r), p, slack);
      bool hl = dl <= lim2, hr = dr <= lim2;
      if (hl && hr) {
        int near = dl <= dr ? l : r, far = dl <= dr ? r : l;
        float dfar = dl <= dr ? dr : dl;
        if (sp < DSTACK) {
          st_n[sp] = far;
          st_d[sp] = dfar;
          ++sp;
        } else {
          overflow = true;
        }
        next = near;
      } else if (hl) {
        next = l;
      } else if (hr) {
        next = r;
      }
    }
    while (next < 0 && sp > 0) {
      --sp;
      if (st_d[sp] <= lim2) next = st_n[sp];
    }
    if (next < 0) break;
    cur = next;
  }
  if (overflow) {
    float e = min_d2_stackless(sv, p, window);
    best = e < best ? e : best;
  }
  return best;
}

// RayScene::bounded_distance (bvh.rs:123-135): the distance when it is at most max_dist.
__device__ __forceinline__ bool bounded_distance(const SceneView &sv, V3 p, float max_dist, float *d) {
  if (!(max_dist > 0.0f)) return false;
  float dist = sqrtf(min_d2(sv, p, max_dist));
  if (!(dist <= max_dist)) return false;
  *d = dist;
  return true;
}

// CpuExecutor::inside's per-point body (cpu.rs:182-188, bvh.rs:146-158): the AABB reject, three
// axis rays with a majority of odd counts, and the unbounded distance of an inside point.
__device__ __forceinline__ bool inside_depth(const SceneView &sv, V3 p, int retest, float *depth) {
  if (p.x < sv.lo[0] || p.x > sv.hi[0] || p.y < sv.lo[1] || p.y > sv.hi[1] || p.z < sv.lo[2] ||
      p.z > sv.hi[2])
    return false;
  // Majority of three odd counts: two agreeing rays decide it, and the third is cast only when
  // the first two disagree.
  int v0 = (int)(crossings_axis(sv, p, 0, retest) & 1u);
  int v1 = (int)(crossings_axis(sv, p, 1, retest) & 1u);
  int votes = v0 + v1;
  if (v0 != v1) votes += (int)(crossings_axis(sv, p, 2, retest) & 1u);
  if (votes < 2) return false;
  *depth = sqrtf(min_d2(sv, p, __builtin_inff()));
  return true;
}

}  // namespace sherd
