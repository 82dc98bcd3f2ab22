// Nearest-neighbour queries on an ICP cloud's point LBVH (icp.hip, huber.hip).
#pragma once

#include "bvh_query.hpp"

namespace sherd {

constexpr uint32_t NONE = 0xFFFFFFFFu;

// kiddo's squared distance from q to point j of the cloud.
__device__ __forceinline__ double point_d2(const double *pts, uint32_t j, double q0, double q1, double q2) {
  const double *p = pts + 3 * (size_t)j;
  return d2_f64(q0, q1, q2, gldd(p), gldd(p + 1), gldd(p + 2));
}

// A point tree's leaf record at `slot`: after the build its first 24 bytes hold the point's f64
// coordinates and `face` its index (k_point_records), all fetched at once.
__device__ __forceinline__ double leaf_d2(const Tri *tris, int slot, double q0, double q1, double q2, uint32_t *j) {
  const Tri *r = tris + slot;
  const double *d = reinterpret_cast<const double *>(r);
  double x = gldd(d), y = gldd(d + 1), z = gldd(d + 2);
  *j = gldu(&r->face);
  return d2_f64(q0, q1, q2, x, y, z);
}

// The nearest point of a cloud to q with d^2 < r2, ties to the lowest index. The cloud's tree is
// an LBVH over its points (each leaf a point, `face` its index); boxes are f32 and grown by the
// slack, which exceeds the rounding of the f32 boxes and of q, so no box hides a point nearer than
// the best so far. Distances are kiddo's f64 ((dx^2 + dy^2) + dz^2) on the f64 points.
//
// `seed` (a point index, or NONE) starts the search from that point's distance when it is inside
// the radius: the answer is the same (every point at most as far is still examined, ties
// included), the walk shorter when the seed is close, as last iteration's partner usually is.
template <class Stack>
__device__ inline bool cloud_nearest_with(Stack &st, const SceneView &sv, const double *pts, double q0, double q1,
                                          double q2, double r2, uint32_t *out_j, double *out_d2, uint32_t seed = NONE) {
  int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  if (cur < 0 || cur >= nodes || !(r2 > 0.0)) return false;
  V3 qf = V3{(float)q0, (float)q1, (float)q2};
  float slack = slack_of(sv.absmax, qf);
  double best = r2;
  uint32_t bj = NONE;
  if (seed != NONE && seed < (uint32_t)sv.n) {
    double d2 = point_d2(pts, seed, q0, q1, q2);
    if (d2 < best) {
      best = d2;
      bj = seed;
    }
  }
  float lim = (float)sqrt(best) + slack;
  float lim2 = lim * lim;
  Node nd = load_node(sv.nodes, cur);
  if (box_d2(nd, qf, slack) > lim2) {
    if (bj == NONE) return false;
    cur = -1;  // (cannot happen with a seed: the root's box holds it) keep the seed
  }
  int sp = 0;
  bool overflow = false;
  for (int iter = 0; cur >= 0 && iter <= nodes; ++iter) {
    int next = -1;
    if (nd.a < 0) {
      uint32_t j;
      double d2 = leaf_d2(sv.tris, ~nd.a, q0, q1, q2, &j);
      if (d2 < best || (d2 == best && j < bj)) {
        best = d2;
        bj = j;
        lim = (float)sqrt(best) + slack;
        lim2 = lim * lim;
      }
    } else if (nd.a > cur && nd.a < nodes) {
      // Both children at once: the left is the next node, the right the parent's `a`.
      int l = cur + 1, r = nd.a;
      Node ln = load_node(sv.nodes, l), rn = load_node(sv.nodes, r);
      float dl = box_d2(ln, qf, slack), dr = box_d2(rn, qf, slack);
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
    // The deferred nodes that did not fit: an exhaustive pre-order walk against the same bound.
    int steps = 0;
    for (int i = sv.root; i >= 0 && i < nodes && steps++ < nodes;) {
      Node nd = load_node(sv.nodes, i);
      if (nd.a >= 0) {
        i = box_d2(nd, qf, slack) <= lim2 ? i + 1 : nd.esc;
      } else {
        uint32_t j;
        double d2 = leaf_d2(sv.tris, ~nd.a, q0, q1, q2, &j);
        if (d2 < best || (d2 == best && j < bj)) {
          best = d2;
          bj = j;
          lim = (float)sqrt(best) + slack;
          lim2 = lim * lim;
        }
        i = nd.esc;
      }
    }
  }
  if (bj == NONE || !(best < r2)) return false;
  *out_j = bj;
  *out_d2 = best;
  return true;
}

__device__ inline bool cloud_nearest(const SceneView &sv, const double *pts, double q0, double q1, double q2,
                                     double r2, uint32_t *out_j, double *out_d2, uint32_t seed = NONE) {
  RayRegStack st;
  return cloud_nearest_with(st, sv, pts, q0, q1, q2, r2, out_j, out_d2, seed);
}


}  // namespace sherd
