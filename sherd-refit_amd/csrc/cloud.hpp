// Nearest-neighbour queries on an ICP cloud's point LBVH (icp.hip, huber.hip).
#pragma once

#include "bvh_query.hpp"

namespace sherd {

constexpr uint32_t NONE = 0xFFFFFFFFu;

// The nearest point of a cloud to q with d^2 < r2, ties to the lowest index. The cloud's tree is
// an LBVH over its points (each leaf a point, `face` its index); boxes are f32 and grown by the
// slack, which exceeds the rounding of the f32 boxes and of q, so no box hides a point nearer than
// the best so far. Distances are kiddo's f64 ((dx^2 + dy^2) + dz^2) on the f64 points.
//
// `seed` (a point index, or NONE) starts the search from that point's distance when it is inside
// the radius: the answer is the same (every point at most as far is still examined, ties
// included), the walk shorter when the seed is close, as last iteration's partner usually is.
__device__ inline bool cloud_nearest(const SceneView &sv, const double *pts, double q0, double q1, double q2,
                                     double r2, uint32_t *out_j, double *out_d2, uint32_t seed = NONE) {
  int nodes = 2 * sv.n - 1;
  int cur = sv.root;
  if (cur < 0 || cur >= nodes || !(r2 > 0.0)) return false;
  V3 qf = V3{(float)q0, (float)q1, (float)q2};
  float slack = slack_of(sv.absmax, qf);
  double best = r2;
  uint32_t bj = NONE;
  if (seed != NONE && seed < (uint32_t)sv.n) {
    double d2 = d2_f64(q0, q1, q2, pts[3 * (size_t)seed], pts[3 * (size_t)seed + 1], pts[3 * (size_t)seed + 2]);
    if (d2 < best) {
      best = d2;
      bj = seed;
    }
  }
  float lim = (float)sqrt(best) + slack;
  float lim2 = lim * lim;
  if (box_d2(load_node(sv.nodes, cur), qf, slack) > lim2) {
    if (bj == NONE) return false;
    cur = -1;  // (cannot happen with a seed: the root's box holds it) keep the seed
  }
  int st_n[DSTACK];
  float st_d[DSTACK];
  int sp = 0;
  bool overflow = false;
  for (int iter = 0; cur >= 0 && iter <= nodes; ++iter) {
    Node nd = load_node(sv.nodes, cur);
    int next = -1;
    if (nd.a < 0) {
      V3 a, b, c;
      uint32_t j;
      load_tri(sv.tris, ~nd.a, &a, &b, &c, &j);
      double d2 = d2_f64(q0, q1, q2, pts[3 * (size_t)j], pts[3 * (size_t)j + 1], pts[3 * (size_t)j + 2]);
      if (d2 < best || (d2 == best && j < bj)) {
        best = d2;
        bj = j;
        lim = (float)sqrt(best) + slack;
        lim2 = lim * lim;
      }
    } else {
      int l = nd.a;
      Node ln = load_node(sv.nodes, l);
      int r = ln.esc;
      float dl = box_d2(ln, qf, slack);
      float dr = __builtin_inff();
      if (r >= 0 && r < nodes) dr = box_d2(load_node(sv.nodes, r), qf, slack);
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
    // The deferred nodes that did not fit: an exhaustive pre-order walk against the same bound.
    int steps = 0;
    for (int i = sv.root; i >= 0 && i < nodes && steps++ < nodes;) {
      Node nd = load_node(sv.nodes, i);
      if (nd.a >= 0) {
        i = box_d2(nd, qf, slack) <= lim2 ? nd.a : nd.esc;
      } else {
        V3 a, b, c;
        uint32_t j;
        load_tri(sv.tris, ~nd.a, &a, &b, &c, &j);
        double d2 = d2_f64(q0, q1, q2, pts[3 * (size_t)j], pts[3 * (size_t)j + 1], pts[3 * (size_t)j + 2]);
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


}  // namespace sherd
