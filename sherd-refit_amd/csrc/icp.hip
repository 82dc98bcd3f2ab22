// R §7's ICP for many candidates over many cloud pairs at once: CpuExecutor::icp_rung
// (crates/sherd-core/src/executor/cpu.rs:70-82) -> icp::register (crates/sherd-core/src/matching/
// icp.rs:308-356), Open3D's loop:
//
//   moved = T0 . S;  (fit, rmse, C) = corr(moved)
//   repeat up to max_iteration: U = update(C); T = U.T; moved = U.moved; (fit', rmse', C) = corr(moved)
//                               stop when |fit' - fit| < 1e-6 and |rmse' - rmse| < 1e-6
//
// A rung is one persistent launch: workgroups take candidates from a queue, longest source cloud
// first, and keep a candidate's whole rung on the device. Each correspondence pass applies the last
// update to the stored moved points, finds the nearest target point strictly inside the radius
// (f64, through the target's point LBVH: kiddo's nearest_within_squared, kdtree.rs:202-218, ties to
// the lowest index), and accumulates what the next update needs in the same pass: the 21 + 6
// normal-equation sums for point-to-plane (icp.rs:599-640), the two point sums for point-to-point,
// whose covariance (icp.rs:539-592) takes a second pass once the means are known.
//
// The sums have one fixed order that depends only on the cloud size: the source is cut into spans
// of SPAN points, lane l of the wave that owns a span adds its points l, l + 64, ... in turn, the
// 64 lanes meet in a fixed shuffle tree, and the span totals are added by a fixed tree in span
// order. So a result is the same for any workgroup size and any assignment of candidates to
// workgroups; it differs from the CPU's sequential source-order sums in rounding only.
#include <hip/hip_runtime.h>

#include "bvh_query.hpp"
#include "cloud.hpp"
#include "kernels.hpp"
#include "linalg.hpp"

namespace sherd {
namespace {

constexpr int NV = ICP_NV;
constexpr int ICP_LSTACK = 8;  // the nearest-point search's deferred nodes, in an LDS column per lane

// Sums v[0..n) over the wave's lanes in a fixed shuffle tree; lane 0 holds the totals.
template <int N>
__device__ __forceinline__ void wave_sum(double *v) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    v[k] = x;
  }
}

// Adds the first nspan span partials (N values each, stride NV) in the fixed order: lane l adds
// spans l, l + 64, ... in turn, then the lanes' tree. Called by the workgroup's first wave; lane 0
// gets the totals in out.
template <int N>
__device__ __forceinline__ void spans_total(const double *part, int nspan, double *out) {
  int lane = threadIdx.x & 63;
  double v[N];
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = 0.0;
  for (int sp = lane; sp < nspan; sp += 64)
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += part[(size_t)sp * NV + k];
  wave_sum<N>(v);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = v[k];
}

// The kernels' sums by estimation method: EST_P2P for point-to-point, EST_PLANE for point-to-plane
// and for identity updates (est 2, or point-to-plane without target normals), which read only the
// count and err^2.
constexpr int EST_P2P = 0, EST_PLANE = 1;

// How many per-pass sums a method keeps: count, err^2, then 27 normal-equation sums for
// point-to-plane or the two point sums of point-to-point.
template <int EST>
constexpr int icp_sums() {
  return EST == EST_PLANE ? NV : 8;
}

// What the next update needs from one corresponding pair (moved source point p, target point j):
// the count and squared distance, then the point-to-plane row products (icp.rs:599-640) or the two
// point sums of point-to-point (icp.rs:557-575). Without normals (tn null) the plane products are
// of a zero normal; such an update is the identity and never reads them.
template <int EST>
__device__ __forceinline__ void icp_accumulate(double *acc, double p0, double p1, double p2, double d2,
                                               const double *tp, const double *tn, uint32_t j) {
  acc[0] += 1.0;
  acc[1] += d2;
  const double *q = tp + 3 * (size_t)j;
  if (EST == EST_PLANE) {
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    if (tn) {
      n0 = gldd(tn + 3 * (size_t)j);
      n1 = gldd(tn + 3 * (size_t)j + 1);
      n2 = gldd(tn + 3 * (size_t)j + 2);
    }
    double row[6] = {p1 * n2 - p2 * n1, p2 * n0 - p0 * n2, p0 * n1 - p1 * n0, n0, n1, n2};
    double r = (p0 - gldd(q)) * n0 + (p1 - gldd(q + 1)) * n1 + (p2 - gldd(q + 2)) * n2;
    int w = 2;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
      for (int l = 0; l <= k; ++l) acc[w++] += row[k] * row[l];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[23 + k] += row[k] * r;
  } else {
    acc[2] += p0;
    acc[3] += p1;
    acc[4] += p2;
    acc[5] += gldd(q);
    acc[6] += gldd(q + 1);
    acc[7] += gldd(q + 2);
  }
}

// One iteration's bookkeeping from a correspondence pass's totals `red` (icp.rs:329-352): fitness
// and rmse, the convergence test against the previous pass, the stop decision and, for
// point-to-plane, the update (icp.rs:454-483: JTJ x = -JTr, Euler ZYX, identity on an unsolvable
// system). Point-to-point leaves U to the covariance pass; an update without normals is the
// identity.
struct IcpIter {
  double fit, rmse;
  int stop, iters, conv;
};

// ws: ICP_WS doubles and 6 ints of working space (the caller's LDS): W 16, the system 36 + 6 + 6,
// the rotation 9, then the solve's 48 (or the Euler product's 36).
constexpr int ICP_WS = 121;
__device__ __forceinline__ void icp_iteration_ws(const double *red, int ns, int it, int max_iter, bool plane, bool p2p,
                                                 IcpIter *s, double *U, double *ws, int *iws) {
  double cnt = red[0];
  double fit = cnt > 0.0 ? cnt / (double)ns : 0.0;
  double rmse = cnt > 0.0 ? sqrt(red[1] / cnt) : 0.0;
  int conv = 0;
  if (it > 0) conv = (fabs(fit - s->fit) < 1e-6 && fabs(rmse - s->rmse) < 1e-6) ? 1 : 0;
  s->fit = fit;
  s->rmse = rmse;
  s->conv = conv;
  s->iters = it;
  s->stop = (conv || it >= max_iter) ? 1 : 0;
  if (!s->stop && plane) {
    double *W = ws, *a = ws + 16, *rhs = ws + 52, *x = ws + 58, *rot = ws + 64;  // (ws + 73: the solve's)
    eye4(W);
    if (cnt > 0.0) {
      int w = 2;
      for (int k = 0; k < 6; ++k)
        for (int l = 0; l <= k; ++l) {
          a[6 * k + l] = red[w];
          a[6 * l + k] = red[w];
          ++w;
        }
      for (int k = 0; k < 6; ++k) rhs[k] = -red[23 + k];
      if (solve_ldlt_ws(a, rhs, x, ws + 73, ws + 109, ws + 115, iws)) {
        euler_zyx_ws(x[0], x[1], x[2], rot, ws + 73);
        for (int r = 0; r < 3; ++r) {
          for (int q = 0; q < 3; ++q) W[4 * r + q] = rot[3 * r + q];
          W[4 * r + 3] = x[3 + r];
        }
      }
    }
    for (int k = 0; k < 16; ++k) U[k] = W[k];
  } else if (!s->stop && !p2p) {
    eye4(U);
  }
}

// Point-to-point's means of the pairs (icp.rs:557-560), from a pass's totals.
struct P2pMeans {
  double inv, mp0, mp1, mp2, mq0, mq1, mq2;
};
__device__ __forceinline__ P2pMeans p2p_means(const double *red) {
  double cnt = red[0];
  double one_over_n = cnt > 0.0 ? 1.0 / cnt : 0.0;
  return P2pMeans{one_over_n,        red[2] * one_over_n, red[3] * one_over_n, red[4] * one_over_n,
                  red[5] * one_over_n, red[6] * one_over_n, red[7] * one_over_n};
}

// One span's covariance terms about the means (icp.rs:561-575), lane l adding points l, l + 64, ...
// in turn; after the wave's tree lane 0 holds the span's nine sums.
__device__ __forceinline__ void p2p_cov_span(const double *moved, const uint32_t *corr, const double *tp, int ns,
                                             int sp, const P2pMeans &m, double *sg) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 9; ++k) sg[k] = 0.0;
  for (int i = sp * ICP_SPAN + lane; i < ns && i < (sp + 1) * ICP_SPAN; i += 64) {
    uint32_t j = corr[i];
    if (j == NONE) continue;
    const double *p = moved + 3 * (size_t)i;
    const double *q = tp + 3 * (size_t)j;
    double a0 = p[0] - m.mp0, a1 = p[1] - m.mp1, a2 = p[2] - m.mp2;
    double b0 = (q[0] - m.mq0) * m.inv, b1 = (q[1] - m.mq1) * m.inv, b2 = (q[2] - m.mq2) * m.inv;
    sg[0] += b0 * a0; sg[1] += b0 * a1; sg[2] += b0 * a2;
    sg[3] += b1 * a0; sg[4] += b1 * a1; sg[5] += b1 * a2;
    sg[6] += b2 * a0; sg[7] += b2 * a1; sg[8] += b2 * a2;
  }
  wave_sum<9>(sg);
}

__device__ __forceinline__ void write_result(const IcpBatchView &v, int c, const double *T, const IcpIter &s,
                                             double count) {
  double *o = v.out + 21 * (size_t)c;
  for (int k = 0; k < 16; ++k) o[k] = T[k];
  o[16] = s.fit;
  o[17] = s.rmse;
  o[18] = count;
  o[19] = (double)s.iters;
  o[20] = (double)s.conv;
  if (v.T_out)
    for (int k = 0; k < 16; ++k) v.T_out[16 * (size_t)c + k] = T[k];
}

template <int B, int EST>
__global__ void __launch_bounds__(B) k_icp_batch(IcpBatchView v) {
  constexpr int W = B / 64;
  constexpr int NS = icp_sums<EST>();
  constexpr bool p2p = EST == EST_P2P;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *part = v.part + (size_t)blockIdx.x * v.part_spans * NV;
  __shared__ double sT[16], sU[16];
  __shared__ double red[NV];
  __shared__ IcpIter s_it;
  __shared__ int s_c;
  // Lane 0's solves work in LDS: as private arrays their pivot-indexed entries went to scratch
  // and their values held registers the whole kernel is allocated by.
  __shared__ double s_ws[ICP_WS > 78 ? ICP_WS : 78];
  __shared__ int s_iws[6];
  __shared__ uint2 s_cstk[ICP_LSTACK][B];
  LdsStack<B, ICP_LSTACK> cst{&s_cstk[0][threadIdx.x]};
  for (;;) {
    if (t == 0) {
      uint32_t k = atomicAdd(v.next, 1u);
      s_c = k < (uint32_t)v.n_cand ? (int)v.order[k] : -1;
    }
    __syncthreads();
    const int c = __builtin_amdgcn_readfirstlane(s_c);
    if (c < 0) break;
    const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
    const int ns = pd.ns;
    const int nspan = (ns + ICP_SPAN - 1) / ICP_SPAN;
    double *moved = v.moved + 3 * v.off[c];
    uint32_t *corr = v.corr + v.off[c];
    const bool radius_ok = pd.r > 0.0;
    const double r2 = pd.r * pd.r;
    const bool plane = v.est == 1 && pd.tn != nullptr;  // otherwise every update is the identity
    if (t < 16) {
      sT[t] = v.T_in[16 * (size_t)c + t];
      sU[t] = sT[t];
    }
    if (t == 0) s_it = IcpIter{0.0, 0.0, 0, 0, 0};
    __syncthreads();
    for (int it = 0;; ++it) {
      const double *M = sU;  // read from LDS where used: no registers held across the searches
      for (int sp = wave; sp < nspan; sp += W) {
        const int end = ns < (sp + 1) * ICP_SPAN ? ns : (sp + 1) * ICP_SPAN;
        // The span's correspondences first, each lane its points l, l + 64, ... Nothing of the sums
        // is live during the searches, which keeps the wave's registers to the tree walk's.
        for (int i = sp * ICP_SPAN + lane; i < end; i += 64) {
          const double *s = it == 0 ? pd.src + 3 * (size_t)i : moved + 3 * (size_t)i;
          double s0 = s[0], s1 = s[1], s2 = s[2];
          double p0 = M[0] * s0 + M[1] * s1 + M[2] * s2 + M[3];
          double p1 = M[4] * s0 + M[5] * s1 + M[6] * s2 + M[7];
          double p2 = M[8] * s0 + M[9] * s1 + M[10] * s2 + M[11];
          moved[3 * (size_t)i] = p0;
          moved[3 * (size_t)i + 1] = p1;
          moved[3 * (size_t)i + 2] = p2;
          uint32_t j = NONE;
          double d2;
          uint32_t seed = it > 0 ? corr[i] : NONE;  // last iteration's partner (this candidate's own)
          if (!radius_ok || !cloud_nearest_with(cst, pd.tree, pd.tp, p0, p1, p2, r2, &j, &d2, seed)) j = NONE;
          corr[i] = j;
        }
        // Then the sums over the same points in the same order, from what the lane just stored (the
        // distance is the search's own arithmetic on the same values).
        double acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = 0.0;
        for (int i = sp * ICP_SPAN + lane; i < end; i += 64) {
          uint32_t j = corr[i];
          if (j == NONE) continue;
          double p0 = moved[3 * (size_t)i], p1 = moved[3 * (size_t)i + 1], p2 = moved[3 * (size_t)i + 2];
          icp_accumulate<EST>(acc, p0, p1, p2, point_d2(pd.tp, j, p0, p1, p2), pd.tp, pd.tn, j);
        }
        wave_sum<NS>(acc);
        if (lane == 0)
#pragma unroll
          for (int k = 0; k < NS; ++k) part[(size_t)sp * NV + k] = acc[k];
      }
      __syncthreads();
      if (wave == 0) spans_total<NS>(part, nspan, red);
      __syncthreads();
      if (t == 0) icp_iteration_ws(red, ns, it, v.max_iter, plane, p2p, &s_it, sU, s_ws, s_iws);
      __syncthreads();
      if (s_it.stop) break;
      if (p2p) {
        // Point-to-point: the covariance of the pairs about their means (icp.rs:557-575).
        const P2pMeans m = p2p_means(red);
        const double cnt = red[0];
        if (cnt > 0.0) {
          for (int sp = wave; sp < nspan; sp += W) {
            double sg[9];
            p2p_cov_span(moved, corr, pd.tp, ns, sp, m, sg);
            if (lane == 0)
#pragma unroll
              for (int k = 0; k < 9; ++k) part[(size_t)sp * NV + k] = sg[k];
          }
          __syncthreads();
          if (wave == 0) spans_total<9>(part, nspan, red + 2);
          __syncthreads();
        }
        if (t == 0) {
          eye4(sU);
          if (cnt > 0.0) {
            double mp[3] = {m.mp0, m.mp1, m.mp2}, mq[3] = {m.mq0, m.mq1, m.mq2};
            umeyama_pose_ws(red + 2, mp, mq, sU, s_ws, s_iws);
          }
        }
        __syncthreads();
      }
      if (t == 0) {
        mul_nn(sU, sT, s_ws, 4);
        for (int k = 0; k < 16; ++k) sT[k] = s_ws[k];
      }
      __syncthreads();
    }
    if (t == 0) write_result(v, c, sT, s_it, red[0]);
    __syncthreads();
  }
}

__global__ void k_icp_reset(uint32_t *next) { *next = 0; }

// ---- the wide form: one candidate over many workgroups ----------------------------------------
// For a few candidates on large clouds (R §9's refinement registers one join at a time, up to
// 150 000 points) one workgroup per candidate leaves the device idle. The wide form runs each pass
// as a grid over every span of every candidate, one thread per source point, and the iteration's
// bookkeeping as a second launch; each launch ends at once for a candidate that has stopped, so the
// host queues max_iteration + 1 passes and the device takes only the ones it needs.
//
// The sums are the batched form's to the bit: the points of a span are found in parallel, then the
// span's first wave adds them lane by lane in the batched order (l, l + 64, ...), and the spans
// meet in the same tree, so a candidate's registration does not depend on which form ran it.

// The queue position whose spans hold span s (span_first is ascending).
__device__ __forceinline__ int span_owner(const IcpWideView &w, int s) {
  int lo = 0, hi = w.b.n_cand - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if ((int)w.span_first[mid] <= s)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void k_icp_wide_init(IcpWideView w) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w.b.n_cand) return;
  IcpWideState &st = w.st[w.b.order[k]];
  const double *T = w.b.T_in + 16 * (size_t)w.b.order[k];
  for (int q = 0; q < 16; ++q) {
    st.T[q] = T[q];
    st.U[q] = T[q];
  }
  st.fit = st.rmse = 0.0;
  st.stop = st.iters = st.conv = 0;
}

template <int EST>
__global__ void __launch_bounds__(ICP_SPAN) k_icp_wide_corr(IcpWideView w, int it) {
  constexpr int NS = icp_sums<EST>();
  __shared__ double s_p[3 * ICP_SPAN];
  __shared__ double s_d2[ICP_SPAN];
  __shared__ uint32_t s_j[ICP_SPAN];
  __shared__ uint2 s_cstk[ICP_LSTACK][ICP_SPAN];
  const int s = w.s0 + (int)blockIdx.x;
  const int k = span_owner(w, s);
  const uint32_t c = w.b.order[k];
  const IcpWideState &st = w.st[c];
  if (st.stop) return;
  const IcpBatchView &v = w.b;
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const int ns = pd.ns;
  const int sp = s - (int)w.span_first[k];
  double *moved = v.moved + 3 * v.off[c];
  uint32_t *corr = v.corr + v.off[c];
  const int t = threadIdx.x;
  const int i = sp * ICP_SPAN + t;
  if (i < ns) {
    double M[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) M[k] = st.U[k];
    const double *src = it == 0 ? pd.src + 3 * (size_t)i : moved + 3 * (size_t)i;
    double s0 = src[0], s1 = src[1], s2 = src[2];
    double p0 = M[0] * s0 + M[1] * s1 + M[2] * s2 + M[3];
    double p1 = M[4] * s0 + M[5] * s1 + M[6] * s2 + M[7];
    double p2 = M[8] * s0 + M[9] * s1 + M[10] * s2 + M[11];
    moved[3 * (size_t)i] = p0;
    moved[3 * (size_t)i + 1] = p1;
    moved[3 * (size_t)i + 2] = p2;
    uint32_t j = NONE;
    double d2 = 0.0;
    uint32_t seed = it > 0 ? corr[i] : NONE;
    LdsStack<ICP_SPAN, ICP_LSTACK> cst{&s_cstk[0][threadIdx.x]};
    if (!(pd.r > 0.0) || !cloud_nearest_with(cst, pd.tree, pd.tp, p0, p1, p2, pd.r * pd.r, &j, &d2, seed)) j = NONE;
    corr[i] = j;
    s_p[3 * t] = p0;
    s_p[3 * t + 1] = p1;
    s_p[3 * t + 2] = p2;
    s_d2[t] = d2;
    s_j[t] = j;
  }
  __syncthreads();
  if (t >= 64) return;
  double acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = 0.0;
  for (int u = t; u < ICP_SPAN && sp * ICP_SPAN + u < ns; u += 64) {
    uint32_t j = s_j[u];
    if (j == NONE) continue;
    icp_accumulate<EST>(acc, s_p[3 * u], s_p[3 * u + 1], s_p[3 * u + 2], s_d2[u], pd.tp, pd.tn, j);
  }
  wave_sum<NS>(acc);
  if (t == 0)
#pragma unroll
    for (int k = 0; k < NS; ++k) w.part[(size_t)s * NV + k] = acc[k];
}

// One wave per candidate: the pass's totals, the iteration's bookkeeping, and for point-to-plane
// (or no normals) the new pose.
__global__ void __launch_bounds__(64) k_icp_wide_step(IcpWideView w, int it) {
  __shared__ double red[NV];
  __shared__ double s_ws[ICP_WS + 32];
  __shared__ int s_iws[6];
  const int kq = w.k0 + (int)blockIdx.x;
  const uint32_t c = w.b.order[kq];
  IcpWideState &st = w.st[c];
  if (st.stop) return;
  const IcpBatchView &v = w.b;
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const int nspan = (pd.ns + ICP_SPAN - 1) / ICP_SPAN;
  spans_total<NV>(w.part + (size_t)w.span_first[kq] * NV, nspan, red);
  if (threadIdx.x != 0) return;
  const bool plane = v.est == 1 && pd.tn != nullptr;
  const bool p2p = v.est == 0;
  IcpIter s{st.fit, st.rmse, st.stop, st.iters, st.conv};
  double *U = s_ws + ICP_WS, *nt = s_ws + ICP_WS + 16;
  for (int k = 0; k < 16; ++k) U[k] = st.U[k];
  icp_iteration_ws(red, pd.ns, it, v.max_iter, plane, p2p, &s, U, s_ws, s_iws);
  st.fit = s.fit;
  st.rmse = s.rmse;
  st.iters = s.iters;
  st.conv = s.conv;
  for (int k = 0; k < NV; ++k) st.red[k] = red[k];
  if (s.stop) {
    write_result(v, (int)c, st.T, s, red[0]);
    st.stop = 1;
    return;
  }
  if (!p2p) {
    mul_nn(U, st.T, nt, 4);
    for (int k = 0; k < 16; ++k) {
      st.U[k] = U[k];
      st.T[k] = nt[k];
    }
  }
}

// Point-to-point's iteration after a correspondence pass, one workgroup per candidate: the
// bookkeeping of k_icp_wide_step, the covariance pass and the Umeyama update in one launch, with
// the batched form's arithmetic in its order. Wave 0 takes the pass's totals and does the bookkeeping; unless the candidate
// stopped, each wave takes spans w, w + UPD_WAVES, ... of the covariance (one wave per span, the
// batched form's loop), and wave 0 adds the span sums and makes the Umeyama update.
constexpr int UPD_WAVES = 4;
__global__ void __launch_bounds__(64 * UPD_WAVES) k_icp_wide_p2p(IcpWideView w, int it) {
  __shared__ double red[NV];
  __shared__ double sig[9];
  __shared__ double s_ws[ICP_WS + 32];
  __shared__ int s_iws[6];
  __shared__ int s_go;
  const int kq = w.k0 + (int)blockIdx.x;
  const uint32_t c = w.b.order[kq];
  IcpWideState &st = w.st[c];
  if (st.stop) return;  // (uniform: one candidate per workgroup)
  const IcpBatchView &v = w.b;
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const int nspan = (pd.ns + ICP_SPAN - 1) / ICP_SPAN;
  double *part = w.part + (size_t)w.span_first[kq] * NV;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave == 0) {
    spans_total<NV>(part, nspan, red);
    if (lane == 0) {
      IcpIter s{st.fit, st.rmse, st.stop, st.iters, st.conv};
      double *U = s_ws + ICP_WS;
      for (int k = 0; k < 16; ++k) U[k] = st.U[k];
      icp_iteration_ws(red, pd.ns, it, v.max_iter, false, true, &s, U, s_ws, s_iws);
      st.fit = s.fit;
      st.rmse = s.rmse;
      st.iters = s.iters;
      st.conv = s.conv;
      for (int k = 0; k < NV; ++k) st.red[k] = red[k];
      if (s.stop) {
        write_result(v, (int)c, st.T, s, red[0]);
        st.stop = 1;
      }
      s_go = s.stop ? 0 : (red[0] > 0.0 ? 2 : 1);  // 0 stopped, 1 identity update, 2 Umeyama
    }
  }
  __syncthreads();
  const int go = s_go;
  if (go == 0) return;
  if (go == 2) {
    const P2pMeans m = p2p_means(red);
    for (int sp = wave; sp < nspan; sp += UPD_WAVES) {
      double sg[9];
      p2p_cov_span(v.moved + 3 * v.off[c], v.corr + v.off[c], pd.tp, pd.ns, sp, m, sg);
      if (lane == 0)
#pragma unroll
        for (int k = 0; k < 9; ++k) part[(size_t)sp * NV + k] = sg[k];
    }
  }
  __syncthreads();
  if (wave != 0) return;
  if (go == 2) spans_total<9>(part, nspan, sig);
  if (lane != 0) return;
  double *U = s_ws + 78, *nt = s_ws + 94;
  if (go == 2) {
    const P2pMeans m = p2p_means(red);
    double mp[3] = {m.mp0, m.mp1, m.mp2}, mq[3] = {m.mq0, m.mq1, m.mq2}, ur[16];
    umeyama_pose_reg(sig, mp, mq, ur);
    for (int k = 0; k < 16; ++k) U[k] = ur[k];
  } else {
    eye4(U);
  }
  mul_nn(U, st.T, nt, 4);
  for (int k = 0; k < 16; ++k) {
    st.U[k] = U[k];
    st.T[k] = nt[k];
  }
}

}  // namespace

template <int EST>
void launch_batch_sized(const IcpBatchView &v, int block, int grid, hipStream_t s) {
  switch (block) {
    case 64:
      hipLaunchKernelGGL((k_icp_batch<64, EST>), dim3(grid), dim3(64), 0, s, v);
      break;
    case 128:
      hipLaunchKernelGGL((k_icp_batch<128, EST>), dim3(grid), dim3(128), 0, s, v);
      break;
    case 512:
      hipLaunchKernelGGL((k_icp_batch<512, EST>), dim3(grid), dim3(512), 0, s, v);
      break;
    default:
      hipLaunchKernelGGL((k_icp_batch<256, EST>), dim3(grid), dim3(256), 0, s, v);
      break;
  }
}

int icp_block_sizes(int *out) {
  static const int sizes[] = {64, 128, 256, 512};
  for (int k = 0; k < 4; ++k) out[k] = sizes[k];
  return 4;
}

hipError_t launch_icp_batch(const IcpBatchView &v, int block, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_icp_reset, dim3(1), dim3(1), 0, s, v.next);
  if (v.est == EST_P2P)
    launch_batch_sized<EST_P2P>(v, block, grid, s);
  else
    launch_batch_sized<EST_PLANE>(v, block, grid, s);
  return hipGetLastError();
}

namespace {
// One chain of a rung's passes over queue positions [k0, k1), spans [s0, s1).
void wide_chain(IcpWideView w, int k0, int k1, int s0, int s1, hipStream_t s, KernelProfiler *prof) {
  w.k0 = k0;
  w.s0 = s0;
  const unsigned nk = (unsigned)(k1 - k0), ns = (unsigned)(s1 - s0);
  if (nk == 0) return;
  for (int it = 0; it <= w.b.max_iter; ++it) {
    if (prof) prof->begin("k_icp_wide_corr", s);
    if (w.b.est == EST_P2P)
      hipLaunchKernelGGL(k_icp_wide_corr<EST_P2P>, dim3(ns), dim3(ICP_SPAN), 0, s, w, it);
    else
      hipLaunchKernelGGL(k_icp_wide_corr<EST_PLANE>, dim3(ns), dim3(ICP_SPAN), 0, s, w, it);
    if (prof) prof->end(s);
    if (w.b.est == EST_P2P && it < w.b.max_iter)
      hipLaunchKernelGGL(k_icp_wide_p2p, dim3(nk), dim3(64 * UPD_WAVES), 0, s, w, it);
    else
      hipLaunchKernelGGL(k_icp_wide_step, dim3(nk), dim3(64), 0, s, w, it);
  }
}
}  // namespace

hipError_t launch_icp_wide(const IcpWideView &w, hipStream_t s, KernelProfiler *prof, const IcpWideSplit &split,
                           const hipStream_t *side, hipEvent_t fork, const hipEvent_t *join) {
  const int n = w.b.n_cand;
  if (n == 0) return hipSuccess;
  IcpWideView w0 = w;
  w0.k0 = w0.s0 = 0;
  hipLaunchKernelGGL(k_icp_wide_init, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, w0);
  if (!side || split.n <= 1) {
    wide_chain(w, 0, n, 0, w.n_spans, s, prof);
    return hipGetLastError();
  }
  hipError_t e = hipEventRecord(fork, s);
  for (int j = 1; j < split.n && e == hipSuccess; ++j) e = hipStreamWaitEvent(side[j - 1], fork, 0);
  if (e != hipSuccess) return e;
  for (int j = 0; j < split.n; ++j)
    wide_chain(w, split.k[j], split.k[j + 1], split.s[j], split.s[j + 1], j == 0 ? s : side[j - 1], prof);
  for (int j = 1; j < split.n && e == hipSuccess; ++j) e = hipEventRecord(join[j - 1], side[j - 1]);
  for (int j = 1; j < split.n && e == hipSuccess; ++j) e = hipStreamWaitEvent(s, join[j - 1], 0);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace sherd
