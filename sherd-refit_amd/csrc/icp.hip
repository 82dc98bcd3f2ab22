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

template <int B>
__global__ void __launch_bounds__(B) k_icp_batch(IcpBatchView v) {
  constexpr int W = B / 64;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *part = v.part + (size_t)blockIdx.x * v.part_spans * NV;
  __shared__ double sT[16], sU[16];
  __shared__ double red[NV];
  __shared__ double s_fit, s_rmse;
  __shared__ int s_stop, s_iters, s_conv, s_c;
  for (;;) {
    if (t == 0) {
      uint32_t k = atomicAdd(v.next, 1u);
      s_c = k < (uint32_t)v.n_cand ? (int)v.order[k] : -1;
    }
    __syncthreads();
    const int c = s_c;
    if (c < 0) break;
    const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
    const int ns = pd.ns;
    const int nspan = (ns + ICP_SPAN - 1) / ICP_SPAN;
    double *moved = v.moved + 3 * v.off[c];
    uint32_t *corr = v.corr + v.off[c];
    const bool radius_ok = pd.r > 0.0;
    const double r2 = pd.r * pd.r;
    const bool plane = v.est == 1 && pd.tn != nullptr;
    const bool p2p = v.est == 0;  // otherwise every update is the identity (no normals)
    if (t < 16) {
      sT[t] = v.T_in[16 * (size_t)c + t];
      sU[t] = sT[t];
    }
    if (t == 0) {
      s_fit = 0.0;
      s_rmse = 0.0;
      s_stop = 0;
      s_iters = 0;
      s_conv = 0;
    }
    __syncthreads();
    for (int it = 0;; ++it) {
      double M[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) M[k] = sU[k];
      for (int sp = wave; sp < nspan; sp += W) {
        double acc[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] = 0.0;
        for (int i = sp * ICP_SPAN + lane; i < ns && i < (sp + 1) * ICP_SPAN; i += 64) {
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
          if (radius_ok && cloud_nearest(pd.tree, pd.tp, p0, p1, p2, r2, &j, &d2, seed)) {
            acc[0] += 1.0;
            acc[1] += d2;
            const double *q = pd.tp + 3 * (size_t)j;
            if (plane) {
              const double *n = pd.tn + 3 * (size_t)j;
              double row[6] = {p1 * n[2] - p2 * n[1], p2 * n[0] - p0 * n[2], p0 * n[1] - p1 * n[0], n[0], n[1], n[2]};
              double r = (p0 - q[0]) * n[0] + (p1 - q[1]) * n[1] + (p2 - q[2]) * n[2];
              int w = 2;
#pragma unroll
              for (int k = 0; k < 6; ++k) {
#pragma unroll
                for (int l = 0; l <= k; ++l) acc[w++] += row[k] * row[l];
              }
#pragma unroll
              for (int k = 0; k < 6; ++k) acc[23 + k] += row[k] * r;
            } else if (p2p) {
              acc[2] += p0;
              acc[3] += p1;
              acc[4] += p2;
              acc[5] += q[0];
              acc[6] += q[1];
              acc[7] += q[2];
            }
          } else {
            j = NONE;
          }
          corr[i] = j;
        }
        wave_sum<NV>(acc);
        if (lane == 0)
#pragma unroll
          for (int k = 0; k < NV; ++k) part[(size_t)sp * NV + k] = acc[k];
      }
      __syncthreads();
      if (wave == 0) spans_total<NV>(part, nspan, red);
      __syncthreads();
      if (t == 0) {
        double cnt = red[0];
        double fit = cnt > 0.0 ? cnt / (double)ns : 0.0;
        double rmse = cnt > 0.0 ? sqrt(red[1] / cnt) : 0.0;
        int conv = 0;
        if (it > 0) conv = (fabs(fit - s_fit) < 1e-6 && fabs(rmse - s_rmse) < 1e-6) ? 1 : 0;
        s_fit = fit;
        s_rmse = rmse;
        s_conv = conv;
        s_iters = it;
        s_stop = (conv || it >= v.max_iter) ? 1 : 0;
        if (!s_stop && plane) {
          // icp.rs:454-483: solve JTJ x = -JTr, Euler update, identity on an unsolvable system.
          double U[16];
          eye4(U);
          if (cnt > 0.0) {
            double a[36], rhs[6], x[6];
            int w = 2;
            for (int k = 0; k < 6; ++k)
              for (int l = 0; l <= k; ++l) {
                a[6 * k + l] = red[w];
                a[6 * l + k] = red[w];
                ++w;
              }
            for (int k = 0; k < 6; ++k) rhs[k] = -red[23 + k];
            if (solve_ldlt(a, rhs, x)) {
              double rot[9];
              euler_zyx(x[0], x[1], x[2], rot);
              for (int r = 0; r < 3; ++r) {
                for (int q = 0; q < 3; ++q) U[4 * r + q] = rot[3 * r + q];
                U[4 * r + 3] = x[3 + r];
              }
            }
          }
          for (int k = 0; k < 16; ++k) sU[k] = U[k];
        } else if (!s_stop && !p2p) {
          eye4(sU);
        }
      }
      __syncthreads();
      if (s_stop) break;
      if (p2p) {
        // Point-to-point: the covariance of the pairs about their means (icp.rs:557-575).
        double cnt = red[0];
        double one_over_n = cnt > 0.0 ? 1.0 / cnt : 0.0;
        double mp0 = red[2] * one_over_n, mp1 = red[3] * one_over_n, mp2 = red[4] * one_over_n;
        double mq0 = red[5] * one_over_n, mq1 = red[6] * one_over_n, mq2 = red[7] * one_over_n;
        if (cnt > 0.0) {
          for (int sp = wave; sp < nspan; sp += W) {
            double sg[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) sg[k] = 0.0;
            for (int i = sp * ICP_SPAN + lane; i < ns && i < (sp + 1) * ICP_SPAN; i += 64) {
              uint32_t j = corr[i];
              if (j == NONE) continue;
              const double *p = moved + 3 * (size_t)i;
              const double *q = pd.tp + 3 * (size_t)j;
              double a0 = p[0] - mp0, a1 = p[1] - mp1, a2 = p[2] - mp2;
              double b0 = (q[0] - mq0) * one_over_n, b1 = (q[1] - mq1) * one_over_n, b2 = (q[2] - mq2) * one_over_n;
              sg[0] += b0 * a0; sg[1] += b0 * a1; sg[2] += b0 * a2;
              sg[3] += b1 * a0; sg[4] += b1 * a1; sg[5] += b1 * a2;
              sg[6] += b2 * a0; sg[7] += b2 * a1; sg[8] += b2 * a2;
            }
            wave_sum<9>(sg);
            if (lane == 0)
#pragma unroll
              for (int k = 0; k < 9; ++k) part[(size_t)sp * NV + k] = sg[k];
          }
          __syncthreads();
          if (wave == 0) spans_total<9>(part, nspan, red + 2);
          __syncthreads();
        }
        if (t == 0) {
          double U[16];
          eye4(U);
          if (cnt > 0.0) {
            double mp[3] = {mp0, mp1, mp2}, mq[3] = {mq0, mq1, mq2};
            umeyama_pose(red + 2, mp, mq, U);
          }
          for (int k = 0; k < 16; ++k) sU[k] = U[k];
        }
        __syncthreads();
      }
      if (t == 0) {
        double nt[16];
        mul_nn(sU, sT, nt, 4);
        for (int k = 0; k < 16; ++k) sT[k] = nt[k];
      }
      __syncthreads();
    }
    if (t == 0) {
      double *o = v.out + 21 * (size_t)c;
      for (int k = 0; k < 16; ++k) o[k] = sT[k];
      o[16] = s_fit;
      o[17] = s_rmse;
      o[18] = red[0];
      o[19] = (double)s_iters;
      o[20] = (double)s_conv;
      if (v.T_out)
        for (int k = 0; k < 16; ++k) v.T_out[16 * (size_t)c + k] = sT[k];
    }
    __syncthreads();
  }
}

__global__ void k_icp_reset(uint32_t *next) { *next = 0; }

}  // namespace

int icp_block_sizes(int *out) {
  static const int sizes[] = {64, 128, 256, 512};
  for (int k = 0; k < 4; ++k) out[k] = sizes[k];
  return 4;
}

hipError_t launch_icp_batch(const IcpBatchView &v, int block, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_icp_reset, dim3(1), dim3(1), 0, s, v.next);
  switch (block) {
    case 64:
      hipLaunchKernelGGL(k_icp_batch<64>, dim3(grid), dim3(64), 0, s, v);
      break;
    case 128:
      hipLaunchKernelGGL(k_icp_batch<128>, dim3(grid), dim3(128), 0, s, v);
      break;
    case 512:
      hipLaunchKernelGGL(k_icp_batch<512>, dim3(grid), dim3(512), 0, s, v);
      break;
    default:
      hipLaunchKernelGGL(k_icp_batch<256>, dim3(grid), dim3(256), 0, s, v);
      break;
  }
  return hipGetLastError();
}

}  // namespace sherd
