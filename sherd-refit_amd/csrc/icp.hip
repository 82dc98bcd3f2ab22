// One rung of R §7's ICP for a batch of candidates: CpuExecutor::icp_rung
// (crates/sherd-core/src/executor/cpu.rs:70-82) -> icp::register (crates/sherd-core/src/matching/
// icp.rs:308-356), Open3D's loop:
//
//   moved = T0 . S;  (fit, rmse, C) = corr(moved)
//   repeat up to max_iteration: U = update(C); T = U.T; moved = U.moved; (fit', rmse', C) = corr(moved)
//                               stop when |fit' - fit| < 1e-6 and |rmse' - rmse| < 1e-6
//
// One 256-thread workgroup per candidate keeps the whole rung on the device. Each correspondence
// pass applies the last update to the stored moved points, finds the nearest target strictly inside
// the radius (f64, the grid standing in for the KD-tree), and accumulates what the next update
// needs in the same pass: the 21 + 6 normal-equation sums for point-to-plane (icp.rs:599-640), the
// two point sums for point-to-point, whose covariance (icp.rs:539-592) takes a second pass once the
// means are known. Sums run per thread in a fixed stride and reduce in a fixed tree, so a rung is
// deterministic run to run; it differs from the CPU's sequential source-order sums in rounding only.
#include <hip/hip_runtime.h>

#include "grid.hpp"
#include "kernels.hpp"
#include "linalg.hpp"

namespace sherd {
namespace {

constexpr int IT = 256;
constexpr int NV = 29;  // count, err^2, then 27 update sums
constexpr uint32_t NONE = 0xFFFFFFFFu;

// Sums NV values over the workgroup in a fixed order: wave trees, then the four waves in order.
__device__ __forceinline__ void block_sum(double *v, double *lds, double *out) {
  int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) lds[wave * NV + k] = x;
  }
  __syncthreads();
  if (t < NV) out[t] = ((lds[t] + lds[NV + t]) + lds[2 * NV + t]) + lds[3 * NV + t];
  __syncthreads();
}

__global__ void __launch_bounds__(IT) k_icp(GridView g, const double *tp, const double *tn, const double *src,
                                            int ns, const double *inits, int est, double distance, int max_iter,
                                            double *moved_all, uint32_t *corr_all, double *out) {
  const int c = blockIdx.x, t = threadIdx.x;
  double *moved = moved_all + (size_t)c * ns * 3;
  uint32_t *corr = corr_all + (size_t)c * ns;
  __shared__ double sT[16], sU[16];
  __shared__ double lds[4 * NV];
  __shared__ double red[NV];
  __shared__ double s_fit, s_rmse;
  __shared__ int s_stop, s_iters, s_conv;
  const bool radius_ok = distance > 0.0;
  const double r2 = distance * distance;
  const bool plane = est == 1 && tn != nullptr;
  const bool p2p = est == 0;  // otherwise every update is the identity (no normals)
  if (t < 16) {
    sT[t] = inits[16 * (size_t)c + t];
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
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    double M[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) M[k] = sU[k];
    for (int i = t; i < ns; i += IT) {
      const double *s = it == 0 ? src + 3 * (size_t)i : moved + 3 * (size_t)i;
      double s0 = s[0], s1 = s[1], s2 = s[2];
      double p0 = M[0] * s0 + M[1] * s1 + M[2] * s2 + M[3];
      double p1 = M[4] * s0 + M[5] * s1 + M[6] * s2 + M[7];
      double p2 = M[8] * s0 + M[9] * s1 + M[10] * s2 + M[11];
      moved[3 * (size_t)i] = p0;
      moved[3 * (size_t)i + 1] = p1;
      moved[3 * (size_t)i + 2] = p2;
      uint32_t j = NONE;
      double d2;
      if (radius_ok && grid_nearest(g, p0, p1, p2, &j, &d2) && d2 < r2) {
        acc[0] += 1.0;
        acc[1] += d2;
        const double *q = tp + 3 * (size_t)j;
        if (plane) {
          const double *n = tn + 3 * (size_t)j;
          double pc0 = p0 - 0.0, pc1 = p1 - 0.0, pc2 = p2 - 0.0;
          double row[6] = {pc1 * n[2] - pc2 * n[1], pc2 * n[0] - pc0 * n[2], pc0 * n[1] - pc1 * n[0],
                           n[0], n[1], n[2]};
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
    block_sum(acc, lds, red);
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
      s_stop = (conv || it >= max_iter) ? 1 : 0;
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
      double sg[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) sg[k] = 0.0;
      if (cnt > 0.0) {
        for (int i = t; i < ns; i += IT) {
          uint32_t j = corr[i];
          if (j == NONE) continue;
          const double *p = moved + 3 * (size_t)i;
          const double *q = tp + 3 * (size_t)j;
          double a0 = p[0] - mp0, a1 = p[1] - mp1, a2 = p[2] - mp2;
          double b0 = (q[0] - mq0) * one_over_n, b1 = (q[1] - mq1) * one_over_n, b2 = (q[2] - mq2) * one_over_n;
          sg[0] += b0 * a0; sg[1] += b0 * a1; sg[2] += b0 * a2;
          sg[3] += b1 * a0; sg[4] += b1 * a1; sg[5] += b1 * a2;
          sg[6] += b2 * a0; sg[7] += b2 * a1; sg[8] += b2 * a2;
        }
      }
      __syncthreads();
      block_sum(sg, lds, red);
      if (t == 0) {
        double U[16];
        eye4(U);
        if (cnt > 0.0) {
          double mp[3] = {mp0, mp1, mp2}, mq[3] = {mq0, mq1, mq2};
          umeyama_pose(red, mp, mq, U);
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
    double *o = out + 21 * (size_t)c;
    for (int k = 0; k < 16; ++k) o[k] = sT[k];
    o[16] = s_fit;
    o[17] = s_rmse;
    o[18] = red[0];
    o[19] = (double)s_iters;
    o[20] = (double)s_conv;
  }
}

}  // namespace

hipError_t launch_icp_rung(const GridView &g, const double *tp, const double *tn, const double *src, int ns,
                           const double *inits, int k, int estimation, double distance, int max_iter,
                           double *moved, uint32_t *corr, double *out, hipStream_t s) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_icp, dim3(k), dim3(IT), 0, s, g, tp, tn, src, ns, inits, estimation, distance, max_iter,
                     moved, corr, out);
  return hipGetLastError();
}

}  // namespace sherd
