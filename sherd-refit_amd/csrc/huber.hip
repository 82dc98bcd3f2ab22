// The robust (Huber) point-to-plane seam energy over a batch of poses, its gradient in the se(3)
// twist, and a batched Levenberg-Marquardt refinement driven by it. New: the reference has no such
// energy; its nearest relative is ICP's point-to-plane normal equations
// (crates/sherd-core/src/matching/icp.rs:599-640), rows [p x n, n] and residuals r = (p - q) . n.
//
// For pose c (base transform T0, twist xi = (omega, v)) and its frozen correspondences i -> j(i):
//   p_i = exp(omega^) T0 s_i + v,  r_i = (p_i - q_j) . n_j,  E = sum rho(r_i)  (Huber, delta)
//   dE/dv = sum psi(r_i) n_j,  dE/domega = J_l(omega)^T sum psi(r_i) (exp(omega^) T0 s_i) x n_j
//
// Precision: the per-point arithmetic and the sums run in T (float or double). Coordinates enter
// shifted by the clouds' centroids (s - c_s, q - c_t, computed in f64 and then rounded to T), so the
// residual's cancellation happens at the scale of the clouds' extent rather than of their position;
// the shift is undone exactly in f64 on the sums (sum psi p x n = sum psi p' x n + (c_t - v) x sum
// psi n). The pose algebra, the left Jacobian and the LM system's 6x6 solve are f64.
//
// One workgroup per pose; sums are per thread in a fixed stride and then a fixed tree, so results
// are deterministic run to run.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cloud.hpp"
#include "kernels.hpp"
#include "linalg.hpp"

namespace sherd {
namespace {

constexpr int HT = 256;
constexpr int H_LSTACK = 8;  // a correspondence search's deferred nodes, in an LDS column per lane

__device__ __forceinline__ void rodrigues(const double *w, double *R) {
  double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double K[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
  double K2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
  double a, b;
  if (th < 1e-8) {
    a = 1.0;
    b = 0.5;
  } else {
    a = sin(th) / th;
    b = (1.0 - cos(th)) / (th * th);
  }
  for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0 ? 1.0 : 0.0) + a * K[k] + b * K2[k];
}

__device__ __forceinline__ void left_jacobian(const double *w, double *J) {
  double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double K[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
  double K2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
  double a, b;
  if (th < 1e-6) {
    a = 0.5;
    b = 1.0 / 6.0;
  } else {
    a = (1.0 - cos(th)) / (th * th);
    b = (th - sin(th)) / (th * th * th);
  }
  for (int k = 0; k < 9; ++k) J[k] = (k % 4 == 0 ? 1.0 : 0.0) + a * K[k] + b * K2[k];
}

template <class T, int N>
__device__ __forceinline__ void block_sum(T *v, T *lds, T *out) {
  int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    T x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) lds[wave * N + k] = x;
  }
  __syncthreads();
  if (t < N) out[t] = ((lds[t] + lds[N + t]) + lds[2 * N + t]) + lds[3 * N + t];
  __syncthreads();
}

template <class T>
__device__ __forceinline__ void huber(T r, T delta, T *rho, T *psi) {
  T a = r < T(0) ? -r : r;
  if (a <= delta) {
    *rho = T(0.5) * r * r;
    *psi = r;
  } else {
    *rho = delta * (a - T(0.5) * delta);
    *psi = r < T(0) ? -delta : delta;
  }
}

// The shifted affine map of a pose: p - c_t = A (s - c_s) + b, from rotation R and translation t.
__device__ __forceinline__ void shifted_map(const double *R, const double *t, const double *cs, const double *ct,
                                            double *A, double *b) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) A[3 * i + j] = R[3 * i + j];
    b[i] = ((R[3 * i] * cs[0] + R[3 * i + 1] * cs[1]) + R[3 * i + 2] * cs[2]) + t[i] - ct[i];
  }
}

// The per-point terms of pose c at the shifted map (A, b) in precision T; `emit` sees
// (r, n', p') of every corresponding point.
template <class T, class F>
__device__ __forceinline__ void for_pairs(const IcpPairDesc &pd, const int32_t *corr, const T *A, const T *b, F emit) {
  for (int i = threadIdx.x; i < pd.ns; i += HT) {
    int32_t j = corr[i];
    if (j < 0) continue;
    const double *s = pd.src + 3 * (size_t)i;
    const double *q = pd.tp + 3 * (size_t)j;
    const double *n = pd.tn + 3 * (size_t)j;
    T s0 = (T)(s[0] - pd.cs[0]), s1 = (T)(s[1] - pd.cs[1]), s2 = (T)(s[2] - pd.cs[2]);
    T p[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) p[k] = ((A[3 * k] * s0 + A[3 * k + 1] * s1) + A[3 * k + 2] * s2) + b[k];
    T qq[3] = {(T)(q[0] - pd.ct[0]), (T)(q[1] - pd.ct[1]), (T)(q[2] - pd.ct[2])};
    T nn[3] = {(T)n[0], (T)n[1], (T)n[2]};
    T r = ((p[0] - qq[0]) * nn[0] + (p[1] - qq[1]) * nn[1]) + (p[2] - qq[2]) * nn[2];
    emit(r, nn, p);
  }
}

// The residual of source point i against target j at the shifted map (A, b), as for_pairs forms it.
template <class T>
__device__ __forceinline__ T pair_residual(const IcpPairDesc &pd, int i, int32_t j, const T *A, const T *b) {
  const double *s = pd.src + 3 * (size_t)i;
  const double *q = pd.tp + 3 * (size_t)j;
  const double *n = pd.tn + 3 * (size_t)j;
  T s0 = (T)(s[0] - pd.cs[0]), s1 = (T)(s[1] - pd.cs[1]), s2 = (T)(s[2] - pd.cs[2]);
  T p[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = ((A[3 * k] * s0 + A[3 * k + 1] * s1) + A[3 * k + 2] * s2) + b[k];
  T qq[3] = {(T)(q[0] - pd.ct[0]), (T)(q[1] - pd.ct[1]), (T)(q[2] - pd.ct[2])};
  T nn[3] = {(T)n[0], (T)n[1], (T)n[2]};
  return ((p[0] - qq[0]) * nn[0] + (p[1] - qq[1]) * nn[1]) + (p[2] - qq[2]) * nn[2];
}

// The change of that residual under a change (dA, db) of the map: (dA (s - c_s) + db) . n.
template <class T>
__device__ __forceinline__ T residual_change(const IcpPairDesc &pd, int i, int32_t j, const T *dA, const T *db) {
  const double *s = pd.src + 3 * (size_t)i;
  const double *n = pd.tn + 3 * (size_t)j;
  T s0 = (T)(s[0] - pd.cs[0]), s1 = (T)(s[1] - pd.cs[1]), s2 = (T)(s[2] - pd.cs[2]);
  T dp[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) dp[k] = ((dA[3 * k] * s0 + dA[3 * k + 1] * s1) + dA[3 * k + 2] * s2) + db[k];
  return (dp[0] * (T)n[0] + dp[1] * (T)n[1]) + dp[2] * (T)n[2];
}

// rho(r + dr) and its change from rho(r): within one branch of the Huber function the change is
// formed from dr itself (dr (r + dr / 2) inside, +-delta dr outside), so it carries no cancellation.
template <class T>
__device__ __forceinline__ void huber_change(T r, T dr, T delta, T *rho1, T *d) {
  T r1 = r + dr, psi, rho0;
  huber(r1, delta, rho1, &psi);
  T a0 = r < T(0) ? -r : r, a1 = r1 < T(0) ? -r1 : r1;
  if (a0 <= delta && a1 <= delta) {
    *d = dr * (r + T(0.5) * dr);
  } else if (a0 > delta && a1 > delta && ((r < T(0)) == (r1 < T(0)))) {
    *d = (r < T(0) ? -delta : delta) * dr;
  } else {
    huber(r, delta, &rho0, &psi);
    *d = *rho1 - rho0;
  }
}

// Energy (and gradient) of every pose at its twist. One workgroup per pose.
template <class T>
__global__ void __launch_bounds__(HT) k_huber_eval(HuberView v, const double *xi, double delta, double *E, double *G) {
  int c = blockIdx.x;
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const int32_t *corr = v.corr + v.off[c];
  __shared__ T sA[9], sb[3];
  __shared__ T lds[4 * 7];
  __shared__ T red[7];
  if (threadIdx.x == 0) {
    const double *T0 = v.T0 + 16 * (size_t)c;
    const double *x = xi + 6 * (size_t)c;
    double Rw[9], R[9], t[3], A[9], b[3];
    rodrigues(x, Rw);
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j)
        R[3 * i + j] = (Rw[3 * i] * T0[j] + Rw[3 * i + 1] * T0[4 + j]) + Rw[3 * i + 2] * T0[8 + j];
      t[i] = ((Rw[3 * i] * T0[3] + Rw[3 * i + 1] * T0[7]) + Rw[3 * i + 2] * T0[11]) + x[3 + i];
    }
    shifted_map(R, t, pd.cs, pd.ct, A, b);
    for (int k = 0; k < 9; ++k) sA[k] = (T)A[k];
    for (int k = 0; k < 3; ++k) sb[k] = (T)b[k];
  }
  __syncthreads();
  T A[9], b[3];
  for (int k = 0; k < 9; ++k) A[k] = sA[k];
  for (int k = 0; k < 3; ++k) b[k] = sb[k];
  T dl = (T)delta;
  T acc[7] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  for_pairs<T>(pd, corr, A, b, [&](T r, const T *n, const T *p) {
    T rho, psi;
    huber(r, dl, &rho, &psi);
    acc[0] += rho;
    acc[1] += psi * (p[1] * n[2] - p[2] * n[1]);
    acc[2] += psi * (p[2] * n[0] - p[0] * n[2]);
    acc[3] += psi * (p[0] * n[1] - p[1] * n[0]);
    acc[4] += psi * n[0];
    acc[5] += psi * n[1];
    acc[6] += psi * n[2];
  });
  block_sum<T, 7>(acc, lds, red);
  if (threadIdx.x == 0) {
    E[c] = (double)red[0];
    if (G) {
      const double *x = xi + 6 * (size_t)c;
      double gv[3] = {(double)red[4], (double)red[5], (double)red[6]};
      // sum psi (R T0 s) x n = sum psi p' x n + (c_t - v) x sum psi n
      double d[3] = {pd.ct[0] - x[3], pd.ct[1] - x[4], pd.ct[2] - x[5]};
      double s1[3] = {(double)red[1] + (d[1] * gv[2] - d[2] * gv[1]), (double)red[2] + (d[2] * gv[0] - d[0] * gv[2]),
                      (double)red[3] + (d[0] * gv[1] - d[1] * gv[0])};
      double J[9];
      left_jacobian(x, J);
      double *g = G + 6 * (size_t)c;
      for (int k = 0; k < 3; ++k) g[k] = (J[k] * s1[0] + J[3 + k] * s1[1]) + J[6 + k] * s1[2];
      for (int k = 0; k < 3; ++k) g[3 + k] = gv[k];
    }
  }
}

// Correspondences of every pose at its transform (icp.rs:361-392): nearest target strictly within
// the pair's radius, -1 otherwise. One workgroup per pose.
__global__ void __launch_bounds__(HT) k_correspond(HuberView v, const double *Tc, int32_t *corr_all) {
  int c = blockIdx.x;
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const double *M = Tc + 16 * (size_t)c;
  int32_t *corr = corr_all + v.off[c];
  double r2 = pd.r * pd.r;
  __shared__ uint2 s_cstk[H_LSTACK][HT];
  LdsStack<HT, H_LSTACK> cst{&s_cstk[0][threadIdx.x]};
  for (int i = threadIdx.x; i < pd.ns; i += HT) {
    const double *s = pd.src + 3 * (size_t)i;
    double p0 = M[0] * s[0] + M[1] * s[1] + M[2] * s[2] + M[3];
    double p1 = M[4] * s[0] + M[5] * s[1] + M[6] * s[2] + M[7];
    double p2 = M[8] * s[0] + M[9] * s[1] + M[10] * s[2] + M[11];
    uint32_t j;
    double d2;
    corr[i] = (pd.r > 0.0 && cloud_nearest_with(cst, pd.tree, pd.tp, p0, p1, p2, r2, &j, &d2)) ? (int32_t)j : -1;
  }
}

// Batched Levenberg-Marquardt on the Huber energy. Each iteration takes the correspondences at the
// current pose, forms the iteratively reweighted normal equations (weights psi(r)/r) with rows
// [p' x n, n] about the target centroid, solves (H + lambda diag H) dx = -g in f64 (icp.rs's
// Eigen-pivoted LDLT), and keeps the step when the energy over the same correspondences drops,
// the drop summed point by point (lambda / 10), else raises lambda x 10 and stays.
//
// An iteration is two launches: k_lm_search finds every point's correspondence for every pose
// still running, one thread per point (each search starts from the point's previous partner, at
// the first iteration its partner at the base pose: the same answer, found sooner), and
// k_lm_step, one workgroup per pose, does the rest. A pose's state (LmState) waits in global
// memory between them; a pose that has stopped returns at once from both.

__global__ void k_lm_init(HuberView v) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= v.k) return;
  LmState &st = v.lm[c];
  for (int q = 0; q < 16; ++q) st.T[q] = v.T0[16 * (size_t)c + q];
  st.lambda = 1e-3;
  st.E = st.Ef = 0.0;
  st.done = 0;
  st.steps = 0;
}

__global__ void __launch_bounds__(HT) k_lm_search(HuberView v, int it, int c0) {
  const int c = c0 + (int)blockIdx.y;
  const LmState &st = v.lm[c];
  if (st.done) return;
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const int i = blockIdx.x * HT + threadIdx.x;
  if (i >= pd.ns) return;
  __shared__ uint2 s_cstk[H_LSTACK][HT];
  LdsStack<HT, H_LSTACK> cst{&s_cstk[0][threadIdx.x]};
  int32_t *corr = v.lm_corr + v.off[c];
  const int32_t *base = v.corr + v.off[c];
  const double *sT = st.T;
  const double *s = pd.src + 3 * (size_t)i;
  double p0 = sT[0] * s[0] + sT[1] * s[1] + sT[2] * s[2] + sT[3];
  double p1 = sT[4] * s[0] + sT[5] * s[1] + sT[6] * s[2] + sT[7];
  double p2 = sT[8] * s[0] + sT[9] * s[1] + sT[10] * s[2] + sT[11];
  uint32_t j;
  double d2;
  const int32_t prev = it > 0 ? corr[i] : base[i];
  const uint32_t seed = prev >= 0 ? (uint32_t)prev : NONE;
  corr[i] =
      (pd.r > 0.0 && cloud_nearest_with(cst, pd.tree, pd.tp, p0, p1, p2, pd.r * pd.r, &j, &d2, seed)) ? (int32_t)j : -1;
}

template <class T>
__global__ void __launch_bounds__(HT) k_lm_step(HuberView v, double delta, int it) {
  const int c = blockIdx.x;
  LmState &st = v.lm[c];
  if (st.done) return;  // (uniform: one pose per workgroup)
  const IcpPairDesc pd = v.pairs[v.cand_pair[c]];
  const int32_t *corr = v.lm_corr + v.off[c];
  __shared__ double sT[16], sTn[16];
  __shared__ T sA[9], sb[3];
  __shared__ T lds[4 * 28];
  __shared__ T red[28];
  __shared__ double s_lambda, s_E, s_Ef;
  __shared__ int s_done, s_steps;
  // Thread 0's 6x6 system and its solve work in LDS: as private arrays their pivot-indexed entries
  // went to scratch.
  __shared__ double s_H[36], s_g[6], s_x[6], s_m[36], s_y[6], s_tmp[6];
  __shared__ int s_perm[6];
  if (threadIdx.x < 16) sT[threadIdx.x] = st.T[threadIdx.x];
  if (threadIdx.x == 0) {
    s_lambda = st.lambda;
    s_done = 0;
    s_steps = st.steps;
    s_E = st.E;
    s_Ef = st.Ef;
    double R[9] = {sT[0], sT[1], sT[2], sT[4], sT[5], sT[6], sT[8], sT[9], sT[10]};
    double t[3] = {sT[3], sT[7], sT[11]}, A[9], b[3];
    shifted_map(R, t, pd.cs, pd.ct, A, b);
    for (int k = 0; k < 9; ++k) sA[k] = (T)A[k];
    for (int k = 0; k < 3; ++k) sb[k] = (T)b[k];
  }
  __syncthreads();
  T dl = (T)delta;
  T A[9], b[3];
  for (int k = 0; k < 9; ++k) A[k] = sA[k];
  for (int k = 0; k < 3; ++k) b[k] = sb[k];
  T acc[28];
#pragma unroll
  for (int k = 0; k < 28; ++k) acc[k] = T(0);
  for_pairs<T>(pd, corr, A, b, [&](T r, const T *n, const T *p) {
    T rho, psi;
    huber(r, dl, &rho, &psi);
    T w = r == T(0) ? T(1) : psi / r;
    T row[6] = {p[1] * n[2] - p[2] * n[1], p[2] * n[0] - p[0] * n[2], p[0] * n[1] - p[1] * n[0], n[0], n[1], n[2]};
    acc[0] += rho;
    int q = 1;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
      for (int l = 0; l <= k; ++l) acc[q++] += w * row[k] * row[l];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[22 + k] += w * r * row[k];
  });
  block_sum<T, 28>(acc, lds, red);
  if (threadIdx.x == 0) {
    s_E = (double)red[0];
    s_Ef = s_E;
    double *H = s_H, *g = s_g, *x = s_x;
    int q = 1;
    for (int k = 0; k < 6; ++k)
      for (int l = 0; l <= k; ++l) {
        H[6 * k + l] = H[6 * l + k] = (double)red[q];
        ++q;
      }
    for (int k = 0; k < 6; ++k) g[k] = -(double)red[22 + k];
    // Marquardt's damping on the diagonal, with a floor of 1e-9 of the mean diagonal so that
    // directions the seam does not constrain (a breakline's tangent) stay damped.
    double tr = 0.0;
    for (int k = 0; k < 6; ++k) tr += H[7 * k];
    for (int k = 0; k < 6; ++k) H[7 * k] += s_lambda * (fmax(H[7 * k], 0.0) + 1e-9 * tr / 6.0 + 1e-300);
    bool ok = s_E > 0.0 && solve_ldlt_ws(H, g, x, s_m, s_y, s_tmp, s_perm);
    if (ok) {
      // A trust region: at most 0.05 rad of rotation and the correspondence radius of
      // translation per step.
      double wn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]), vn = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
      double f = fmin(1.0, fmin(0.05 / fmax(wn, 1e-300), pd.r / fmax(vn, 1e-300)));
      for (int k = 0; k < 6; ++k) x[k] *= f;
    }
    if (!ok) {
      s_done = 1;
    } else {
      // The step as a rotation about the target centroid and a translation.
      double Rw[9];
      rodrigues(x, Rw);
      double tw[3];
      for (int i = 0; i < 3; ++i)
        tw[i] = pd.ct[i] - ((Rw[3 * i] * pd.ct[0] + Rw[3 * i + 1] * pd.ct[1]) + Rw[3 * i + 2] * pd.ct[2]) + x[3 + i];
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 4; ++j)
          sTn[4 * i + j] = (Rw[3 * i] * sT[j] + Rw[3 * i + 1] * sT[4 + j]) + Rw[3 * i + 2] * sT[8 + j];
        sTn[4 * i + 3] += tw[i];
      }
      sTn[12] = sTn[13] = sTn[14] = 0.0;
      sTn[15] = 1.0;
      // The step as a change of the shifted map, (A2 - A, b2 - b), taken in f64 and rounded once,
      // so each residual's change is computed from the small difference itself.
      double R0[9] = {sT[0], sT[1], sT[2], sT[4], sT[5], sT[6], sT[8], sT[9], sT[10]};
      double t0[3] = {sT[3], sT[7], sT[11]}, A0[9], b0[3];
      shifted_map(R0, t0, pd.cs, pd.ct, A0, b0);
      double R[9] = {sTn[0], sTn[1], sTn[2], sTn[4], sTn[5], sTn[6], sTn[8], sTn[9], sTn[10]};
      double t[3] = {sTn[3], sTn[7], sTn[11]}, A2[9], b2[3];
      shifted_map(R, t, pd.cs, pd.ct, A2, b2);
      for (int k = 0; k < 9; ++k) sA[k] = (T)(A2[k] - A0[k]);
      for (int k = 0; k < 3; ++k) sb[k] = (T)(b2[k] - b0[k]);
    }
  }
  __syncthreads();
  if (!s_done) {
    // The trial pose's energy over the same correspondences, and its change summed point by point
    // (sum of rho(r') - rho(r), not the difference of two sums: the change stays resolved in f32
    // long after the two totals agree to their last bit). r' = r + dr, with dr = (dA s' + db) . n
    // from the map's change.
    T dA[9], db[3];
    for (int k = 0; k < 9; ++k) dA[k] = sA[k];
    for (int k = 0; k < 3; ++k) db[k] = sb[k];
    T e[2] = {T(0), T(0)};
    for (int i = threadIdx.x; i < pd.ns; i += HT) {
      int32_t j = corr[i];
      if (j < 0) continue;
      T r0 = pair_residual<T>(pd, i, j, A, b), dr = residual_change<T>(pd, i, j, dA, db);
      T rho1, d;
      huber_change(r0, dr, dl, &rho1, &d);
      e[0] += rho1;
      e[1] += d;
    }
    block_sum<T, 2>(e, lds, red);
    if (threadIdx.x == 0) {
      double en = (double)red[0], de = (double)red[1];
      if (de < 0.0) {
        for (int k = 0; k < 16; ++k) sT[k] = sTn[k];
        s_Ef = en;
        s_lambda = fmax(s_lambda * 0.1, 1e-12);
        s_steps = it + 1;
        if (-de <= 1e-12 * s_E) s_done = 1;
      } else {
        s_lambda *= 10.0;
        if (s_lambda > 1e12) s_done = 1;
      }
    }
  }
  if (threadIdx.x == 0) {
    for (int k = 0; k < 16; ++k) st.T[k] = sT[k];
    st.lambda = s_lambda;
    st.E = s_E;
    st.Ef = s_Ef;
    st.steps = s_steps;
    st.done = s_done;
  }
}

__global__ void k_lm_final(HuberView v, double *T_out, double *E_out, uint32_t *steps_out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= v.k) return;
  const LmState &st = v.lm[c];
  for (int q = 0; q < 16; ++q) T_out[16 * (size_t)c + q] = st.T[q];
  E_out[c] = st.Ef;
  steps_out[c] = (uint32_t)st.steps;
}

}  // namespace

hipError_t launch_correspond(const HuberView &v, const double *Tc, int32_t *corr, hipStream_t s) {
  if (v.k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_correspond, dim3(v.k), dim3(HT), 0, s, v, Tc, corr);
  return hipGetLastError();
}

hipError_t launch_huber_eval(const HuberView &v, const double *xi, double delta, int f64, double *E, double *G,
                             hipStream_t s) {
  if (v.k <= 0) return hipSuccess;
  if (f64)
    hipLaunchKernelGGL(k_huber_eval<double>, dim3(v.k), dim3(HT), 0, s, v, xi, delta, E, G);
  else
    hipLaunchKernelGGL(k_huber_eval<float>, dim3(v.k), dim3(HT), 0, s, v, xi, delta, E, G);
  return hipGetLastError();
}

hipError_t launch_huber_lm(const HuberView &v, double delta, int iters, int f64, double *T_out, double *E_out,
                           uint32_t *steps, hipStream_t s) {
  if (v.k <= 0) return hipSuccess;
  const unsigned kb = (unsigned)((v.k + 63) / 64);
  hipLaunchKernelGGL(k_lm_init, dim3(kb), dim3(64), 0, s, v);
  const unsigned spans = (unsigned)((v.max_ns + HT - 1) / HT);
  for (int it = 0; it < iters; ++it) {
    // (poses in groups of at most 65535, the grid's y extent)
    for (int c0 = 0; spans > 0 && c0 < v.k; c0 += 65535)
      hipLaunchKernelGGL(k_lm_search, dim3(spans, (unsigned)std::min(65535, v.k - c0)), dim3(HT), 0, s, v, it, c0);
    if (f64)
      hipLaunchKernelGGL(k_lm_step<double>, dim3(v.k), dim3(HT), 0, s, v, delta, it);
    else
      hipLaunchKernelGGL(k_lm_step<float>, dim3(v.k), dim3(HT), 0, s, v, delta, it);
  }
  hipLaunchKernelGGL(k_lm_final, dim3(kb), dim3(64), 0, s, v, T_out, E_out, steps);
  return hipGetLastError();
}

}  // namespace sherd
