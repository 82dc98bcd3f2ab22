// R §5.1-5.5 for many pairs at once on the device (crates/sherd-core/src/matching/pair.rs:170-264):
// the hypotheses of paired breakline frames, their coarse scores, the walk order and the greedy
// suppression; after the breakline ladder (icp.hip), the stage-1 re-score and the second
// suppression.
//
//   k_hyp_count / k_hyp_fill   hypotheses.rs:149-165 build_with: per row i of the A subset, the
//                              B-subset positions j with |dih_A + dih_B - 180| < tol, row-major
//                              (np.where's order), placed by an exclusive scan over all rows
//   k_hyp_pose                 hypotheses.rs:187-212 pose: R = RA.RB^T with B's tangent and
//                              in-plane axis negated, tau = P_A - R.P_B, sums in j = 0, 1, 2 order
//   k_stage_coarse             coarse.rs score_of / cpu.rs:136-169 for every item (a hypothesis or a
//                              stage-1 pose), one wave each, against its own pair's target grid
//   k_order_keys               nms.rs:50-58 order_by_score: descending score, ties by ascending
//                              index. Scores are count / n with one n per pair, so the order of the
//                              counts is the order of the scores; one 64-bit key (pair, count,
//                              index) sorted once orders every pair's items at once
//   k_nms                      nms.rs:62-84: the greedy walk, 64 candidates at a time by one wave:
//                              each lane tests its candidate against the kept poses and the earlier
//                              lanes' candidates, then the lanes are decided in walk order from
//                              those masks, which is the sequential loop's decision for each
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "sort.hpp"

namespace sherd {
namespace {

constexpr int ST = 256;
unsigned blocks(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// The pair owning global row / item r, from an ascending first-row table of n_pairs + 1 entries.
__device__ __forceinline__ int owner(const uint32_t *first, int n_pairs, uint32_t r) {
  int lo = 0, hi = n_pairs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= r)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(ST) k_hyp_count(const StagePair *sp, const uint32_t *row_first, int n_pairs,
                                                  uint32_t rows, uint32_t *cnt) {
  uint32_t r = blockIdx.x * ST + threadIdx.x;
  if (r > rows) return;
  if (r == rows) {
    cnt[r] = 0;  // the scan's total
    return;
  }
  int pi = owner(row_first, n_pairs, r);
  const StagePair &p = sp[pi];
  uint32_t i = r - row_first[pi];
  double da = p.a.dih[p.ia[i]];
  uint32_t c = 0;
  for (int j = 0; j < p.nib; ++j)
    if (fabs(da + p.b.dih[p.ib[j]] - 180.0) < p.dihedral_tol) ++c;
  cnt[r] = c;
}

__global__ void __launch_bounds__(ST) k_hyp_fill(const StagePair *sp, const uint32_t *row_first, int n_pairs,
                                                 uint32_t rows, const uint32_t *off, uint32_t *pa, uint32_t *pb,
                                                 uint32_t *hpair) {
  uint32_t r = blockIdx.x * ST + threadIdx.x;
  if (r >= rows) return;
  int pi = owner(row_first, n_pairs, r);
  const StagePair &p = sp[pi];
  uint32_t i = r - row_first[pi];
  double da = p.a.dih[p.ia[i]];
  uint32_t at = off[r];
  for (int j = 0; j < p.nib; ++j)
    if (fabs(da + p.b.dih[p.ib[j]] - 180.0) < p.dihedral_tol) {
      pa[at] = i;
      pb[at] = (uint32_t)j;
      hpair[at] = (uint32_t)pi;
      ++at;
    }
}

__global__ void __launch_bounds__(ST) k_hyp_pose(const StagePair *sp, const uint32_t *pa, const uint32_t *pb,
                                                 const uint32_t *hpair, uint32_t n, double *pose) {
  uint32_t h = blockIdx.x * ST + threadIdx.x;
  if (h >= n) return;
  const StagePair &p = sp[hpair[h]];
  uint32_t ka = p.ia[pa[h]], kb = p.ib[pb[h]];
  const double *A[3] = {p.a.t + 3 * (size_t)ka, p.a.ns + 3 * (size_t)ka, p.a.f + 3 * (size_t)ka};
  const double *B[3] = {p.b.t + 3 * (size_t)kb, p.b.ns + 3 * (size_t)kb, p.b.f + 3 * (size_t)kb};
  double *o = pose + 12 * (size_t)h;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) o[3 * i + k] = A[0][i] * -B[0][k] + A[1][i] * B[1][k] + A[2][i] * -B[2][k];
  const double *pA = p.a.p + 3 * (size_t)ka, *pB = p.b.p + 3 * (size_t)kb;
  for (int i = 0; i < 3; ++i) o[9 + i] = pA[i] - (o[3 * i] * pB[0] + o[3 * i + 1] * pB[1] + o[3 * i + 2] * pB[2]);
}

// which = 0: the coarse probe at the coarse radius; 1: B's subset at the stage-1 radius.
__global__ void __launch_bounds__(ST) k_stage_coarse(const StagePair *sp, const uint32_t *ipair, const double *pose,
                                                     uint32_t n, int which, uint32_t *counts, double *scores) {
  uint32_t h = (blockIdx.x * ST + threadIdx.x) >> 6;
  if (h >= n) return;
  const StagePair &p = sp[ipair[h]];
  const double *pp = which ? p.b.sub_p : p.probe_p;
  const double *pn = which ? p.b.sub_n : p.probe_n;
  int np = which ? p.b.n_sub : p.n_probe;
  double radius = which ? p.stage1_r : p.coarse_r;
  uint32_t c = 0;
  if (np > 0 && p.a.n > 0 && radius > 0.0)
    c = coarse_agree_wave(p.a.grid, p.a.ns, pp, pn, np, pose + 12 * (size_t)h, radius, 0.7);
  if ((threadIdx.x & 63) == 0) {
    counts[h] = c;
    scores[h] = np > 0 ? (double)c / (double)np : 0.0;  // an empty probe scores zero (cpu.rs:47-55)
  }
}

__global__ void __launch_bounds__(ST) k_order_keys(const uint32_t *ipair, const uint32_t *item_first,
                                                   const uint32_t *counts, uint32_t n, uint64_t *keys,
                                                   uint32_t *vals) {
  uint32_t h = blockIdx.x * ST + threadIdx.x;
  if (h >= n) return;
  uint32_t p = ipair[h];
  keys[h] = ((uint64_t)p << 48) | ((uint64_t)(0xFFFFu - counts[h]) << 32) | (uint64_t)(h - item_first[p]);
  vals[h] = h;
}

// trace(a^T b) as nms.rs:97-104 sums it: per column i, (a0i b0i + a1i b1i) + a2i b2i, added in turn.
__device__ __forceinline__ bool same_candidate(const double *x, const double *y, double tol) {
  double d0 = x[9] - y[9], d1 = x[10] - y[10], d2 = x[11] - y[11];
  if (!(sqrt(d0 * d0 + d1 * d1 + d2 * d2) < tol)) return false;
  double tr = 0.0;
  for (int i = 0; i < 3; ++i) tr += x[i] * y[i] + x[3 + i] * y[3 + i] + x[6 + i] * y[6 + i];
  return tr > 2.9;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int l) {
  uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, l), hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// One wave per problem.
__global__ void __launch_bounds__(64) k_nms(const NmsProblem *probs) {
  const NmsProblem &q = probs[blockIdx.x];
  const int lane = threadIdx.x;
  uint32_t K = 0;
  bool stop = q.n == 0;
  for (uint32_t s = 0; !stop && s < (uint32_t)q.n; s += 64) {
    const uint32_t me = s + lane;
    const bool valid = me < (uint32_t)q.n;
    uint32_t h = valid ? q.order[me] : 0u;
    const double *mine = q.pose + 12 * (size_t)h;
    bool below = valid && q.score[h] < q.floor;
    bool dup_kept = false;
    if (valid)
      for (uint32_t k = 0; k < K && !dup_kept; ++k) dup_kept = same_candidate(mine, q.pose + 12 * (size_t)q.kept[k], q.tol);
    uint64_t dup_before = 0;
    const uint32_t nb = (uint32_t)q.n - s < 64u ? (uint32_t)q.n - s : 64u;
    for (uint32_t m = 0; m < nb; ++m) {
      uint32_t hm = q.order[s + m];
      if (valid && m < (uint32_t)lane && same_candidate(mine, q.pose + 12 * (size_t)hm, q.tol))
        dup_before |= 1ull << m;
    }
    // The walk, in order: stop at the first score under the floor, keep what is a duplicate of
    // nothing kept, stop once topk are kept (nms.rs:68-82).
    uint64_t kept_mask = 0;
    uint32_t K0 = K;
    for (uint32_t l = 0; l < nb; ++l) {
      if (__shfl((int)below, (int)l)) {
        stop = true;
        break;
      }
      bool dk = __shfl((int)dup_kept, (int)l);
      uint64_t db = shfl64(dup_before, (int)l);
      if (!dk && !(db & kept_mask)) {
        kept_mask |= 1ull << l;
        ++K;
      }
      if (K >= (uint32_t)q.topk) {
        stop = true;
        break;
      }
    }
    if ((kept_mask >> lane) & 1ull) q.kept[K0 + (uint32_t)__popcll(kept_mask & ((1ull << lane) - 1ull))] = h;
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
  }
  if (lane == 0) *q.n_kept = K;
}

__global__ void __launch_bounds__(ST) k_pose12(const double *T, uint32_t n, double *pose) {
  uint32_t h = blockIdx.x * ST + threadIdx.x;
  if (h >= n) return;
  const double *t = T + 16 * (size_t)h;
  double *o = pose + 12 * (size_t)h;
  for (int i = 0; i < 3; ++i) {
    for (int k = 0; k < 3; ++k) o[3 * i + k] = t[4 * i + k];
    o[9 + i] = t[4 * i + 3];
  }
}

__global__ void __launch_bounds__(ST) k_gather12(const double *pose, const uint32_t *idx, uint32_t n, double *out) {
  uint32_t k = blockIdx.x * ST + threadIdx.x;
  if (k >= 12 * n) return;
  out[k] = pose[12 * (size_t)idx[k / 12] + k % 12];
}

}  // namespace

hipError_t launch_gather12(const double *pose, const uint32_t *idx, uint32_t n, double *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather12, dim3(blocks((size_t)n * 12, ST)), dim3(ST), 0, s, pose, idx, n, out);
  return hipGetLastError();
}

hipError_t launch_hyp_count(const StagePair *sp, const uint32_t *row_first, int n_pairs, uint32_t rows, uint32_t *cnt,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_hyp_count, dim3(blocks((size_t)rows + 1, ST)), dim3(ST), 0, s, sp, row_first, n_pairs, rows,
                     cnt);
  return hipGetLastError();
}

hipError_t launch_hyp_build(const StagePair *sp, const uint32_t *row_first, int n_pairs, uint32_t rows,
                            const uint32_t *off, uint32_t n_hyp, uint32_t *pa, uint32_t *pb, uint32_t *hpair,
                            double *pose, hipStream_t s) {
  if (rows > 0)
    hipLaunchKernelGGL(k_hyp_fill, dim3(blocks(rows, ST)), dim3(ST), 0, s, sp, row_first, n_pairs, rows, off, pa, pb,
                       hpair);
  if (n_hyp > 0) hipLaunchKernelGGL(k_hyp_pose, dim3(blocks(n_hyp, ST)), dim3(ST), 0, s, sp, pa, pb, hpair, n_hyp, pose);
  return hipGetLastError();
}

hipError_t launch_stage_coarse(const StagePair *sp, const uint32_t *ipair, const double *pose, uint32_t n, int which,
                               uint32_t *counts, double *scores, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_stage_coarse, dim3(blocks((size_t)n * 64, ST)), dim3(ST), 0, s, sp, ipair, pose, n, which,
                     counts, scores);
  return hipGetLastError();
}

hipError_t launch_order(const uint32_t *ipair, const uint32_t *item_first, const uint32_t *counts, uint32_t n,
                        uint64_t *keys, uint32_t *vals, uint64_t *tk, uint32_t *tv, uint32_t *hist, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_order_keys, dim3(blocks(n, ST)), dim3(ST), 0, s, ipair, item_first, counts, n, keys, vals);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return rsort::sort_pairs<uint64_t>(keys, vals, tk, tv, hist, (int)n, 0, 64, s);
}

size_t order_hist_words(uint32_t n) { return rsort::hist_words((int)std::max<uint32_t>(n, 1)); }

hipError_t launch_nms(const NmsProblem *probs, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nms, dim3(n), dim3(64), 0, s, probs);
  return hipGetLastError();
}

hipError_t launch_pose12(const double *T, uint32_t n, double *pose, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pose12, dim3(blocks(n, ST)), dim3(ST), 0, s, T, n, pose);
  return hipGetLastError();
}

}  // namespace sherd
