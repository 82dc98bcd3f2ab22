// R §6 verification on the device: verify::verify (crates/sherd-core/src/matching/verify.rs:590-626)
// and verify::accept (:632-638) for a batch of candidates over many pairs at once.
//
// The statistics are counts, medians, a maximum and a minimum, so none depends on the order the
// per-point results arrive in: counts are atomic increments, medians are exact selections (radix
// select on the f64 total order), the deepest excursion and the closest approach are atomic
// max/min on the bits of non-negative floats. The per-point values themselves are the executor's
// arithmetic (bvh_query.hpp, grid.hpp), so every score equals the CPU restatement's bit for bit.
//
// Kernels, each over all candidates of the batch:
//   k_prep      pose inverse per candidate (verify.rs:739-798), counters cleared, the apart flag
//   k_frac      R §6.1: both sides' fracture samples against the other's fracture scene
//   k_seam      R §6.2: A's breakline against the moved B breakline (through B's breakline grid),
//               voxel set of the seam points
//   k_cont      R §6.3: B's margin against A's margin grid, steps and normal agreements
//   k_pen       R §6.4: surface samples inside the other mesh's box, listed
//   k_pen_groups, k_pen_rays  those samples' parity rays in full waves: the inside ones counted and listed with
//               their nearest crossing
//   k_pen_depth the listed samples' depths: deeper-than-pen count, deepest
//   k_pen_seed  R §6.4's fallback when nothing is inside, the closest approach: per (candidate,
//               side) the exact distance of the sample of least estimate, the first window
//   k_pen_close the blocks of samples nearest first, eight workgroups per (candidate, side)
//   k_pen_hard  the closest-approach queries too long for one lane, one wave each
//   k_median    the four medians (numpy semantics, verify.rs:693-700), each pair after its section
//   k_final     Scores and the accept verdict
// The closest approach runs twice: for the candidates whose sample sets both miss the other mesh's
// box (nothing can be inside) beside k_pen, for the others after it (beside k_pen_depth).
#include <hip/hip_runtime.h>

#include "bvh_query.hpp"
#include "grid.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "verify.hpp"

namespace sherd {
namespace {

enum { S_T, S_RES, S_COARSE, S_STAGE1, S_TIGHT, S_FACING, S_GAP, S_SEAM, S_NEAR, S_PEN, S_NMS, S_ICP };

constexpr int VT = 256;
constexpr int PEN_STEPS = 64;      // nodes one lane walks for a closest-approach query before a wave takes over
constexpr int HARD_STACK = 2048;   // pending nodes of a wave's cooperative query (LDS, 8 KiB per wave)
constexpr int HARD_BLOCKS = 1024;  // workgroups of the cooperative pass
// Steps (up to 64 nodes each) one wave may spend on one closest-approach query: sixteen times the
// 2^21 nodes an exhaustive walk of a 2^20-triangle mesh visits. A query that needs more sets
// FLAG_GAVE_UP and the batch fails with a status (the caller answers it on the CPU), so the kernel
// always ends in bounded time.
constexpr int WAVE_STEPS = 1 << 19;
constexpr int DEPTH_VT = 64;        // one wave per workgroup of k_pen_depth: its 16 KB of LDS stacks fit beside
                                    // the side streams' kernels on a CU (256 lanes held 64 KB, 2 waves per SIMD)
constexpr int DEPTH_BLOCKS = 4096;  // workgroups of k_pen_depth, which loop over the listed samples
// k_pen_depth's deferred nodes: its samples fill few waves, so a deep LDS stack (64 KiB per
// workgroup) costs no occupancy they could use, and spares the overflow walk.
constexpr int DEPTH_LSTACK = 32;


__device__ __forceinline__ void apply_f64(const double *T, const double *p, double *o) {
  o[0] = T[0] * p[0] + T[1] * p[1] + T[2] * p[2] + T[3];
  o[1] = T[4] * p[0] + T[5] * p[1] + T[6] * p[2] + T[7];
  o[2] = T[8] * p[0] + T[9] * p[1] + T[10] * p[2] + T[11];
}
__device__ __forceinline__ void rotate_f64(const double *T, const double *p, double *o) {
  o[0] = T[0] * p[0] + T[1] * p[1] + T[2] * p[2];
  o[1] = T[4] * p[0] + T[5] * p[1] + T[6] * p[2];
  o[2] = T[8] * p[0] + T[9] * p[1] + T[10] * p[2];
}
__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// The (candidate, side) task and the point of this thread in a launch of per_task workgroups per
// task. The tasks are ordered by the fragment they query (VerifyBatch::task), and xcd_logical keeps
// a run of them on one XCD, so the workgroups that walk one mesh share that XCD's L2.
// The bounding sphere of the 64-sample block holding sample i (a wave's lanes are one block),
// moved by T: its centre in f64 and its radius.
struct MovedBlock {
  double c[3];
  double r;
};
__device__ __forceinline__ MovedBlock moved_block(const double *blk, int i, const double *T) {
  const double *q = blk + 4 * (size_t)(i >> 6);
  MovedBlock m;
  apply_f64(T, q, m.c);
  m.r = q[3];
  return m;
}
__device__ __forceinline__ V3 centre_f32(const MovedBlock &m) {
  return V3{(float)m.c[0], (float)m.c[1], (float)m.c[2]};
}
// How far every point of the block is from the box [lo, hi] at least, less the pruning slack (which
// covers the f32 rounding of each moved sample and of the distances computed from it): the
// Euclidean gap from the centre to the box, minus the radius and the slack. Positive: no sample
// of the block reaches the box, or anything in it, by that much.
__device__ __forceinline__ double block_box_gap(const MovedBlock &m, const float *lo, const float *hi, float absmax) {
  double g2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double g = fmax(fmax((double)lo[k] - m.c[k], m.c[k] - (double)hi[k]), 0.0);
    g2 += g * g;
  }
  return sqrt(g2) - m.r - (double)slack_of(absmax, centre_f32(m));
}
// A lower bound of every block sample's distance to the mesh (the distance bound at the centre,
// less the radius and the slack; 0 when that is negative).
__device__ __forceinline__ float block_bound(const SceneView &sv, const MovedBlock &m) {
  V3 cf = centre_f32(m);
  double b = (double)dist_bound(sv.db, cf) - m.r - (double)slack_of(sv.absmax, cf);
  return b > 0.0 ? (float)b * 0.99999f : 0.0f;
}

// ---- the tree's cut as a distance bound --------------------------------------------------------
// Far from a mesh the distance-bound grid only knows the distance to its own box; the nearest of
// the cut's boxes (DistBound::cut) is a far tighter bound there. Two uses: a filter that only
// skips what a traversal would prune anyway, so every result is unchanged, and an estimate that
// orders work (which block or sample to seed with).

// What a block of samples can reach through the cut within w: lane l tests box l (the wave's
// lanes hold the cut, one box each, in *lo / *hi), and the ballot is the boxes some sample of
// the block may find a triangle within w in. A box is dropped when its gap from the centre, less
// the radius and three times the slack (which covers the growth of the box by a sample's slack,
// the rounding of the moved samples and of these distances), exceeds w. Every lane calls it; all
// ones without a cut.
__device__ __forceinline__ uint64_t cut_near_block(const SceneView &sv, const MovedBlock &m, float w, float *lo,
                                                   float *hi) {
  const int lane = threadIdx.x & 63;
  const int n = sv.db.n_cut;
  if (n == 0) return ~0ull;
  bool near = false;
  if (lane < n) {
    const float *b = sv.db.cut + 8 * lane;
    float g2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
      lo[k] = b[k];
      hi[k] = b[4 + k];
      const float x = (float)m.c[k];
      const float g = fmaxf(fmaxf(lo[k] - x, x - hi[k]), 0.0f);
      g2 += g * g;
    }
    const float margin = 3.0f * (slack_of(sv.absmax, centre_f32(m)) + 1e-5f * (float)m.r);
    near = sqrtf(g2) - (float)m.r - margin <= w;
  }
  return __ballot(near);
}

// The smallest squared distance from q to the boxes of `mask` grown by pad, in box_d2's own
// arithmetic: when it exceeds a traversal's (window + pad)^2 the traversal, which tests each of
// these nodes on the way to any triangle below them, measures nothing. *far2 gets the smallest
// squared distance from q to the farthest point of one of these boxes: every triangle of a box
// lies within that of q, so q's distance to the mesh does too (both 0 / +inf without a cut).
// Every lane calls it, with the boxes in lo / hi of the lanes (cut_near_block, cut_lanes).
__device__ __forceinline__ float cut_d2(const SceneView &sv, uint64_t mask, const float *lo, const float *hi, V3 q,
                                        float pad, float *far2 = nullptr) {
  if (far2) *far2 = __builtin_inff();
  if (sv.db.n_cut == 0) return 0.0f;
  float best = __builtin_inff(), fbest = __builtin_inff();
  while (mask) {
    const int k = __ffsll((unsigned long long)mask) - 1;
    mask &= mask - 1;
    Node nd;
    for (int j = 0; j < 3; ++j) {
      nd.lo[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lo[j]), k));
      nd.hi[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hi[j]), k));
    }
    const float d2 = box_d2(nd, q, pad);
    best = d2 < best ? d2 : best;
    if (far2) {
      const float f0 = fmaxf(q.x - nd.lo[0], nd.hi[0] - q.x), f1 = fmaxf(q.y - nd.lo[1], nd.hi[1] - q.y),
                  f2 = fmaxf(q.z - nd.lo[2], nd.hi[2] - q.z);
      fbest = fminf(fbest, (f0 * f0 + f1 * f1) + f2 * f2);
    }
  }
  if (far2) *far2 = fbest;
  return best;
}

// A search radius that certainly reaches a triangle, from a far-corner distance of the cut
// (cut_d2's far2), allowing for the f32 rounding of that distance and of the triangle's.
__device__ __forceinline__ float cut_reach(float far2, float pad) { return sqrtf(far2) * 1.0001f + pad; }

// The cut with one box per lane (lanes past its size hold an empty box). Every lane calls it.
struct CutLanes {
  float lo[3], hi[3];
};
__device__ __forceinline__ CutLanes cut_lanes(const SceneView &sv) {
  const int lane = threadIdx.x & 63;
  CutLanes c;
  for (int k = 0; k < 3; ++k) {
    c.lo[k] = lane < sv.db.n_cut ? sv.db.cut[8 * lane + k] : __builtin_inff();
    c.hi[k] = lane < sv.db.n_cut ? sv.db.cut[8 * lane + 4 + k] : -__builtin_inff();
  }
  return c;
}
// The smallest squared distance from each lane's x (a point, or a block's centre) to the cut's
// boxes, read from the lanes that hold them. Every lane calls it.
__device__ __forceinline__ float cut_lanes_d2(const SceneView &sv, const CutLanes &c, V3 x) {
  float best = __builtin_inff();
  for (int j = 0; j < sv.db.n_cut; ++j) {
    float g2 = 0.0f;
    const float xs[3] = {x.x, x.y, x.z};
    for (int k = 0; k < 3; ++k) {
      const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.lo[k]), j));
      const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.hi[k]), j));
      const float g = fmaxf(fmaxf(lo - xs[k], xs[k] - hi), 0.0f);
      g2 += g * g;
    }
    best = fminf(best, g2);
  }
  return best;
}
// The boxes that can be the nearest for some sample of a block: those whose gap from the block
// (the centre's gap less the radius) is no more than the smallest far reach of any box (the
// farthest point of the box from the centre, plus the radius). The nearest box of any sample of
// the block is one of them, up to rounding, which an estimate may ignore. One box per lane (the
// lanes of cut_lanes); every lane calls it.
__device__ __forceinline__ uint64_t cut_seed_mask(const SceneView &sv, const CutLanes &c, const MovedBlock &m) {
  const int lane = threadIdx.x & 63;
  if (sv.db.n_cut == 0) return 0;
  const V3 x = centre_f32(m);
  const float xs[3] = {x.x, x.y, x.z}, r = (float)m.r;
  float g2 = 0.0f, f2 = 0.0f;
  for (int k = 0; k < 3; ++k) {
    const float g = fmaxf(fmaxf(c.lo[k] - xs[k], xs[k] - c.hi[k]), 0.0f);
    const float f = fmaxf(xs[k] - c.lo[k], c.hi[k] - xs[k]);
    g2 += g * g;
    f2 += f * f;
  }
  const bool real = lane < sv.db.n_cut;
  const float ub = real ? sqrtf(f2) + r : __builtin_inff();
  const float ub_min = __uint_as_float(wave_min(__float_as_uint(ub)));
  return __ballot(real && sqrtf(g2) - r <= ub_min);
}

// The seed's estimate for a moved sample: the grid's bound and the distance to the nearest cut
// box among `mask` (cut_seed_mask of its block), the larger.
__device__ __forceinline__ float sample_key_mask(const SceneView &sv, const CutLanes &c, uint64_t mask, V3 q) {
  const float b = dist_bound(sv.db, q);
  if (sv.db.n_cut == 0) return b;
  return fmaxf(b, sqrtf(cut_d2(sv, mask, c.lo, c.hi, q, 0.0f)));
}

__device__ __forceinline__ MovedBlock moved_sphere(const double *s, const double *T) {
  MovedBlock m;
  apply_f64(T, s, m.c);
  m.r = s[3];
  return m;
}

// The per-sample passes give a workgroup CHUNKS consecutive chunks of VT samples of one
// (candidate, side) task. One: a workgroup walking eight chunks ran the vessel's passes 2.8x
// slower (fewer waves in flight to hide the dependent loads of the traversals).
constexpr int CHUNKS = 1;
struct TaskGroup {
  int c, side, chunk0;
};
__device__ __forceinline__ TaskGroup task_group(const VerifyBatch &vb, int per_task) {
  uint32_t lb = xcd_logical(blockIdx.x, gridDim.x);
  uint32_t cs = vb.task[lb / (uint32_t)per_task];
  return TaskGroup{(int)(cs >> 1), (int)(cs & 1u), (int)(lb % (uint32_t)per_task) * CHUNKS};
}

__global__ void k_prep(VerifyBatch vb) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= vb.n_cand) return;
  CandDesc &cd = vb.cand[c];
  pose_inverse(cd.T, cd.Ti);
  // Apart: k_pen's whole-set tests dismiss both sides (no sample can be inside either mesh), so
  // both closest approaches are needed whatever k_pen finds, and they can start before it.
  {
    const PairDesc &pd = vb.pair[cd.pair];
    const FragDesc &A = vb.frag[pd.a], &B = vb.frag[pd.b];
    cd.apart = pd.pen_ok &&
               block_box_gap(moved_sphere(A.s_all, cd.Ti), B.mesh.lo, B.mesh.hi, B.mesh.absmax) > 0.0 &&
               block_box_gap(moved_sphere(B.s_all, cd.T), A.mesh.lo, A.mesh.hi, A.mesh.absmax) > 0.0;
  }
  CandAcc &a = vb.acc[c];
  for (int s = 0; s < 2; ++s) {
    a.facing[s] = a.tight[s] = a.touching[s] = 0;
    a.inside[s] = a.deeper[s] = 0;
    a.in_box[s] = 0;
    a.deepest[s] = 0u;
    a.closest[s] = 0x7f7fffffu;  // f32::MAX, the window of the fallback (verify.rs:570-577)
  }
  a.n_cont = 0;
  if (c == 0) vb.hard_n[0] = vb.hard_n[1] = vb.hard_n[2] = vb.hard_n[3] = vb.hard_n[4] = vb.hard_n[5] = 0;
  a.seam_cells = 0;
  a.flags = 0;
}

// R §6.1 one_side inputs (verify.rs:318-349): side 0 is A's samples moved by T^-1 against B's
// fracture, side 1 is B's samples moved by T against A's fracture.
__global__ void __launch_bounds__(VT) k_frac(VerifyBatch vb, int per_task) {
  __shared__ uint2 s_stk[LSTACK][VT];
  const TaskGroup g = task_group(vb, per_task);
  const int c = g.c, side = g.side;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
  const FragDesc &other = vb.frag[side == 0 ? pd.b : pd.a];
  const double *T = side == 0 ? cd.Ti : cd.T;
  const double facing_r = pd.sc[S_FACING];
  // Samples beyond the facing distance of the fracture scene's box add nothing (no facing, tight
  // or touching sample): the whole set, then each wave's block, is tested first.
  if (own.n_pf == 0 ||
      block_box_gap(moved_sphere(own.pf_all, T), other.frac.lo, other.frac.hi, other.frac.absmax) > facing_r)
    return;
  CandAcc &acc = vb.acc[c];
  LdsStack<VT> st{&s_stk[0][threadIdx.x]};
  for (int k = 0; k < CHUNKS; ++k) {
    const int base = (g.chunk0 + k) * VT;
    if (base >= own.n_pf) break;  // (uniform)
    const int i = base + threadIdx.x, i0 = i & ~63;
    if (i0 >= own.n_pf ||
        block_box_gap(moved_block(own.pf_blk, i0, T), other.frac.lo, other.frac.hi, other.frac.absmax) > facing_r)
      continue;  // (uniform across the wave)
    bool facing = false, tight = false, touching = false;
    double d64 = 0.0;
    if (i < own.n_pf) {
      const double *p = own.pf + 3 * (size_t)i;
      V3 q = moved_f32(T, p[0], p[1], p[2]);
      float d;
      float window = (float)facing_r;
      d64 = bounded_distance_with(st, other.frac, q, window, &d) ? (double)d : __builtin_inf();
      facing = d64 < facing_r;
      tight = facing && d64 < pd.sc[S_TIGHT];
      touching = d64 < 2.0 * pd.sc[S_TIGHT];
    }
    uint32_t slot = wave_slot(&acc.facing[side], facing);
    if (facing) vb.keys[cd.seg[side] + slot] = total_key(d64);
    uint32_t nt = (uint32_t)__popcll(__ballot(tight)), nh = (uint32_t)__popcll(__ballot(touching));
    if ((threadIdx.x & 63) == 0) {
      if (nt) atomicAdd(&acc.tight[side], nt);
      if (nh) atomicAdd(&acc.touching[side], nh);
    }
  }
}

// The nearest point of B's breakline moved by T to p (A's frame), with kiddo's f64 distance to
// T b and ties to the lowest index: exactly what a scan over all of B would pick, provided that
// point lies within the breakline grid's reach. The grid is over B's own frame and is entered at
// T^-1 p; a point within the reach of p in A's frame is within the reach (plus rounding, which the
// reach's margin covers) of T^-1 p in B's, so it is among the 27 cells searched.
__device__ inline bool seam_nearest(const GridView &g, const double *T, const double *Ti, const double *bp,
                                    const double *p, double reach, uint32_t *out_j, double *out_d2) {
  if (g.n <= 0) return false;
  double q[3];
  apply_f64(Ti, p, q);
  // Nothing of B within the reach (the caller keeps only points strictly inside it): no scan.
  if (grid_far(g, q[0], q[1], q[2], reach)) return false;
  double fx = floor((q[0] - g.origin[0]) / g.cell);
  double fy = floor((q[1] - g.origin[1]) / g.cell);
  double fz = floor((q[2] - g.origin[2]) / g.cell);
  if (!(fx >= -1.0 && fx <= (double)GRID_MAXC && fy >= -1.0 && fy <= (double)GRID_MAXC && fz >= -1.0 &&
        fz <= (double)GRID_MAXC))
    return false;
  double best = __builtin_inf();
  uint32_t bj = 0xFFFFFFFFu;
  grid_scan27(g, (int64_t)fx, (int64_t)fy, (int64_t)fz, [&](uint32_t k) {
    uint32_t oi = gldu(g.idx + k);
    double m[3];
    apply_f64(T, bp + 3 * (size_t)oi, m);
    double d2 = d2_f64(p[0], p[1], p[2], m[0], m[1], m[2]);
    if (d2 < best || (d2 == best && oi < bj)) {
      best = d2;
      bj = oi;
    }
  });
  if (bj == 0xFFFFFFFFu) return false;
  *out_j = bj;
  *out_d2 = best;
  return true;
}

// Clears each candidate's voxel set.
__global__ void __launch_bounds__(VT) k_seam_clear(VerifyBatch vb) {
  const CandDesc &cd = vb.cand[blockIdx.x];
  unsigned long long *table = (unsigned long long *)(vb.seam_table + cd.seam_off);
  for (uint32_t s = threadIdx.x; s <= cd.seam_mask; s += VT) table[s] = ~0ull;
}

// R §6.2 seam_points / seam_cells (verify.rs:378-442): for each point of A's breakline, the nearest
// moved point of B's breakline (f64, ties to the lowest index); a seam point when it is strictly
// inside sc.seam and the two shell normals agree; the score counts distinct t/3 voxels of them.
__global__ void __launch_bounds__(VT) k_seam(VerifyBatch vb) {
  const int c = blockIdx.y;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  const FragDesc &A = vb.frag[pd.a];
  const FragDesc &B = vb.frag[pd.b];
  const double seam = pd.sc[S_SEAM];
  if (A.n_brk == 0 || B.brk.n <= 0 || !(seam > 0.0)) return;
  // A's whole breakline, moved into B's frame, farther than the seam from B's: nothing to find.
  {
    const MovedBlock all = moved_sphere(A.brk_all, cd.Ti);
    if (grid_far(B.brk, all.c[0], all.c[1], all.c[2], seam + all.r)) return;
  }
  const double voxel = pd.sc[S_T] / 3.0;
  const int64_t H = 1ll << 20;
  unsigned long long *table = (unsigned long long *)(vb.seam_table + cd.seam_off);
  const uint32_t mask = cd.seam_mask;
  for (int k = 0; k < CHUNKS; ++k) {
    const int base = ((int)blockIdx.x * CHUNKS + k) * VT;
    if (base >= A.n_brk) break;  // (uniform)
    const int i = base + threadIdx.x;
    if (i >= A.n_brk) continue;
    const double *p = A.brk_p + 3 * (size_t)i;
    uint32_t j;
    double d2;
    if (!seam_nearest(B.brk, cd.T, cd.Ti, B.brk_p, p, seam, &j, &d2)) continue;
    if (!(sqrt(d2) < seam)) continue;
    double nb[3];
    rotate_f64(cd.T, B.brk_ns + 3 * (size_t)j, nb);
    if (!(dot3(A.brk_ns + 3 * (size_t)i, nb) > 0.7)) continue;
    int64_t x = (int64_t)floor(p[0] / voxel), y = (int64_t)floor(p[1] / voxel), z = (int64_t)floor(p[2] / voxel);
    if (x < -H || x >= H || y < -H || y >= H || z < -H || z >= H) {
      atomicOr(&vb.acc[c].flags, FLAG_VOXEL_RANGE);
      continue;
    }
    uint64_t key = (uint64_t)(x + H) | ((uint64_t)(y + H) << 21) | ((uint64_t)(z + H) << 42);
    uint32_t slot = grid_hash(key, mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      unsigned long long prev = atomicCAS(&table[slot], ~0ull, (unsigned long long)key);
      if (prev == ~0ull) {
        atomicAdd(&vb.acc[c].seam_cells, 1u);
        break;
      }
      if (prev == key) break;
      slot = (slot + 1) & mask;
    }
  }
}

// R §6.3 continuity_scores (verify.rs:450-474).
__global__ void __launch_bounds__(VT) k_cont(VerifyBatch vb) {
  const int c = blockIdx.y;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  const FragDesc &A = vb.frag[pd.a];
  const FragDesc &B = vb.frag[pd.b];
  const double near = pd.sc[S_NEAR];
  if (B.n_margin == 0 || A.n_margin == 0) return;
  // B's whole margin, moved into A's frame, farther than `near` from A's: nothing to pair.
  {
    const MovedBlock all = moved_sphere(B.margin_all, cd.T);
    if (grid_far(A.margin, all.c[0], all.c[1], all.c[2], near + all.r)) return;
  }
  CandAcc &acc = vb.acc[c];
  for (int k = 0; k < CHUNKS; ++k) {
    const int base = ((int)blockIdx.x * CHUNKS + k) * VT;
    if (base >= B.n_margin) break;  // (uniform)
    const int i = base + threadIdx.x;
    bool take = false;
    double step = 0.0, agree = 0.0;
    if (i < B.n_margin) {
      double moved[3];
      apply_f64(cd.T, B.margin_p + 3 * (size_t)i, moved);
      uint32_t j;
      double dist;
      if (!grid_far(A.margin, moved[0], moved[1], moved[2], near) &&
          grid_nearest_below(A.margin, moved[0], moved[1], moved[2], near, &j, &dist)) {
        const double *np = A.margin_p + 3 * (size_t)j, *nn = A.margin_n + 3 * (size_t)j;
        double delta[3] = {moved[0] - np[0], moved[1] - np[1], moved[2] - np[2]};
        step = fabs(dot3(delta, nn));
        double turned[3];
        rotate_f64(cd.T, B.margin_n + 3 * (size_t)i, turned);
        agree = dot3(turned, nn);
        take = true;
      }
    }
    uint32_t slot = wave_slot(&acc.n_cont, take);
    if (take) {
      vb.keys[cd.seg[2] + slot] = total_key(step);
      vb.keys[cd.seg[3] + slot] = total_key(agree);
    }
  }
}

// R §6.4 one_penetration's first pass (verify.rs:541-566), RayScene::inside's AABB reject: side 0
// is A's surface samples moved by T^-1 against B's mesh, side 1 is B's moved by T against A's mesh.
// The samples inside the mesh's box (a tenth of those k_pen reads on the collection) are counted
// and listed in the task's own segment of boxq (one counter per task, so the waves' atomics spread
// over the tasks), for k_pen_rays, which casts their rays one group of 64 entries per wave.
__global__ void __launch_bounds__(VT) k_pen(VerifyBatch vb, int per_task) {
  const TaskGroup g = task_group(vb, per_task);
  const int c = g.c, side = g.side;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
  const FragDesc &other = vb.frag[side == 0 ? pd.b : pd.a];
  if (!pd.pen_ok || own.n_s == 0) return;  // (uniform: a workgroup belongs to one task)
  const SceneView &sv = other.mesh;
  const double *T = side == 0 ? cd.Ti : cd.T;
  // No sample can be inside when the whole set lies outside the mesh's box (by more than the
  // rounding), nor a block's when its sphere does: those skip the moves.
  if (block_box_gap(moved_sphere(own.s_all, T), sv.lo, sv.hi, sv.absmax) > 0.0) return;
  CandAcc &acc = vb.acc[c];
  uint32_t *list = vb.boxq + (size_t)(2 * c + side) * (size_t)vb.box_cap;
  for (int k = 0; k < CHUNKS; ++k) {
    const int base = (g.chunk0 + k) * VT;
    if (base >= own.n_s) break;  // (uniform)
    const int i = base + threadIdx.x, i0 = i & ~63;
    if (i0 >= own.n_s || block_box_gap(moved_block(own.s_blk, i0, T), sv.lo, sv.hi, sv.absmax) > 0.0)
      continue;  // (uniform per wave: a wave is one block)
    bool in = false;
    if (i < own.n_s) {
      const double *p = own.s + 3 * (size_t)i;
      const V3 q = moved_f32(T, p[0], p[1], p[2]);
      in = !(q.x < sv.lo[0] || q.x > sv.hi[0] || q.y < sv.lo[1] || q.y > sv.hi[1] || q.z < sv.lo[2] ||
             q.z > sv.hi[2]);
    }
    const uint32_t slot = wave_slot(&acc.in_box[side], in);
    if (in) list[slot] = (uint32_t)i;
  }
}

// The groups of 64 entries of every task's in-box list, in the tasks' order (by the fragment they
// query), into boxw, and their number into hard_n[5]: one workgroup, each thread a run of tasks,
// their offsets by a prefix sum in LDS into gstart, then the groups written by whole waves. (A
// global counter that every group took a slot of cost more than this launch: one address's atomics
// are serialised across the device.)
constexpr int GROUPS_VT = 1024;
__global__ void __launch_bounds__(GROUPS_VT) k_pen_groups(VerifyBatch vb) {
  __shared__ uint32_t s_sum[GROUPS_VT];
  const int nt = 2 * vb.n_cand, per = (nt + GROUPS_VT - 1) / GROUPS_VT;
  const int t0 = (int)threadIdx.x * per, t1 = t0 + per < nt ? t0 + per : nt;
  auto groups = [&](int j) {
    const uint32_t cs = vb.task[j];
    return (vb.acc[cs >> 1].in_box[cs & 1u] + 63u) >> 6;
  };
  uint32_t mine = 0;
  for (int j = t0; j < t1; ++j) mine += groups(j);
  s_sum[threadIdx.x] = mine;
  __syncthreads();
  for (int off = 1; off < GROUPS_VT; off <<= 1) {  // inclusive scan
    const uint32_t v = (int)threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
    __syncthreads();
    s_sum[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t at = s_sum[threadIdx.x] - mine;
  if (threadIdx.x == GROUPS_VT - 1) vb.hard_n[5] = s_sum[threadIdx.x];
  for (int j = t0; j < t1; ++j) {  // each task's first group
    vb.gstart[j] = at;
    at += groups(j);
  }
  if (t0 < t1 && t1 == nt) vb.gstart[nt] = at;
  __syncthreads();
  // The groups written by whole waves: a wave takes 64 tasks at a time, one per lane, and its lanes
  // then write each of those tasks' groups as one contiguous run (in-box counts are skewed: one
  // thread per task range, writing its groups in turn, waited on its largest tasks for 0.19 ms).
  const uint32_t lane = threadIdx.x & 63;
  for (int j0 = (int)(threadIdx.x >> 6) * 64; j0 < nt; j0 += GROUPS_VT) {
    const int j = j0 + (int)lane;
    const uint32_t st = j < nt ? vb.gstart[j] : 0u, g = j < nt ? vb.gstart[j + 1] - st : 0u;
    const uint32_t cs = j < nt ? vb.task[j] : 0u;
    uint64_t todo = __ballot(g != 0u);
    while (todo) {
      const int l = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t sl = __shfl(st, l), gl = __shfl(g, l), cl = __shfl(cs, l);
      for (uint32_t q = lane; q < gl; q += 64) vb.boxw[sl + q] = ((uint64_t)cl << 32) | q;
    }
  }
}

// R §6.4's parity half for the samples k_pen listed, one group of 64 entries per wave (so the rays
// run in full waves whatever share of a task's samples is in the box), the waves of the grid
// taking the groups in turn (a logical workgroup's waves take consecutive groups, on one XCD):
// each lane's sample moved again, three axis rays with a majority (inside_parity); inside samples
// are counted and listed with their nearest ray crossing for k_pen_depth.
constexpr int RAY_BLOCKS = 2048;
__global__ void __launch_bounds__(VT) k_pen_rays(VerifyBatch vb) {
  const uint32_t n = vb.hard_n[5];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t lb = xcd_logical(blockIdx.x, gridDim.x);
  for (uint32_t w = (lb * VT + threadIdx.x) >> 6; w < n; w += (gridDim.x * VT) >> 6) {
    const uint64_t item = vb.boxw[w];
    const uint32_t task = (uint32_t)(item >> 32), k = 64u * (uint32_t)item + lane;
    const int c = (int)(task >> 1), side = (int)(task & 1u);
    CandAcc &acc = vb.acc[c];
    bool in = false;
    float t = 0.0f;
    uint32_t i = 0;
    if (k < acc.in_box[side]) {
      const CandDesc &cd = vb.cand[c];
      const PairDesc &pd = vb.pair[cd.pair];
      const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
      const SceneView &sv = vb.frag[side == 0 ? pd.b : pd.a].mesh;
      const double *T = side == 0 ? cd.Ti : cd.T;
      i = vb.boxq[(size_t)task * (size_t)vb.box_cap + k];
      const double *p = own.s + 3 * (size_t)i;
      in = inside_parity(sv, moved_f32(T, p[0], p[1], p[2]), vb.retest, &t);
    }
    const uint32_t ni = (uint32_t)__popcll(__ballot(in));
    if (lane == 0 && ni) atomicAdd(&acc.inside[side], ni);
    const uint32_t slot = wave_slot(&vb.hard_n[2], in);
    if (in) {
      vb.deep[slot] = ((uint64_t)task << 32) | i;
      vb.deep_t[slot] = t;
    }
  }
}

// R §6.4's depths of the inside samples k_pen listed, one lane each, so the searches run in full
// waves: the distance to the nearest triangle (the same search RayScene::inside's depth makes),
// the count deeper than pen and the deepest per side.
__global__ void __launch_bounds__(DEPTH_VT) k_pen_depth(VerifyBatch vb) {
  __shared__ uint2 s_stk[DEPTH_LSTACK][DEPTH_VT];
  LdsStack<DEPTH_VT, DEPTH_LSTACK> st{&s_stk[0][threadIdx.x]};
  const uint32_t n = vb.hard_n[2];
  for (uint32_t k = blockIdx.x * DEPTH_VT + threadIdx.x; k < n; k += gridDim.x * DEPTH_VT) {
    const uint64_t e = vb.deep[k];
    const int c = (int)(e >> 33), side = (int)((e >> 32) & 1u), i = (int)(uint32_t)e;
    const CandDesc &cd = vb.cand[c];
    const PairDesc &pd = vb.pair[cd.pair];
    const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
    const SceneView &sv = vb.frag[side == 0 ? pd.b : pd.a].mesh;
    const double *T = side == 0 ? cd.Ti : cd.T;
    const double *p = own.s + 3 * (size_t)i;
    const V3 q = moved_f32(T, p[0], p[1], p[2]);
    CandAcc &acc = vb.acc[c];
    // The search bound: the nearest crossing's, or the distance-bound grid's upper bound when that
    // is smaller (a ray along a wall can cross it far away).
    const float depth = depth_within(st, sv, q, fminf(depth_bound(sv, q, vb.deep_t[k]), dist_upper(sv.db, q)));
    const bool deeper = (double)depth > pd.sc[S_PEN];
    if (deeper) atomicAdd(&acc.deeper[side], 1u);
    if (__float_as_uint(depth)) atomicMax(&acc.deepest[side], __float_as_uint(depth));
  }
}

// Whether side `side`'s closest approach is measured: when nothing of it is inside the other mesh.
// pen_depth is max(-min_a, -min_b) / t (verify.rs:494), min being -deepest on a side with a sample
// inside and the closest approach (at least 0) on a side without (k_final). When the other side has
// a sample inside at a positive depth this side's term is never read, but it is measured anyway:
// that needs the depths (k_pen_depth), and without them the closest approach starts as soon as
// k_pen has counted the inside samples.
__device__ __forceinline__ bool closest_needed(const CandAcc &acc, int side) { return acc.inside[side] == 0; }

// The window a side's closest approach may prune with: the best distance attained so far on this
// side, and on the other side too when that side has nothing inside. Only min(closest_a, closest_b)
// is read then, so a side whose own minimum is above the other's may stop at any attained value
// above it, and a side whose minimum is below is found exactly, every window being at least it.
__device__ __forceinline__ float closest_window(CandAcc &acc, int side) {
  uint32_t w = __hip_atomic_load(&acc.closest[side], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (acc.inside[1 - side] == 0) {
    uint32_t o = __hip_atomic_load(&acc.closest[1 - side], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w = o < w ? o : w;
  }
  return __uint_as_float(w);
}

// ---- R §6.4's closest approach when nothing is inside (verify.rs:567-578) --------------------------
//
// min(sd) is then the closest approach, a running minimum over the samples whose window shrinks to
// the best attained so far (DistReduce::Min, cpu.rs:100-111), from f32::MAX.

// The closest approach per task, one workgroup each. A seed sample's exact distance is the first
// window (see below); then the task's 64-sample blocks are taken in order of a lower bound (the
// larger of the grid's block bound and the cut's, with the margins cut_near_block uses), so the
// nearest blocks shrink the window first, and once the least bound left exceeds the window every
// remaining block is known to hold nothing below it. The waves share the ordered blocks, each
// claiming the least one left; a task with more blocks than CLOSE_MAXB takes them in index order,
// each tested the same way.
//
// In a block, one lane per sample: a sample whose grid bound, or distance to every cut box the
// block can reach, already exceeds the window (by more than the slack, which covers the rounding
// of the moved point) cannot lower the minimum; the others search (min_d2) no farther than the
// window and their own upper bounds. A query that needs more than PEN_STEPS nodes is finished by
// a whole wave (k_pen_hard); what it met so far is a distance it attains, so it may still lower
// the window.
constexpr int CLOSE_MAXB = 2048;
constexpr int SEED_BLOCKS = 64;
constexpr int CLOSE_G = 8;  // workgroups per task of the block walk
constexpr uint32_t CLAIMED = 0xffffffffu;  // above every key (non-negative f32 bits)

__device__ __forceinline__ float close_key(const SceneView &sv, const CutLanes &c, const MovedBlock &m) {
  float b = block_bound(sv, m);
  if (sv.db.n_cut > 0) {
    const float margin = 3.0f * (slack_of(sv.absmax, centre_f32(m)) + 1e-5f * (float)m.r);
    b = fmaxf(b, sqrtf(cut_lanes_d2(sv, c, centre_f32(m))) - (float)m.r - margin);
  }
  return fmaxf(b, 0.0f);
}

// The (candidate, side) task of a closest-approach launch, or false when it has nothing to do: not
// in this pass (apart or not), its closest approach not read, or no samples.
struct CloseTask {
  int c, side;
  const FragDesc *own, *other;
  const double *T;
};
__device__ __forceinline__ bool close_task(const VerifyBatch &vb, uint32_t cs, int apart, CloseTask *t) {
  t->c = (int)(cs >> 1);
  t->side = (int)(cs & 1u);
  const CandDesc &cd = vb.cand[t->c];
  const PairDesc &pd = vb.pair[cd.pair];
  if (!pd.pen_ok || cd.apart != apart) return false;
  if (!closest_needed(vb.acc[t->c], t->side)) return false;
  t->own = &vb.frag[t->side == 0 ? pd.a : pd.b];
  t->other = &vb.frag[t->side == 0 ? pd.b : pd.a];
  t->T = t->side == 0 ? cd.Ti : cd.T;
  return t->own->n_s > 0;
}

// The seed of a task, one workgroup: the sample of least estimate (the grid's bound or the nearest
// cut box, the larger) over its blocks in key order while a block's key, less the grid's cell
// diagonal by which a sample's bound may fall below its block's, does not exceed the best estimate
// so far (at most SEED_BLOCKS blocks, four at a time, one per wave); its exact distance, by one whole
// wave (min_d2_wave, its frontier in LDS), is the first window of the task's block walk. Which
// sample seeds changes the work, never the minimum. A task with more than CLOSE_MAXB blocks has no
// seed.
// LDS (dynamic): the keys of the batch's largest sample set's blocks, their key order, then
// SEED_FRONT ints of frontier.
constexpr int SEED_FRONT = 1024;
__global__ void __launch_bounds__(VT) k_pen_seed(VerifyBatch vb, int apart, int max_nb) {
  extern __shared__ uint32_t s_dyn[];
  uint32_t *s_key = s_dyn, *s_ord = s_dyn + max_nb;
  int *s_front = reinterpret_cast<int *>(s_dyn + 2 * max_nb);
  __shared__ unsigned long long s_best;  // (estimate bits << 32 | sample) of the best so far
  __shared__ uint32_t s_bh;              // its estimate bits as the round began
  __shared__ int s_stop;
  CloseTask tk;
  if (!close_task(vb, vb.task[xcd_logical(blockIdx.x, gridDim.x)], apart, &tk)) return;
  const FragDesc &own = *tk.own, &other = *tk.other;
  const double *T = tk.T;
  CandAcc &acc = vb.acc[tk.c];
  const int side = tk.side, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nb = (own.n_s + 63) >> 6;
  if (nb > CLOSE_MAXB || nb > max_nb) return;
  const CutLanes cut = cut_lanes(other.mesh);
  for (int base = 64 * w; base < nb; base += VT) {  // (uniform per wave)
    const int b = base + lane;
    const float k = close_key(other.mesh, cut, moved_block(own.s_blk, 64 * (b < nb ? b : 0), T));
    if (b < nb) s_key[b] = __float_as_uint(k);
  }
  if (threadIdx.x == 0) {
    s_best = ~0ull;
    s_bh = CLAIMED;
  }
  __syncthreads();
  // Key order (ties by block index): each block's rank by counting the blocks before it.
  for (int j = threadIdx.x; j < nb; j += VT) {
    const uint32_t kj = s_key[j];
    int r = 0;
    for (int i = 0; i < nb; ++i) {
      const uint32_t ki = s_key[i];
      r += (ki < kj || (ki == kj && i < j)) ? 1 : 0;
    }
    s_ord[r] = (uint32_t)j;
  }
  __syncthreads();
  const float cell = 1.8f * other.mesh.db.h;
  const int last = nb < SEED_BLOCKS ? nb : SEED_BLOCKS;
  for (int base = 0; base < last; base += VT / 64) {
    const uint32_t bh = s_bh;
    const int k = base + w;
    bool take = false;
    int blk = 0;
    if (k < last) {
      blk = (int)s_ord[k];
      take = bh == CLAIMED || !(__uint_as_float(s_key[blk]) - cell > __uint_as_float(bh));
    }
    if (take) {  // (uniform per wave)
      const int i = 64 * blk + lane;
      const uint64_t mask = cut_seed_mask(other.mesh, cut, moved_block(own.s_blk, 64 * blk, T));
      const double *pi = own.s + 3 * (size_t)(i < own.n_s ? i : 0);
      const float e = sample_key_mask(other.mesh, cut, mask, moved_f32(T, pi[0], pi[1], pi[2]));
      const uint64_t v = i < own.n_s ? (((uint64_t)__float_as_uint(e) << 32) | (uint32_t)i) : ~0ull;
      const uint32_t hi = wave_min((uint32_t)(v >> 32));
      const uint32_t lo = wave_min((uint32_t)(v >> 32) == hi ? (uint32_t)v : 0xffffffffu);
      if (lane == 0) atomicMin(&s_best, ((unsigned long long)hi << 32) | lo);
    }
    __syncthreads();
    // Blocks go by in key order: once one is past the best estimate, so is every later one. (One
    // thread decides between two barriers, so every wave leaves the loop together.)
    if (threadIdx.x == 0) {
      const uint32_t nh = (uint32_t)(s_best >> 32);
      const int kl = base + VT / 64;
      s_bh = nh;
      s_stop = kl >= last || (nh != CLAIMED && __uint_as_float(s_key[s_ord[kl]]) - cell > __uint_as_float(nh));
    }
    __syncthreads();
    if (s_stop) break;
  }
  const uint32_t seed = (uint32_t)s_best;
  if (w != 0 || seed == 0xffffffffu) return;
  const double *p = own.s + 3 * (size_t)seed;
  const V3 q = moved_f32(T, p[0], p[1], p[2]);
  const float window = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(closest_window(acc, side))));
  float far2;
  (void)cut_d2(other.mesh, other.mesh.db.n_cut ? ~0ull >> (64 - other.mesh.db.n_cut) : 0ull, cut.lo, cut.hi, q, 0.0f,
               &far2);
  // The greedy descent's triangle is at a distance the seed attains: the search needs to reach no
  // farther (a relative 2^-20 above covers the root's rounding).
  const float radius = fminf(fminf(fminf(window, dist_upper(other.mesh.db, q)),
                                   cut_reach(far2, slack_of(other.mesh.absmax, q))),
                             sqrtf(greedy_d2(other.mesh, q)) * (1.0f + 0x1p-20f));
  bool complete;
  const float dist = sqrtf(min_d2_wave(other.mesh, q, radius, s_front, SEED_FRONT, WAVE_STEPS, &complete));
  if (lane == 0 && dist <= window) atomicMin(&acc.closest[side], __float_as_uint(dist));
  if (lane == 0 && !complete) atomicOr(&acc.flags, FLAG_GAVE_UP);
}

// The block walk, CLOSE_G workgroups per task after its seed: workgroup g takes blocks g,
// g + CLOSE_G, ... (a spread of the fragment's Morton-ordered blocks), its waves claiming the one of
// least key left, so a task's blocks are walked by CLOSE_G x 4 waves (a task far from the other
// mesh has most of its blocks within the bound). Every workgroup stops once its least key exceeds
// the task's window, which all of them lower.
__global__ void __launch_bounds__(VT) k_pen_close(VerifyBatch vb, int apart) {
  __shared__ uint2 s_stk[LSTACK][VT];
  __shared__ uint32_t s_key[CLOSE_MAXB / CLOSE_G];
  __shared__ uint32_t s_next;
  const uint32_t t = xcd_logical(blockIdx.x, gridDim.x);
  const int g = (int)(t % CLOSE_G);
  CloseTask tk;
  if (!close_task(vb, vb.task[t / CLOSE_G], apart, &tk)) return;  // (uniform: the workgroup is one task)
  const FragDesc &own = *tk.own, &other = *tk.other;
  const double *T = tk.T;
  const int c = tk.c, side = tk.side;
  CandAcc &acc = vb.acc[c];
  // The queue of queries too long for one lane: the apart pass's own.
  uint64_t *queue = apart ? vb.hard_a : vb.hard;
  uint32_t *queued = vb.hard_n + (apart ? 3 : 0);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nb = (own.n_s + 63) >> 6;
  const int nl = nb > g ? (nb - g + CLOSE_G - 1) / CLOSE_G : 0;  // this workgroup's blocks: g + CLOSE_G j
  if (nl == 0) return;
  const bool ordered = nb <= CLOSE_MAXB;
  const CutLanes cut = cut_lanes(other.mesh);
  if (ordered) {
    for (int base = 64 * w; base < nl; base += VT) {  // (uniform per wave)
      const int j = base + lane, b = g + CLOSE_G * (j < nl ? j : 0);
      const float k = close_key(other.mesh, cut, moved_block(own.s_blk, 64 * b, T));
      if (j < nl) s_key[j] = __float_as_uint(k);
    }
  }
  if (threadIdx.x == 0) s_next = 0;
  __syncthreads();
  LdsStack<VT> st{&s_stk[0][threadIdx.x]};
  for (;;) {
    int b;
    float w0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(closest_window(acc, side))));
    if (ordered) {
      uint32_t kmin = CLAIMED, bmin = 0xffffffffu;
      for (int j = lane; j < nl; j += 64) {
        const uint32_t k = __hip_atomic_load(&s_key[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k < kmin) {
          kmin = k;
          bmin = (uint32_t)j;
        }
      }
      const uint32_t khi = wave_min(kmin);
      if (khi == CLAIMED || __uint_as_float(khi) > w0) break;  // (uniform)
      const int j = (int)wave_min(kmin == khi ? bmin : 0xffffffffu);
      uint32_t old = 0;
      if (lane == 0) old = atomicExch(&s_key[j], CLAIMED);
      if (__builtin_amdgcn_readfirstlane(__shfl(old, 0)) == CLAIMED) continue;  // another wave took it
      b = g + CLOSE_G * j;
    } else {
      uint32_t nx = 0;
      if (lane == 0) nx = atomicAdd(&s_next, 1u);
      const int j = (int)__builtin_amdgcn_readfirstlane(__shfl(nx, 0));
      if (j >= nl) break;
      b = g + CLOSE_G * j;
    }
    const int i0 = 64 * b, i = i0 + lane;
    const MovedBlock mb = moved_block(own.s_blk, i0, T);
    if (block_bound(other.mesh, mb) > w0) continue;
    // The cut's boxes the block can reach within the window; none: the next block.
    float clo[3], chi[3];
    const uint64_t near = cut_near_block(other.mesh, mb, w0, clo, chi);
    if (near == 0) continue;
    const bool have = i < own.n_s;
    V3 q = V3{0.0f, 0.0f, 0.0f};
    if (have) {
      const double *p = own.s + 3 * (size_t)i;
      q = moved_f32(T, p[0], p[1], p[2]);
    }
    const float pad = slack_of(other.mesh.absmax, q);
    float far2;
    const float near_d2 = cut_d2(other.mesh, near, clo, chi, q, pad, &far2);
    uint32_t best = 0xffffffffu;
    bool hard = false;
    if (have) {
      float window = closest_window(acc, side);
      const float lim = window + pad;
      if (dist_bound(other.mesh.db, q) <= lim && window > 0.0f && near_d2 <= lim * lim) {
        bool complete;
        const float radius = fminf(fminf(window, dist_upper(other.mesh.db, q)), cut_reach(far2, pad));
        float dist = sqrtf(min_d2_with(st, other.mesh, q, radius, PEN_STEPS, &complete));
        if (dist <= window) best = __float_as_uint(dist);
        hard = !complete;
      }
    }
    uint32_t slot = wave_slot(queued, hard);
    if (hard) queue[slot] = ((uint64_t)c << 33) | ((uint64_t)side << 32) | (uint32_t)i;
    best = wave_min(best);
    if (lane == 0 && best != 0xffffffffu) atomicMin(&acc.closest[side], best);
  }
}

// The k-th smallest key (0-based) of n by an MSD radix select over 8-bit digits, which stops as soon
// as the chosen bucket holds a single key (that key is then found by one more scan). *rank gets the
// k-th key's rank among the keys equal to it: 0 exactly when fewer than k + 1 keys are at most it
// and k keys are below it.
template <class K>
__device__ K select_kth(const K *keys, int n, int k, uint32_t *hist, K *s_prefix, int *s_k, uint32_t *s_cnt,
                        int *rank) {
  constexpr int PASSES = (int)sizeof(K);
  K prefix = 0, pmask = 0;
  if (threadIdx.x == 0) *s_k = k;
  for (int pass = 0; pass < PASSES; ++pass) {
    int shift = 8 * (PASSES - 1) - 8 * pass;
    for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      K key = keys[i];
      if ((key & pmask) == prefix) atomicAdd(&hist[(uint32_t)(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      // The digit whose bucket holds the k-th key: lane l owns bins 4l..4l+3; a wave prefix sum
      // over the lanes' totals finds the lane, then that lane walks its four bins.
      int lane = threadIdx.x;
      int kk = *s_k;
      uint32_t h[4] = {hist[4 * lane], hist[4 * lane + 1], hist[4 * lane + 2], hist[4 * lane + 3]};
      uint32_t tot = h[0] + h[1] + h[2] + h[3], incl = tot;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += v;
      }
      uint32_t excl = incl - tot;
      if ((uint32_t)kk >= excl && (uint32_t)kk < incl) {
        int r = kk - (int)excl, d = 0;
        while (r >= (int)h[d]) {  // ends within the four bins: r < tot
          r -= (int)h[d];
          ++d;
        }
        *s_k = r;
        *s_cnt = h[d];
        *s_prefix = prefix | ((K)(4 * lane + d) << shift);
      }
    }
    __syncthreads();
    prefix = *s_prefix;
    pmask |= ((K)255 << shift);
    const uint32_t cnt = *s_cnt;
    __syncthreads();
    if (cnt == 1u && pass + 1 < PASSES) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        K key = keys[i];
        if ((key & pmask) == prefix) *s_prefix = key;
      }
      __syncthreads();
      prefix = *s_prefix;
      break;
    }
  }
  *rank = *s_k;
  __syncthreads();
  return prefix;
}

// The closest-approach queries k_pen_close left unfinished, one per wave (min_d2_wave), taken from
// the queue until it is empty.
__global__ void __launch_bounds__(VT) k_pen_hard(VerifyBatch vb, int apart) {
  __shared__ int stk[VT / 64][HARD_STACK];
  int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t *queue = apart ? vb.hard_a : vb.hard;
  uint32_t *counts = vb.hard_n + (apart ? 3 : 0);  // [0] queued, [1] taken
  uint32_t n = counts[0];
  for (;;) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(&counts[1], 1u);
    idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
    if (idx >= n) break;
    uint64_t e = queue[idx];
    int c = (int)(e >> 33), side = (int)((e >> 32) & 1u), i = (int)(uint32_t)e;
    const CandDesc &cd = vb.cand[c];
    const PairDesc &pd = vb.pair[cd.pair];
    CandAcc &acc = vb.acc[c];
    const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
    const FragDesc &other = vb.frag[side == 0 ? pd.b : pd.a];
    const double *T = side == 0 ? cd.Ti : cd.T;
    const double *p = own.s + 3 * (size_t)i;
    V3 q = moved_f32(T, p[0], p[1], p[2]);
    float window = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(closest_window(acc, side))));
    // The window may have shrunk since the query was queued: one beyond it, by the grid's bound
    // or by every box of the cut (a lane each), is dropped unwalked.
    const float pad = slack_of(other.mesh.absmax, q), lim = window + pad;
    float bd2 = other.mesh.db.n_cut == 0 ? 0.0f : __builtin_inff(), fd2 = __builtin_inff();
    if (lane < other.mesh.db.n_cut) {
      const float *b = other.mesh.db.cut + 8 * lane;
      Node nd;
      for (int k = 0; k < 3; ++k) {
        nd.lo[k] = b[k];
        nd.hi[k] = b[4 + k];
      }
      bd2 = box_d2(nd, q, pad);
      const float f0 = fmaxf(q.x - nd.lo[0], nd.hi[0] - q.x), f1 = fmaxf(q.y - nd.lo[1], nd.hi[1] - q.y),
                  f2 = fmaxf(q.z - nd.lo[2], nd.hi[2] - q.z);
      fd2 = (f0 * f0 + f1 * f1) + f2 * f2;
    }
    const float cut_min = __uint_as_float(wave_min(__float_as_uint(bd2)));
    const float far2 = __uint_as_float(wave_min(__float_as_uint(fd2)));
    if (!(dist_bound(other.mesh.db, q) <= lim) || cut_min > lim * lim) continue;
    // The search needs to reach no farther than the point's own upper bounds (its distance is
    // within them: the grid's, and the cut's nearest far corner), which prunes from the first
    // node on.
    const float radius = fminf(fminf(window, dist_upper(other.mesh.db, q)), cut_reach(far2, pad));
    bool complete;
    float dist = sqrtf(min_d2_wave(other.mesh, q, radius, stk[w], HARD_STACK, WAVE_STEPS, &complete));
    if (lane == 0 && dist <= window) atomicMin(&acc.closest[side], __float_as_uint(dist));
    if (lane == 0 && !complete) atomicOr(&acc.flags, FLAG_GAVE_UP);
  }
}

// Rust's f64::midpoint.
__device__ double midpoint(double a, double b) {
  const double LO = 2.2250738585072014e-308 * 2.0;
  const double HI = 1.7976931348623157e308 / 2.0;
  double aa = fabs(a), ab = fabs(b);
  if (aa <= HI && ab <= HI) return (a + b) / 2.0;
  if (aa < LO) return a + (b / 2.0);
  if (ab < LO) return (a / 2.0) + b;
  return (a / 2.0) + (b / 2.0);
}

// verify.rs:693-700 numpy's median over each candidate's four segments.
// Statistics seg0 and seg0 + 1 of every candidate: the facing distances of R §6.1 (0, 1), or R
// §6.3's steps and agreements (2, 3).
constexpr int MEDIAN_VT = 512;
__global__ void __launch_bounds__(MEDIAN_VT) k_median(VerifyBatch vb, int seg0) {
  int c = blockIdx.x >> 1, seg = seg0 + (int)(blockIdx.x & 1);
  CandAcc &acc = vb.acc[c];
  int n = seg < 2 ? (int)acc.facing[seg] : (int)acc.n_cont;
  __shared__ uint32_t hist[256];
  __shared__ uint64_t s_prefix;
  __shared__ unsigned long long s_max;
  __shared__ int s_k;
  __shared__ uint32_t s_cnt;
  __shared__ uint64_t lds_keys[MEDIAN_LDS];
  if (n == 0) {
    if (threadIdx.x == 0) acc.med[seg] = __builtin_nan("");
    return;
  }
  const uint64_t *keys = vb.keys + vb.cand[c].seg[seg];
  if (n <= MEDIAN_LDS) {
    for (int i = threadIdx.x; i < n; i += MEDIAN_VT) lds_keys[i] = keys[i];
    __syncthreads();
    keys = lds_keys;
  }
  int rank;
  const uint64_t hi = select_kth(keys, n, n / 2, hist, &s_prefix, &s_k, &s_cnt, &rank);
  double med;
  if (n % 2 == 1) {
    med = from_total_key(hi);
  } else {
    // The lower middle key: the upper one when a key equal to it comes before it, else the
    // largest key below it (exactly n / 2 keys are below it then).
    uint64_t lo = hi;
    if (rank == 0) {
      if (threadIdx.x == 0) s_max = 0ull;
      __syncthreads();
      unsigned long long m = 0ull;
      for (int i = threadIdx.x; i < n; i += MEDIAN_VT) {
        const uint64_t key = keys[i];
        if (key < hi && key > m) m = key;
      }
      atomicMax(&s_max, m);
      __syncthreads();
      lo = s_max;
    }
    med = midpoint(from_total_key(lo), from_total_key(hi));
  }
  if (threadIdx.x == 0) acc.med[seg] = med;
}

// Scores (verify.rs:73-160, 590-626) and accept (verify.rs:632-638) per candidate.
__global__ void k_final(VerifyBatch vb) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= vb.n_cand) return;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  const FragDesc &A = vb.frag[pd.a];
  const FragDesc &B = vb.frag[pd.b];
  const CandAcc &acc = vb.acc[c];
  const double *sc = pd.sc;
  double side[2][3];
  const FragDesc *own[2] = {&A, &B};
  for (int s = 0; s < 2; ++s) {
    if (acc.facing[s] < 20) {  // MIN_FACING (verify.rs:61)
      side[s][0] = 0.0;
      side[s][1] = 1.0;
      side[s][2] = 0.0;
      continue;
    }
    side[s][0] = (double)acc.tight[s] / (double)acc.facing[s];
    side[s][1] = acc.med[s] / sc[S_T];
    double share = (double)acc.touching[s] / (double)own[s]->n_pf;
    side[s][2] = share * own[s]->frac_area / (sc[S_T] * sc[S_T]);
  }
  double *o = vb.scores + (size_t)VS_COUNT * c;
  for (int k = 0; k < VS_COUNT; ++k) o[k] = 0.0;
  o[VS_TIGHT_A] = side[0][0];
  o[VS_TIGHT_B] = side[1][0];
  o[VS_TIGHT] = fmin(side[0][0], side[1][0]);
  o[VS_GAP_A] = side[0][1];
  o[VS_GAP_B] = side[1][1];
  o[VS_GAP] = fmax(side[0][1], side[1][1]);
  o[VS_CONTACT_A] = side[0][2];
  o[VS_CONTACT_B] = side[1][2];
  o[VS_CONTACT] = fmin(side[0][2], side[1][2]);
  o[VS_SEAM] = acc.seam_cells == 0 ? 0.0 : (double)acc.seam_cells / 3.0;
  o[VS_GAP_LIMIT] = sc[S_GAP] / sc[S_T];
  o[VS_TIGHT_DELTA] = sc[S_TIGHT] / sc[S_T];
  o[VS_CONT] = 1.0;
  o[VS_CONT_N] = -1.0;
  if (!vb.full) {
    o[VS_PARTIAL] = 1.0;
  } else {
    if (A.n_margin > 0 && B.n_margin > 0 && acc.n_cont > 20) {  // MIN_NEAR (verify.rs:63)
      o[VS_CONT] = acc.med[2] / sc[S_T];
      o[VS_CONT_N] = acc.med[3];
    }
    if (!pd.pen_ok) {
      o[VS_PEN_UNAVAILABLE] = 1.0;
    } else {
      double pen[2], mn[2];
      for (int s = 0; s < 2; ++s) {
        int n = own[s]->n_s;
        if (n == 0) {
          pen[s] = 0.0;
          mn[s] = 0.0;
          continue;
        }
        pen[s] = (double)acc.deeper[s] / (double)n;
        mn[s] = acc.inside[s] > 0 ? -(double)__uint_as_float(acc.deepest[s]) : (double)__uint_as_float(acc.closest[s]);
      }
      o[VS_PEN] = fmax(pen[0], pen[1]);
      o[VS_PEN_DEPTH] = fmax(-mn[0], -mn[1]) / sc[S_T];
    }
  }
  bool ok = o[VS_TIGHT] >= 0.25 && o[VS_GAP] * sc[S_T] <= sc[S_GAP] && o[VS_PEN] <= 0.005 && o[VS_SEAM] >= 3.0 &&
            o[VS_CONT_N] >= 0.8;
  vb.accepted[c] = ok ? 1 : 0;
}

inline int blocks_for(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

hipError_t launch_verify(const VerifyBatch &vb, int max_pf, int max_s, int max_margin, int max_brk, hipStream_t s,
                         const SideStreams *ss, KernelProfiler *prof) {
  if (vb.n_cand <= 0) return hipSuccess;
  int nc = vb.n_cand;
  auto B = [&](const char *k, hipStream_t st) {
    if (prof && prof->on) prof->begin(k, st);
  };
  auto E = [&](hipStream_t st) {
    if (prof && prof->on) prof->end(st);
  };
  B("k_prep", s);
  hipLaunchKernelGGL(k_prep, dim3(blocks_for(nc, 64)), dim3(64), 0, s, vb);
  E(s);
  const bool pen = vb.full && max_s > 0;
  const int per_s = blocks_for(max_s, VT);
  const int per_sg = blocks_for(per_s, CHUNKS);  // workgroups per task of the per-sample passes
  // R §6.4's closest approach of the apart candidates needs nothing of k_pen: it starts at once on
  // the first side stream (ahead of R §6.1 there), and fills what the inside passes leave idle.
  const hipStream_t sa = ss ? ss->side[0] : s;
  // The seed's LDS: keys for the largest sample set's blocks (none past CLOSE_MAXB), the frontier.
  const int seed_nb = std::min(blocks_for(max_s, 64), CLOSE_MAXB);
  const size_t seed_lds = sizeof(uint32_t) * 2 * (size_t)seed_nb + sizeof(int) * SEED_FRONT;
  if (pen) {
    if (ss) {
      hipError_t e = hipEventRecord(ss->early, s);
      if (e == hipSuccess) e = hipStreamWaitEvent(sa, ss->early, 0);
      if (e != hipSuccess) return e;
    }
    B("k_pen_seed_apart", sa);
    hipLaunchKernelGGL(k_pen_seed, dim3(2 * nc), dim3(VT), seed_lds, sa, vb, 1, seed_nb);
    E(sa);
    B("k_pen_close_apart", sa);
    hipLaunchKernelGGL(k_pen_close, dim3(2 * nc * CLOSE_G), dim3(VT), 0, sa, vb, 1);
    E(sa);
    B("k_pen_hard_apart", sa);
    hipLaunchKernelGGL(k_pen_hard, dim3(HARD_BLOCKS), dim3(VT), 0, sa, vb, 1);
    E(sa);
  }
  if (pen) {
    // R §6.4's inside pass first and alone: the box reject, then the rays of the samples it kept.
    B("k_pen", s);
    hipLaunchKernelGGL(k_pen, dim3(2 * nc * per_sg), dim3(VT), 0, s, vb, per_sg);
    E(s);
    B("k_pen_groups", s);
    hipLaunchKernelGGL(k_pen_groups, dim3(1), dim3(GROUPS_VT), 0, s, vb);
    E(s);
    B("k_pen_rays", s);
    hipLaunchKernelGGL(k_pen_rays, dim3(RAY_BLOCKS), dim3(VT), 0, s, vb);
    E(s);
  }
  // The sections below read the poses k_prep completed and write disjoint accumulators and key
  // segments, so they may run side by side with the rest of R §6.4, whose closest-approach
  // passes leave most of the device idle; k_final waits for all of them.
  hipStream_t s1 = s, s2 = s, s3 = s;
  if (ss) {
    hipError_t e = hipEventRecord(ss->fork, s);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipStreamWaitEvent(ss->side[k], ss->fork, 0);
    if (e != hipSuccess) return e;
    s1 = ss->side[0];
    s2 = ss->side[1];
    s3 = ss->side[2];
  }
  if (max_pf > 0) {
    int per = blocks_for(blocks_for(max_pf, VT), CHUNKS);
    B("k_frac", s1);
    hipLaunchKernelGGL(k_frac, dim3(2 * nc * per), dim3(VT), 0, s1, vb, per);
    E(s1);
  }
  // Each statistic's medians right after the section that fills its keys, on that section's stream.
  B("k_median", s1);
  hipLaunchKernelGGL(k_median, dim3(2 * nc), dim3(MEDIAN_VT), 0, s1, vb, 0);
  E(s1);
  // The depths of the inside samples need only k_pen's list: ahead of R §6.2 on its stream.
  if (pen) {
    B("k_pen_depth", s2);
    hipLaunchKernelGGL(k_pen_depth, dim3(DEPTH_BLOCKS), dim3(DEPTH_VT), 0, s2, vb);
    E(s2);
  }
  B("k_seam_clear", s2);
  hipLaunchKernelGGL(k_seam_clear, dim3(nc), dim3(VT), 0, s2, vb);
  E(s2);
  if (max_brk > 0) {
    B("k_seam", s2);
    hipLaunchKernelGGL(k_seam, dim3(blocks_for(blocks_for(max_brk, VT), CHUNKS), nc), dim3(VT), 0, s2, vb);
    E(s2);
  }
  if (vb.full && max_margin > 0) {
    B("k_cont", s3);
    hipLaunchKernelGGL(k_cont, dim3(blocks_for(blocks_for(max_margin, VT), CHUNKS), nc), dim3(VT), 0, s3, vb);
    E(s3);
  }
  B("k_median", s3);
  hipLaunchKernelGGL(k_median, dim3(2 * nc), dim3(MEDIAN_VT), 0, s3, vb, 2);
  E(s3);
  if (pen) {
    B("k_pen_seed", s);
    hipLaunchKernelGGL(k_pen_seed, dim3(2 * nc), dim3(VT), seed_lds, s, vb, 0, seed_nb);
    E(s);
    B("k_pen_close", s);
    hipLaunchKernelGGL(k_pen_close, dim3(2 * nc * CLOSE_G), dim3(VT), 0, s, vb, 0);
    E(s);
    B("k_pen_hard", s);
    hipLaunchKernelGGL(k_pen_hard, dim3(HARD_BLOCKS), dim3(VT), 0, s, vb, 0);
    E(s);
  }
  if (ss) {
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipEventRecord(ss->join[k], ss->side[k]);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipStreamWaitEvent(s, ss->join[k], 0);
    if (e != hipSuccess) return e;
  }
  B("k_final", s);
  hipLaunchKernelGGL(k_final, dim3(blocks_for(nc, 64)), dim3(64), 0, s, vb);
  E(s);
  return hipGetLastError();
}

}  // namespace sherd
