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
//   k_prep     pose inverse per candidate (verify.rs:739-798), counters cleared
//   k_frac     R §6.1: both sides' fracture samples against the other's fracture scene
//   k_seam     R §6.2: A's breakline against the moved B breakline (through B's breakline grid),
//              voxel set of the seam points
//   k_cont     R §6.3: B's margin against A's margin grid, steps and normal agreements
//   k_pen      R §6.4: surface samples inside the other mesh, deeper-than-pen count, deepest
//   k_pen_lb   R §6.4's fallback when nothing is inside: each side's sample of smallest bound
//   k_pen_seed the exact distance of that sample, one wave each: a first attained window
//   k_pen_min  the closest approach
//   k_pen_hard the closest-approach queries too long for one lane, one wave each
//   k_median   the four medians (numpy semantics, verify.rs:693-700)
//   k_final    Scores and the accept verdict
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
// k_pen's depth searches start from an inside point with a window as wide as the nearest ray
// crossing, so they defer more nodes than the windowed searches: a deeper LDS stack (27 KB of LDS
// per workgroup, which keeps the five waves per SIMD its registers allow).
constexpr int PEN_LSTACK = 12;

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

struct TaskPos {
  int c, side, i;
};
__device__ __forceinline__ TaskPos task_pos(const VerifyBatch &vb, int per_task) {
  uint32_t lb = xcd_logical(blockIdx.x, gridDim.x);
  uint32_t cs = vb.task[lb / (uint32_t)per_task];
  return TaskPos{(int)(cs >> 1), (int)(cs & 1u), (int)((lb % (uint32_t)per_task) * VT + threadIdx.x)};
}

__global__ void k_prep(VerifyBatch vb) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= vb.n_cand) return;
  CandDesc &cd = vb.cand[c];
  pose_inverse(cd.T, cd.Ti);
  CandAcc &a = vb.acc[c];
  for (int s = 0; s < 2; ++s) {
    a.facing[s] = a.tight[s] = a.touching[s] = 0;
    a.inside[s] = a.deeper[s] = 0;
    a.in_box[s] = 0;
    a.deepest[s] = 0u;
    a.closest[s] = 0x7f7fffffu;  // f32::MAX, the window of the fallback (verify.rs:570-577)
  }
  a.n_cont = 0;
  if (c == 0) vb.hard_n[0] = vb.hard_n[1] = 0;
  a.seam_cells = 0;
  a.flags = 0;
  a.seed[0] = a.seed[1] = ~0ull;
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

// R §6.4 one_penetration's first pass (verify.rs:541-566): side 0 is A's surface samples moved by
// T^-1 against B's mesh, side 1 is B's moved by T against A's mesh.
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
  // Most samples fall outside the other mesh's box (RayScene::inside's reject). Per chunk, the
  // workgroup gathers the ones inside it in LDS and its first lanes take them, one each, so the
  // rays and depth searches run in as few waves as there are such samples.
  __shared__ float s_q[3][VT];
  __shared__ uint32_t s_n;
  __shared__ uint2 s_stk[PEN_LSTACK][VT];
  CandAcc &acc = vb.acc[c];
  LdsStack<VT, PEN_LSTACK> st{&s_stk[0][threadIdx.x]};
  for (int k = 0; k < CHUNKS; ++k) {
    const int base = (g.chunk0 + k) * VT;
    if (base >= own.n_s) break;  // (uniform)
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int i = base + threadIdx.x, i0 = i & ~63;
    const bool skip = i0 >= own.n_s || block_box_gap(moved_block(own.s_blk, i0, T), sv.lo, sv.hi, sv.absmax) > 0.0;
    if (i < own.n_s && !skip) {
      const double *p = own.s + 3 * (size_t)i;
      V3 q = moved_f32(T, p[0], p[1], p[2]);
      if (!(q.x < sv.lo[0] || q.x > sv.hi[0] || q.y < sv.lo[1] || q.y > sv.hi[1] || q.z < sv.lo[2] ||
            q.z > sv.hi[2])) {
        uint32_t s = atomicAdd(&s_n, 1u);
        s_q[0][s] = q.x;
        s_q[1][s] = q.y;
        s_q[2][s] = q.z;
      }
    }
    __syncthreads();
    const uint32_t n = s_n;
    if (threadIdx.x == 0 && n) atomicAdd(&acc.in_box[side], n);
    if (threadIdx.x < ((n + 63u) & ~63u)) {  // whole waves only
      bool in = false, deeper = false;
      float depth = 0.0f;
      if (threadIdx.x < n) {
        V3 q = V3{s_q[0][threadIdx.x], s_q[1][threadIdx.x], s_q[2][threadIdx.x]};
        in = inside_depth_with(st, sv, q, vb.retest, &depth);
        deeper = in && (double)depth > pd.sc[S_PEN];
      }
      uint32_t deepest = wave_max(in ? __float_as_uint(depth) : 0u);
      uint32_t ni = (uint32_t)__popcll(__ballot(in)), nd = (uint32_t)__popcll(__ballot(deeper));
      if ((threadIdx.x & 63) == 0) {
        if (ni) atomicAdd(&acc.inside[side], ni);
        if (nd) atomicAdd(&acc.deeper[side], nd);
        if (deepest) atomicMax(&acc.deepest[side], deepest);
      }
    }
    __syncthreads();  // the gathered samples are read before the next chunk writes them
  }
}

// Whether side `side`'s closest approach can change the scores. pen_depth is
// max(-min_a, -min_b) / t (verify.rs:494), min being -deepest on a side with a sample inside and the
// closest approach (at least 0) on a side without: when the other side has a sample inside at a
// positive depth, this side's term is at most 0 < that depth and is never read (at depth 0 exactly
// it still is, for the sign of a zero).
__device__ __forceinline__ bool closest_needed(const CandAcc &acc, int side) {
  return acc.inside[side] == 0 && !(acc.inside[1 - side] != 0 && acc.deepest[1 - side] != 0u);
}

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

// The seed's first half: each side's 64-sample block of smallest bound (the distance bound at the
// moved centre less the radius), (bound bits << 32) | block, kept by an atomic minimum. One lane
// per block.
__global__ void __launch_bounds__(VT) k_pen_lb(VerifyBatch vb, int per_task) {
  TaskPos tp = task_pos(vb, per_task);
  const int c = tp.c, side = tp.side, b = tp.i;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  if (!pd.pen_ok) return;
  CandAcc &acc = vb.acc[c];
  if (!closest_needed(acc, side)) return;
  const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
  const FragDesc &other = vb.frag[side == 0 ? pd.b : pd.a];
  uint64_t key = ~0ull;
  if (64 * (size_t)b < (size_t)own.n_s) {
    const double *T = side == 0 ? cd.Ti : cd.T;
    key = ((uint64_t)__float_as_uint(block_bound(other.mesh, moved_block(own.s_blk, 64 * b, T))) << 32) | (uint32_t)b;
  }
  const uint32_t hi = wave_min((uint32_t)(key >> 32));
  const uint32_t lo = wave_min((uint32_t)(key >> 32) == hi ? (uint32_t)key : 0xffffffffu);
  if ((threadIdx.x & 63) == 0 && hi != 0xffffffffu)
    atomicMin((unsigned long long *)&acc.seed[side], ((unsigned long long)hi << 32) | lo);
}

// The exact pass, one lane per sample: a sample whose distance bound already exceeds the window (by
// more than the slack, which covers the rounding of the moved point) cannot lower the minimum; the
// others search (min_d2). A query that needs more than PEN_STEPS nodes is finished by a whole wave
// (k_pen_hard); what it met so far is a distance it attains, so it may still lower the window.
__global__ void __launch_bounds__(VT) k_pen_min(VerifyBatch vb, int per_task) {
  __shared__ uint2 s_stk[LSTACK][VT];
  const TaskGroup g = task_group(vb, per_task);
  const int c = g.c, side = g.side;
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  if (!pd.pen_ok) return;
  CandAcc &acc = vb.acc[c];
  if (!closest_needed(acc, side)) return;
  const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
  const FragDesc &other = vb.frag[side == 0 ? pd.b : pd.a];
  const double *T = side == 0 ? cd.Ti : cd.T;
  if (own.n_s == 0) return;
  LdsStack<VT> st{&s_stk[0][threadIdx.x]};
  for (int k = 0; k < CHUNKS; ++k) {
    const int base = (g.chunk0 + k) * VT;
    if (base >= own.n_s) break;  // (uniform)
    const int i = base + threadIdx.x, i0 = i & ~63;
    if (i0 >= own.n_s) break;
    // A block every sample of which is farther than the window (read once for the wave) cannot
    // lower the minimum: the wave goes on to its next block. Uniform across the wave.
    const float w0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(closest_window(acc, side))));
    if (block_bound(other.mesh, moved_block(own.s_blk, i0, T)) > w0) continue;
    uint32_t best = 0xffffffffu;
    bool hard = false;
    if (i < own.n_s) {
      const double *p = own.s + 3 * (size_t)i;
      V3 q = moved_f32(T, p[0], p[1], p[2]);
      float window = closest_window(acc, side);
      if (dist_bound(other.mesh.db, q) <= window + slack_of(other.mesh.absmax, q) && window > 0.0f) {
        bool complete;
        float dist = sqrtf(min_d2_with(st, other.mesh, q, fminf(window, dist_upper(other.mesh.db, q)), PEN_STEPS,
                                       &complete));
        if (dist <= window) best = __float_as_uint(dist);
        hard = !complete;
      }
    }
    uint32_t slot = wave_slot(&vb.hard_n[0], hard);
    if (hard) vb.hard[slot] = ((uint64_t)c << 33) | ((uint64_t)side << 32) | (uint32_t)i;
    best = wave_min(best);
    if ((threadIdx.x & 63) == 0 && best != 0xffffffffu) atomicMin(&acc.closest[side], best);
  }
}

// The k-th smallest key (0-based) of n by an 8-pass MSD radix select.
template <class K>
__device__ K select_kth(const K *keys, int n, int k, uint32_t *hist, K *s_prefix, int *s_k) {
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
      uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
      uint32_t tot = h0 + h1 + h2 + h3, incl = tot;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += v;
      }
      uint32_t excl = incl - tot;
      if ((uint32_t)kk >= excl && (uint32_t)kk < incl) {
        int r = kk - (int)excl, d = 4 * lane;
        if (r >= (int)h0) {
          r -= (int)h0;
          ++d;
          if (r >= (int)h1) {
            r -= (int)h1;
            ++d;
            if (r >= (int)h2) {
              r -= (int)h2;
              ++d;
            }
          }
        }
        *s_k = r;
        *s_prefix = prefix | ((K)d << shift);
      }
    }
    __syncthreads();
    prefix = *s_prefix;
    pmask |= ((K)255 << shift);
    __syncthreads();
  }
  return prefix;
}

// The closest approach is a minimum over every sample, and a search is cheap once a small
// window is known. Before the exact pass, each task's sample of smallest distance bound is found
// (k_pen_lb over blocks, then k_pen_seed over the samples of the blocks that can hold it), and the
// same wave measures that sample's exact distance (min_d2_wave): a distance the samples really attain, so at least the true minimum,
// and, the bound being within a couple of grid cells of the distance, close to it. The minimum that
// comes out is the same whichever sample seeds it; only the work changes.
__global__ void __launch_bounds__(VT) k_pen_seed(VerifyBatch vb) {
  __shared__ int stk[VT / 64][HARD_STACK];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = (int)blockIdx.x * (VT / 64) + w;
  if (t >= 2 * vb.n_cand) return;  // (uniform across the wave; the kernel has no block barrier)
  const uint32_t cs = vb.task[t];
  const int c = (int)(cs >> 1), side = (int)(cs & 1u);
  const CandDesc &cd = vb.cand[c];
  const PairDesc &pd = vb.pair[cd.pair];
  if (!pd.pen_ok) return;
  CandAcc &acc = vb.acc[c];
  if (!closest_needed(acc, side)) return;
  const FragDesc &own = vb.frag[side == 0 ? pd.a : pd.b];
  const FragDesc &other = vb.frag[side == 0 ? pd.b : pd.a];
  const double *T = side == 0 ? cd.Ti : cd.T;
  const uint64_t key = acc.seed[side];
  if (key == ~0ull) return;
  // The task's sample of smallest distance bound, (bound bits, index) least: first over the block
  // of smallest block bound, then over every other block whose block bound (less the bound grid's
  // cell diagonal, by which a sample's bound may fall below its block's) does not exceed the best
  // so far; the blocks are tested 64 at a time, each qualifying one by the whole wave.
  const int bstar = (int)(uint32_t)key;
  const int nb = (own.n_s + 63) >> 6;
  const float cell = 1.8f * other.mesh.db.h;
  uint32_t khi, seed;
  {
    const int i = 64 * bstar + lane;
    uint64_t k2 = ~0ull;
    if (i < own.n_s) {
      const double *pi = own.s + 3 * (size_t)i;
      k2 = ((uint64_t)__float_as_uint(dist_bound(other.mesh.db, moved_f32(T, pi[0], pi[1], pi[2]))) << 32) | (uint32_t)i;
    }
    khi = wave_min((uint32_t)(k2 >> 32));
    seed = wave_min((uint32_t)(k2 >> 32) == khi ? (uint32_t)k2 : 0xffffffffu);
  }
  if (seed == 0xffffffffu) return;  // (uniform)
  for (int base = 0; base < nb; base += 64) {
    const int b = base + lane;
    const bool take = b < nb && b != bstar &&
                      block_bound(other.mesh, moved_block(own.s_blk, 64 * b, T)) - cell <= __uint_as_float(khi);
    uint64_t m = __ballot(take);
    while (m) {
      const int bb = base + __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int i = 64 * bb + lane;
      uint64_t k2 = ~0ull;
      if (i < own.n_s) {
        const double *pi = own.s + 3 * (size_t)i;
        k2 = ((uint64_t)__float_as_uint(dist_bound(other.mesh.db, moved_f32(T, pi[0], pi[1], pi[2]))) << 32) | (uint32_t)i;
      }
      const uint64_t cur = ((uint64_t)khi << 32) | seed;
      k2 = k2 < cur ? k2 : cur;
      khi = wave_min((uint32_t)(k2 >> 32));
      seed = wave_min((uint32_t)(k2 >> 32) == khi ? (uint32_t)k2 : 0xffffffffu);
    }
  }
  const double *p = own.s + 3 * (size_t)seed;
  V3 q = moved_f32(T, p[0], p[1], p[2]);
  float window = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(closest_window(acc, side))));
  // Searched no farther than the seed's upper bound: the exact distance is within it.
  const float radius = fminf(window, dist_upper(other.mesh.db, q));
  bool complete;
  float dist = sqrtf(min_d2_wave(other.mesh, q, radius, stk[w], HARD_STACK, WAVE_STEPS, &complete));
  if (lane == 0 && dist <= window) atomicMin(&acc.closest[side], __float_as_uint(dist));
  if (lane == 0 && !complete) atomicOr(&acc.flags, FLAG_GAVE_UP);
}

// The closest-approach queries k_pen_min left unfinished, one per wave (min_d2_wave), taken from
// the queue until it is empty.
__global__ void __launch_bounds__(VT) k_pen_hard(VerifyBatch vb) {
  __shared__ int stk[VT / 64][HARD_STACK];
  int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t n = vb.hard_n[0];
  for (;;) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(&vb.hard_n[1], 1u);
    idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
    if (idx >= n) break;
    uint64_t e = vb.hard[idx];
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
    // The window may have shrunk since the query was queued: one beyond it is dropped unwalked.
    if (!(dist_bound(other.mesh.db, q) <= window + slack_of(other.mesh.absmax, q))) continue;
    // The search needs to reach no farther than the point's own upper bound (its distance is
    // within it), which prunes from the first node on.
    const float radius = fminf(window, dist_upper(other.mesh.db, q));
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
__global__ void __launch_bounds__(VT) k_median(VerifyBatch vb) {
  int c = blockIdx.x >> 2, seg = blockIdx.x & 3;
  CandAcc &acc = vb.acc[c];
  int n = seg < 2 ? (int)acc.facing[seg] : (int)acc.n_cont;
  __shared__ uint32_t hist[256];
  __shared__ uint64_t s_prefix;
  __shared__ int s_k;
  __shared__ uint64_t lds_keys[MEDIAN_LDS];
  if (n == 0) {
    if (threadIdx.x == 0) acc.med[seg] = __builtin_nan("");
    return;
  }
  const uint64_t *keys = vb.keys + vb.cand[c].seg[seg];
  if (n <= MEDIAN_LDS) {
    for (int i = threadIdx.x; i < n; i += VT) lds_keys[i] = keys[i];
    __syncthreads();
    keys = lds_keys;
  }
  double med;
  if (n % 2 == 1) {
    med = from_total_key(select_kth(keys, n, n / 2, hist, &s_prefix, &s_k));
  } else {
    double lo = from_total_key(select_kth(keys, n, n / 2 - 1, hist, &s_prefix, &s_k));
    double hi = from_total_key(select_kth(keys, n, n / 2, hist, &s_prefix, &s_k));
    med = midpoint(lo, hi);
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
  if (pen) {
    // R §6.4's inside pass first and alone: it is the longest kernel and fills the device.
    B("k_pen", s);
    hipLaunchKernelGGL(k_pen, dim3(2 * nc * per_sg), dim3(VT), 0, s, vb, per_sg);
    E(s);
  }
  // The sections below read the poses k_prep completed and write disjoint accumulators and key
  // segments, so they may run side by side with the rest of R §6.4, whose closest-approach
  // passes leave most of the device idle; k_median waits for all of them.
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
  if (pen) {
    B("k_pen_lb", s);
    const int per_b = blocks_for(blocks_for(max_s, 64), VT);
    hipLaunchKernelGGL(k_pen_lb, dim3(2 * nc * per_b), dim3(VT), 0, s, vb, per_b);
    E(s);
    B("k_pen_seed", s);
    hipLaunchKernelGGL(k_pen_seed, dim3(blocks_for(2L * nc, VT / 64)), dim3(VT), 0, s, vb);
    E(s);
    B("k_pen_min", s);
    hipLaunchKernelGGL(k_pen_min, dim3(2 * nc * per_sg), dim3(VT), 0, s, vb, per_sg);
    E(s);
    B("k_pen_hard", s);
    hipLaunchKernelGGL(k_pen_hard, dim3(HARD_BLOCKS), dim3(VT), 0, s, vb);
    E(s);
  }
  if (ss) {
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipEventRecord(ss->join[k], ss->side[k]);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipStreamWaitEvent(s, ss->join[k], 0);
    if (e != hipSuccess) return e;
  }
  B("k_median", s);
  hipLaunchKernelGGL(k_median, dim3(4 * nc), dim3(VT), 0, s, vb);
  E(s);
  B("k_final", s);
  hipLaunchKernelGGL(k_final, dim3(blocks_for(nc, 64)), dim3(64), 0, s, vb);
  E(s);
  return hipGetLastError();
}

}  // namespace sherd
