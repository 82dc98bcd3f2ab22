// R §3.4.4-3.4.8 on the device: the clean-up of the cone vote's mask (crates/sherd-core/src/
// fragment/segment.rs:403-480 majority_filter and refine_boundary, :544-635 the steps in order;
// mesh/adjacency.rs:138-158 face_adjacency; mesh/components.rs:120-205 drop_small_components).
//
//   k_seg_geom        f64 face normals, areas and centroids (mesh/geometry.rs:70-96), face indices
//                     checked
//   k_seg_edge_keys   the 3m edge rows slot-major, keyed (lo, hi); a stable radix sort; then
//   k_seg_pairs       every two consecutive rows with one key are an adjacent pair
//   k_seg_ball<S>     one wave per voxel representative: the centroids within r through the f64
//                     neighbour grid (kiddo's (dx^2 + dy^2) + dz^2 <= r^2), put in ascending index
//                     order by rank, then summed in that order as PointTree::within + the
//                     reference's loop do: the majority's area sums (S = false), or the shell's
//                     area-weighted normal (S = true, shell faces only; every face when no mask
//                     is given: R §3.4.2's smoothed normals)
//   k_cc_*            components of mask == target: hooking of roots over the pairs and pointer
//                     jumping until nothing changes; each component's exact area (fixed point,
//                     four 32-bit digits summed in 64-bit counters, a wave of one component adding
//                     once); the faces of a component below the bound flipped
//   k_grow_*          the growth passes: candidates (fracture faces next to a shell face) marked
//                     from the mask as the pass starts, then the eligible ones flipped
#include <hip/hip_runtime.h>

#include "grid.hpp"
#include "kernels.hpp"

namespace sherd {
namespace {

constexpr int SEG_WAVES = 4;

__global__ void k_seg_geom(const double *V, int nv, const uint32_t *F, int nf, double *N, double *A, double *C,
                           uint32_t *bad) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const uint32_t i0 = F[3 * (size_t)f], i1 = F[3 * (size_t)f + 1], i2 = F[3 * (size_t)f + 2];
  double *n = N + 3 * (size_t)f, *c = C + 3 * (size_t)f;
  if (i0 >= (uint32_t)nv || i1 >= (uint32_t)nv || i2 >= (uint32_t)nv) {
    n[0] = n[1] = n[2] = 0.0;
    c[0] = c[1] = c[2] = 0.0;
    A[f] = 0.0;
    atomicOr(bad, 1u);
    return;
  }
  const double *a = V + 3 * (size_t)i0, *b = V + 3 * (size_t)i1, *d = V + 3 * (size_t)i2;
  const double e1x = b[0] - a[0], e1y = b[1] - a[1], e1z = b[2] - a[2];
  const double e2x = d[0] - a[0], e2y = d[1] - a[1], e2z = d[2] - a[2];
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double len = sqrt((nx * nx + ny * ny) + nz * nz);
  const double den = fmax(len, 1e-12);
  n[0] = nx / den;
  n[1] = ny / den;
  n[2] = nz / den;
  A[f] = 0.5 * len;
  c[0] = (a[0] + b[0] + d[0]) / 3.0;
  c[1] = (a[1] + b[1] + d[1]) / 3.0;
  c[2] = (a[2] + b[2] + d[2]) / 3.0;
}

// Row i = slot * nf + f holds edge (F[f][slot], F[f][(slot + 1) % 3]), keyed lo * nv + hi.
__global__ void k_seg_edge_keys(const uint32_t *F, int nf, uint32_t nv, uint64_t *key, uint32_t *pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nf) return;
  const int slot = i / nf, f = i - slot * nf;
  const uint32_t a = F[3 * (size_t)f + slot], b = F[3 * (size_t)f + (slot + 1) % 3];
  const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  key[i] = (uint64_t)lo * nv + hi;
  pos[i] = (uint32_t)i;
}

__global__ void k_seg_pairs(const uint64_t *key, const uint32_t *pos, int m, int nf, uint32_t *fa, uint32_t *fb,
                            uint32_t *count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool take = false;
  uint32_t a = 0, b = 0;
  if (i + 1 < m && key[i] == key[i + 1]) {
    take = true;
    a = pos[i] % (uint32_t)nf;
    b = pos[i + 1] % (uint32_t)nf;
  }
  uint32_t s = wave_slot(count, take);
  if (take) {
    fa[s] = a;
    fb[s] = b;
  }
}

// The table slot of an occupied cell, or count 0.
__device__ __forceinline__ GridSlot find_cell(const GridView &g, int64_t x, int64_t y, int64_t z) {
  GridSlot none{GRID_EMPTY, 0u, 0u};
  if (x < 0 || y < 0 || z < 0 || x >= GRID_MAXC || y >= GRID_MAXC || z >= GRID_MAXC) return none;
  const uint64_t want = grid_pack(x, y, z);
  uint32_t slot = grid_hash(want, g.mask);
  GridSlot s = load_slot(g.table, slot);
  for (uint32_t probe = 0; s.key != want && s.key != GRID_EMPTY && probe < g.mask; ++probe) {
    slot = (slot + 1) & g.mask;
    s = load_slot(g.table, slot);
  }
  return s.key == want ? s : none;
}

// Lane k's value, read by every lane (k uniform across the wave).
__device__ __forceinline__ double lane_read(double x, uint32_t k) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)k);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint32_t lane_read_u32(uint32_t x, uint32_t k) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)k);
}

// One wave per representative q (face rep[q]).
template <bool SHELL>
__global__ void __launch_bounds__(64 * SEG_WAVES)
    k_seg_ball(GridView g, const double *C, const double *A, const double *N, const uint8_t *frac, int nf,
               const uint32_t *rep, int n_rep, double r, uint8_t *maj, double *ref, uint8_t *has, uint32_t *err) {
  __shared__ uint32_t list[SEG_WAVES][SEG_BALL_CAP];
  __shared__ uint32_t sorted[SEG_WAVES][SEG_BALL_CAP];
  __shared__ uint32_t cnt[SEG_WAVES];
  __shared__ uint32_t runs[SEG_WAVES][2][32];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = blockIdx.x * SEG_WAVES + wv;
  uint32_t f = 0;
  bool live = q < n_rep;
  if (live) {
    f = rep[q];
    if (f >= (uint32_t)nf) {
      if (lane == 0) atomicOr(err, 1u);
      live = false;
    }
  }
  // Lanes 0..26 find the 27 cells' runs of points; then the runs, laid end to end, are tested 64
  // points at a time across the whole wave.
  if (lane == 0) cnt[wv] = 0;
  if (lane < 32) runs[wv][0][lane] = runs[wv][1][lane] = 0;
  double qx = 0.0, qy = 0.0, qz = 0.0;
  if (live) {
    qx = C[3 * (size_t)f];
    qy = C[3 * (size_t)f + 1];
    qz = C[3 * (size_t)f + 2];
  }
  if (live && lane < 27) {
    const int64_t cx = (int64_t)floor((qx - g.origin[0]) / g.cell) + (lane / 9 - 1);
    const int64_t cy = (int64_t)floor((qy - g.origin[1]) / g.cell) + ((lane / 3) % 3 - 1);
    const int64_t cz = (int64_t)floor((qz - g.origin[2]) / g.cell) + (lane % 3 - 1);
    const GridSlot s = find_cell(g, cx, cy, cz);
    runs[wv][0][lane] = s.count;
    runs[wv][1][lane] = s.start;
  }
  __syncthreads();
  if (live) {
    const double r2 = r * r;
    uint32_t total = 0;
    for (int k = 0; k < 27; ++k) total += runs[wv][0][k];
    uint32_t k = 0, before = 0;  // the run holding point idx and the points ahead of it
    for (uint32_t idx = lane; idx < total; idx += 64) {
      while (idx >= before + runs[wv][0][k]) before += runs[wv][0][k++];
      const uint32_t at = runs[wv][1][k] + (idx - before);
      const double *p = g.pts + 3 * (size_t)at;
      const double d2 = d2_f64(qx, qy, qz, gldd(p), gldd(p + 1), gldd(p + 2));
      if (!(d2 <= r2)) continue;
      const uint32_t j = gldu(g.idx + at);
      if (SHELL && frac && frac[j]) continue;
      const uint32_t slot = atomicAdd(&cnt[wv], 1u);
      if (slot < SEG_BALL_CAP) list[wv][slot] = j;
    }
  }
  __syncthreads();
  const uint32_t n = live ? cnt[wv] : 0u;
  const bool over = n > (uint32_t)SEG_BALL_CAP;
  if (live && over && lane == 0) atomicOr(err, 2u);
  if (live && !over) {
    // Distinct indices: an element's rank is how many are smaller.
    for (uint32_t i = lane; i < n; i += 64) {
      const uint32_t e = list[wv][i];
      uint32_t rank = 0;
      for (uint32_t k = 0; k < n; ++k) rank += list[wv][k] < e;
      sorted[wv][rank] = e;
    }
  }
  __syncthreads();
  if (!live || over) return;
  // The sums in ascending index order, 64 faces at a time: every lane fetches one face's terms
  // (one memory round trip for 64 faces), then the whole wave adds them in order from lane reads,
  // so each lane ends with the same sums the reference's sequential loop makes.
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, sw = 0.0;
  for (uint32_t b = 0; b < n; b += 64) {
    const uint32_t i = b + lane;
    double a = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
    uint32_t wet = 0;
    if (i < n) {
      const uint32_t j = sorted[wv][i];
      a = A[j];
      if (SHELL) {
        p0 = N[3 * (size_t)j] * a;
        p1 = N[3 * (size_t)j + 1] * a;
        p2 = N[3 * (size_t)j + 2] * a;
      } else {
        wet = frac[j] ? 1u : 0u;
      }
    }
    const uint32_t cnt = n - b < 64u ? n - b : 64u;
    for (uint32_t k = 0; k < cnt; ++k) {
      const double ak = lane_read(a, k);
      sw += ak;
      if (SHELL) {
        s0 += lane_read(p0, k);
        s1 += lane_read(p1, k);
        s2 += lane_read(p2, k);
      } else if (lane_read_u32(wet, k)) {
        s0 += ak;
      }
    }
  }
  if (lane != 0) return;
  if (!SHELL) {
    maj[q] = s0 > 0.5 * sw ? 1 : 0;  // wet > 0.5 total
  } else {
    const double den = fmax(sw, 1e-12);
    const double m0 = s0 / den, m1 = s1 / den, m2 = s2 / den;
    const double norm = fmax(sqrt((m0 * m0 + m1 * m1) + m2 * m2), 1e-12);
    ref[3 * (size_t)q] = m0 / norm;
    ref[3 * (size_t)q + 1] = m1 / norm;
    ref[3 * (size_t)q + 2] = m2 / norm;
    has[q] = n > 0 ? 1 : 0;
  }
}

// A face takes its nearest representative's answer.
__global__ void k_seg_take_majority(const uint32_t *near, int nf, int n_rep, const uint8_t *maj, uint8_t *frac,
                                    uint32_t *err) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const uint32_t r = near[f];
  if (r >= (uint32_t)n_rep) {
    atomicOr(err, 4u);
    frac[f] = 0;
    return;
  }
  frac[f] = maj[r];
}

// The fixed reference normal of each face and whether the growth may take it
// (segment.rs:465-470: has_ref and n . ref > cos).
__global__ void k_seg_take_reference(const uint32_t *near, int nf, int n_rep, const double *ref_g, const uint8_t *has_g,
                                     const double *N, double cos_grow, double *ref, uint8_t *has, uint8_t *elig) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const uint32_t r = near[f];
  if (r >= (uint32_t)n_rep) return;  // already reported by k_seg_take_majority
  const double r0 = ref_g[3 * (size_t)r], r1 = ref_g[3 * (size_t)r + 1], r2 = ref_g[3 * (size_t)r + 2];
  ref[3 * (size_t)f] = r0;
  ref[3 * (size_t)f + 1] = r1;
  ref[3 * (size_t)f + 2] = r2;
  has[f] = has_g[r];
  const double *n = N + 3 * (size_t)f;
  elig[f] = (has_g[r] && (n[0] * r0 + n[1] * r1) + n[2] * r2 > cos_grow) ? 1 : 0;
}

// A face's smoothed normal is its representative's ball normal (segment.rs:545-546).
__global__ void k_seg_take_normal(const uint32_t *near, int nf, int n_rep, const double *ref_g, double *ns,
                                  uint32_t *err) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const uint32_t r = near[f];
  if (r >= (uint32_t)n_rep) {
    atomicOr(err, 4u);
    return;
  }
  ns[3 * (size_t)f] = ref_g[3 * (size_t)r];
  ns[3 * (size_t)f + 1] = ref_g[3 * (size_t)r + 1];
  ns[3 * (size_t)f + 2] = ref_g[3 * (size_t)r + 2];
}

// ---- components -------------------------------------------------------------------------------

__global__ void k_cc_init(int n, uint32_t *parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = (uint32_t)i;
}

// Parents only ever point lower, so a walk up ends at a root even while others write.
__device__ __forceinline__ uint32_t cc_root(const uint32_t *parent, uint32_t x) {
  uint32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (p != x) {
    x = p;
    p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return x;
}

// The higher of two different roots is hooked under the lower. A hook that lands on a node no
// longer a root can drop an earlier hook; the pass is repeated until no pair joins two roots.
__global__ void k_cc_hook(const uint32_t *fa, const uint32_t *fb, int np, const uint8_t *mask, uint8_t target,
                          uint32_t *parent, uint32_t *changed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t a = fa[i], b = fb[i];
  if (mask[a] != target || mask[b] != target) return;
  const uint32_t ra = cc_root(parent, a), rb = cc_root(parent, b);
  if (ra == rb) return;
  const uint32_t lo = ra < rb ? ra : rb, hi = ra < rb ? rb : ra;
  atomicMin(&parent[hi], lo);
  *changed = 1u;
}

__global__ void k_cc_compress(int n, uint32_t *parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = cc_root(parent, (uint32_t)i);
  if (r != (uint32_t)i) atomicMin(&parent[i], r);
}

// floor(a 2^sh) as (hi, lo) 64-bit halves; exact for every a with a 2^sh < 2^117.
__device__ __forceinline__ void fixed_of(double a, int sh, uint64_t *hi, uint64_t *lo) {
  const double x = ldexp(a, sh);
  const double h = floor(x * 0x1p-64);
  *hi = (uint64_t)h;
  *lo = (uint64_t)(x - h * 0x1p64);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Each component's area as four 32-bit digits of its fixed-point sum, each digit column summed in
// a 64-bit counter (acc[4 r + k]); exact, so the order of the additions does not matter.
__global__ void k_cc_area(int n, const uint8_t *mask, uint8_t target, const uint32_t *parent, const double *A,
                          int sh, unsigned long long *acc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n && mask[i] == target;
  uint32_t root = 0;
  uint64_t d[4] = {0, 0, 0, 0};
  if (valid) {
    root = parent[i];
    uint64_t hi, lo;
    fixed_of(A[i], sh, &hi, &lo);
    d[0] = lo & 0xFFFFFFFFull;
    d[1] = lo >> 32;
    d[2] = hi & 0xFFFFFFFFull;
    d[3] = hi >> 32;
  }
  const uint64_t act = __ballot(valid);
  if (act == 0) return;
  const int lead = __ffsll((unsigned long long)act) - 1;
  const uint32_t r0 = (uint32_t)__shfl((int)root, lead);
  if (__ballot(valid && root == r0) == act) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t s = wave_sum_u64(d[k]);
      if ((int)(threadIdx.x & 63) == lead && s) atomicAdd(&acc[4 * (size_t)r0 + k], (unsigned long long)s);
    }
  } else if (valid) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (d[k]) atomicAdd(&acc[4 * (size_t)root + k], (unsigned long long)d[k]);
  }
}

// Faces of a component whose exact area is below the bound (bound_hi, bound_lo) take the other label.
__global__ void k_cc_flip(int n, uint8_t *mask, uint8_t target, const uint32_t *parent, const unsigned long long *acc,
                          uint64_t bound_hi, uint64_t bound_lo) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || mask[i] != target) return;
  const size_t r = parent[i];
  const unsigned __int128 s = (unsigned __int128)acc[4 * r] + ((unsigned __int128)acc[4 * r + 1] << 32) +
                              ((unsigned __int128)acc[4 * r + 2] << 64) + ((unsigned __int128)acc[4 * r + 3] << 96);
  const unsigned __int128 b = ((unsigned __int128)bound_hi << 64) | bound_lo;
  if (s < b) mask[i] = target ? 0 : 1;
}

// ---- growth -------------------------------------------------------------------------------------

// flags[0]: some face is a candidate; flags[1]: some candidate flipped.
__global__ void k_grow_mark(const uint32_t *fa, const uint32_t *fb, int np, const uint8_t *frac, uint8_t *cand,
                            uint32_t *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t a = fa[i], b = fb[i];
  const uint8_t la = frac[a], lb = frac[b];
  if (la == lb) return;
  cand[la ? a : b] = 1;
  flags[0] = 1u;
}

__global__ void k_grow_flip(int n, uint8_t *frac, uint8_t *cand, const uint8_t *elig, uint32_t *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !cand[i]) return;
  cand[i] = 0;
  if (elig[i]) {
    frac[i] = 0;
    flags[1] = 1u;
  }
}

inline dim3 blocks(long n) { return dim3((unsigned)((n > 0 ? n : 1) + 255) / 256); }

}  // namespace

hipError_t launch_seg_geom(const double *V, int nv, const uint32_t *F, int nf, double *N, double *A, double *C,
                           uint32_t *bad, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_geom, blocks(nf), dim3(256), 0, s, V, nv, F, nf, N, A, C, bad);
  return hipGetLastError();
}

hipError_t launch_seg_edge_keys(const uint32_t *F, int nf, uint32_t nv, uint64_t *key, uint32_t *pos, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_edge_keys, blocks(3l * nf), dim3(256), 0, s, F, nf, nv, key, pos);
  return hipGetLastError();
}

hipError_t launch_seg_pairs(const uint64_t *key, const uint32_t *pos, int m, int nf, uint32_t *fa, uint32_t *fb,
                            uint32_t *count, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_pairs, blocks(m), dim3(256), 0, s, key, pos, m, nf, fa, fb, count);
  return hipGetLastError();
}

hipError_t launch_seg_majority(const GridView &g, const double *C, const double *A, const uint8_t *frac_raw, int nf,
                               const uint32_t *rep, int n_rep, const uint32_t *near, double r, uint8_t *maj,
                               uint8_t *frac, uint32_t *err, hipStream_t s) {
  if (n_rep > 0)
    hipLaunchKernelGGL(k_seg_ball<false>, dim3((unsigned)((n_rep + SEG_WAVES - 1) / SEG_WAVES)), dim3(64 * SEG_WAVES),
                       0, s, g, C, A, (const double *)nullptr, frac_raw, nf, rep, n_rep, r, maj, (double *)nullptr,
                       (uint8_t *)nullptr, err);
  hipLaunchKernelGGL(k_seg_take_majority, blocks(nf), dim3(256), 0, s, near, nf, n_rep, maj, frac, err);
  return hipGetLastError();
}

hipError_t launch_seg_reference(const GridView &g, const double *C, const double *A, const double *N,
                                const uint8_t *frac, int nf, const uint32_t *rep, int n_rep, const uint32_t *near,
                                double r, double cos_grow, double *ref_g, uint8_t *has_g, double *ref, uint8_t *has,
                                uint8_t *elig, uint32_t *err, hipStream_t s) {
  if (n_rep > 0)
    hipLaunchKernelGGL(k_seg_ball<true>, dim3((unsigned)((n_rep + SEG_WAVES - 1) / SEG_WAVES)), dim3(64 * SEG_WAVES),
                       0, s, g, C, A, N, frac, nf, rep, n_rep, r, (uint8_t *)nullptr, ref_g, has_g, err);
  hipLaunchKernelGGL(k_seg_take_reference, blocks(nf), dim3(256), 0, s, near, nf, n_rep, ref_g, has_g, N, cos_grow, ref,
                     has, elig);
  return hipGetLastError();
}

hipError_t launch_seg_smoothed(const GridView &g, const double *C, const double *A, const double *N, int nf,
                               const uint32_t *rep, int n_rep, const uint32_t *near, double r, double *ref_g,
                               uint8_t *has_g, double *ns, uint32_t *err, hipStream_t s) {
  if (n_rep > 0)
    hipLaunchKernelGGL(k_seg_ball<true>, dim3((unsigned)((n_rep + SEG_WAVES - 1) / SEG_WAVES)), dim3(64 * SEG_WAVES),
                       0, s, g, C, A, N, (const uint8_t *)nullptr, nf, rep, n_rep, r, (uint8_t *)nullptr, ref_g, has_g,
                       err);
  hipLaunchKernelGGL(k_seg_take_normal, blocks(nf), dim3(256), 0, s, near, nf, n_rep, ref_g, ns, err);
  return hipGetLastError();
}

hipError_t launch_cc_init(int n, uint32_t *parent, hipStream_t s) {
  hipLaunchKernelGGL(k_cc_init, blocks(n), dim3(256), 0, s, n, parent);
  return hipGetLastError();
}

hipError_t launch_cc_round(const uint32_t *fa, const uint32_t *fb, int np, const uint8_t *mask, uint8_t target, int n,
                           uint32_t *parent, uint32_t *changed, hipStream_t s) {
  if (np > 0) hipLaunchKernelGGL(k_cc_hook, blocks(np), dim3(256), 0, s, fa, fb, np, mask, target, parent, changed);
  hipLaunchKernelGGL(k_cc_compress, blocks(n), dim3(256), 0, s, n, parent);
  return hipGetLastError();
}

hipError_t launch_cc_drop(int n, uint8_t *mask, uint8_t target, const uint32_t *parent, const double *A, int sh,
                          unsigned long long *acc, uint64_t bound_hi, uint64_t bound_lo, hipStream_t s) {
  hipLaunchKernelGGL(k_cc_area, blocks(n), dim3(256), 0, s, n, mask, target, parent, A, sh, acc);
  hipLaunchKernelGGL(k_cc_flip, blocks(n), dim3(256), 0, s, n, mask, target, parent, acc, bound_hi, bound_lo);
  return hipGetLastError();
}

hipError_t launch_grow_pass(const uint32_t *fa, const uint32_t *fb, int np, int n, uint8_t *frac, uint8_t *cand,
                            const uint8_t *elig, uint32_t *flags, hipStream_t s) {
  if (np > 0) hipLaunchKernelGGL(k_grow_mark, blocks(np), dim3(256), 0, s, fa, fb, np, frac, cand, flags);
  hipLaunchKernelGGL(k_grow_flip, blocks(n), dim3(256), 0, s, n, frac, cand, elig, flags);
  return hipGetLastError();
}

}  // namespace sherd
