// C ABI for R §3.4.4-3.4.8, the clean-up of the cone vote's mask (crates/sherd-core/src/fragment/
// segment.rs:544-635): the majority filter, the two island passes, the shell's reference normal,
// the boundary growth and the last island pass, all on the device. Host code: validation, the
// buffer layout, and the loops that stop when the device reports nothing left to do (component
// hooking, growth passes).
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"
#include "sort.hpp"

using namespace sherd;

namespace {

enum { SL_SEG = 35 };

constexpr double MAJORITY_T = 0.25;       // segment.rs:67
constexpr double REFERENCE_T = 0.5;       // segment.rs:69
constexpr double MIN_FRACTURE_T2 = 0.5;   // segment.rs:71
constexpr double MIN_SHELL_T2 = 2.0;      // segment.rs:73
constexpr int MAX_GROWTH_PASSES = 60;     // segment.rs:75

// Bump allocation of one call's buffers inside one grow-only scratch slot.
struct Arena {
  size_t used = 0;
  template <class T>
  size_t take(size_t n) {
    size_t at = used;
    used += (sizeof(T) * (n ? n : 1) + 255) & ~(size_t)255;
    return at;
  }
};

// ceil(a 2^sh) as (hi, lo) 64-bit halves: the bound an exact fixed-point area sum is compared with.
void fixed_ceil(double a, int sh, uint64_t *hi, uint64_t *lo) {
  const double x = std::ldexp(a, sh);
  const double h = std::floor(std::ldexp(x, -64));
  const double rem = x - h * 0x1p64;
  unsigned __int128 v = ((unsigned __int128)(uint64_t)h << 64) | (uint64_t)rem;
  if ((double)(uint64_t)rem != rem) v += 1;
  *hi = (uint64_t)(v >> 64);
  *lo = (uint64_t)v;
}

struct SegDev {
  Ctx *c;
  int nf, np;
  const uint32_t *fa, *fb;
  const double *A;
  uint32_t *parent, *flags;
  unsigned long long *acc;
  int sh;  // 64 - E, 2^(E-1) <= t^2 < 2^E
};

// mesh/components.rs:170-205 drop_small_components(mask, target, min_area): components by hooking
// until no pair joins two roots, then each component's exact area against the bound.
int drop_small(const SegDev &d, uint8_t *mask, uint8_t target, double min_area) {
  hipStream_t s = d.c->stream;
  HIPCK(launch_cc_init(d.nf, d.parent, s), "components");
  for (int round = 0;; ++round) {
    if (round > d.nf + 1) return fail(SHERD_E_LIMIT, "components did not settle");
    HIPCK(hipMemsetAsync(d.flags + 1, 0, sizeof(uint32_t), s), "clear");
    HIPCK(launch_cc_round(d.fa, d.fb, d.np, mask, target, d.nf, d.parent, d.flags + 1, s), "components");
    uint32_t changed = 0;
    HIPCK(hipMemcpyAsync(&changed, d.flags + 1, sizeof changed, hipMemcpyDeviceToHost, s), "download");
    HIPCK(hipStreamSynchronize(s), "synchronize");
    if (!changed) break;
  }
  uint64_t bhi, blo;
  fixed_ceil(min_area, d.sh, &bhi, &blo);
  HIPCK(hipMemsetAsync(d.acc, 0, sizeof(unsigned long long) * 4 * (size_t)d.nf, s), "clear");
  HIPCK(launch_cc_drop(d.nf, mask, target, d.parent, d.A, d.sh, d.acc, bhi, blo, s), "island areas");
  return SHERD_OK;
}

}  // namespace

extern "C" {

int sherd_segment_cleanup(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf, double thick,
                          const uint8_t *frac_raw, const uint32_t *rep, size_t n_rep, const uint32_t *near,
                          double boundary_angle, uint8_t *labels, sherd_seg_trace *trace) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (trace) trace->passes = 0;
  if (nf == 0) return SHERD_OK;
  if (!V || !F || !frac_raw || !rep || !near || !labels) return fail(SHERD_E_ARG, "null buffer");
  if (nf > (size_t)INT_MAX / 3 || nv > (size_t)UINT32_MAX || n_rep > (size_t)INT_MAX)
    return fail(SHERD_E_LIMIT, "mesh too large");
  // segment.rs:520-529: no wall, no segmentation: every face shell.
  if (!std::isfinite(thick) || thick <= 0.0) {
    std::memset(labels, 0, nf);
    return SHERD_OK;
  }
  if (n_rep == 0) return fail(SHERD_E_ARG, "no voxel representatives");
  if (!std::isfinite(boundary_angle)) return fail(SHERD_E_ARG, "boundary angle not finite");
  const int n = (int)nf, m = 3 * n;
  hipStream_t s = c->stream;

  Arena ar;
  const size_t oV = ar.take<double>(3 * nv), oF = ar.take<uint32_t>(3 * nf), oN = ar.take<double>(3 * nf),
               oA = ar.take<double>(nf), oC = ar.take<double>(3 * nf), oRaw = ar.take<uint8_t>(nf),
               oFrac = ar.take<uint8_t>(nf), oRep = ar.take<uint32_t>(n_rep), oNear = ar.take<uint32_t>(nf),
               oMaj = ar.take<uint8_t>(n_rep), oKey = ar.take<uint64_t>(m), oPos = ar.take<uint32_t>(m),
               oTk = ar.take<uint64_t>(m), oTv = ar.take<uint32_t>(m), oHist = ar.take<uint32_t>(rsort::hist_words(m)),
               oFa = ar.take<uint32_t>(m), oFb = ar.take<uint32_t>(m), oPar = ar.take<uint32_t>(nf),
               oAcc = ar.take<unsigned long long>(4 * nf), oRefG = ar.take<double>(3 * n_rep),
               oHasG = ar.take<uint8_t>(n_rep), oRef = ar.take<double>(3 * nf), oHas = ar.take<uint8_t>(nf),
               oElig = ar.take<uint8_t>(nf), oCand = ar.take<uint8_t>(nf), oFlags = ar.take<uint32_t>(8);
  hipError_t e;
  char *base = (char *)c->scratch(SL_SEG, ar.used, &e);
  if (!base) return hip_fail(e, "scratch");
  double *dV = (double *)(base + oV), *dN = (double *)(base + oN), *dA = (double *)(base + oA),
         *dC = (double *)(base + oC), *dRefG = (double *)(base + oRefG), *dRef = (double *)(base + oRef);
  uint32_t *dF = (uint32_t *)(base + oF), *dRep = (uint32_t *)(base + oRep), *dNear = (uint32_t *)(base + oNear),
           *dPos = (uint32_t *)(base + oPos), *dTv = (uint32_t *)(base + oTv), *dHist = (uint32_t *)(base + oHist),
           *dFa = (uint32_t *)(base + oFa), *dFb = (uint32_t *)(base + oFb), *dPar = (uint32_t *)(base + oPar),
           *dFlags = (uint32_t *)(base + oFlags);
  uint64_t *dKey = (uint64_t *)(base + oKey), *dTk = (uint64_t *)(base + oTk);
  uint8_t *dRaw = (uint8_t *)(base + oRaw), *dFrac = (uint8_t *)(base + oFrac), *dMaj = (uint8_t *)(base + oMaj),
          *dHasG = (uint8_t *)(base + oHasG), *dHas = (uint8_t *)(base + oHas), *dElig = (uint8_t *)(base + oElig),
          *dCand = (uint8_t *)(base + oCand);
  unsigned long long *dAcc = (unsigned long long *)(base + oAcc);

  // flags: [0] input errors (k_seg_geom's face index, the balls' bits), [1] components changed,
  // [2..3] growth's candidate / flipped.
  HIPCK(hipMemsetAsync(dFlags, 0, 8 * sizeof(uint32_t), s), "clear");
  HIPCK(hipMemsetAsync(dCand, 0, nf, s), "clear");
  HIPCK(hipMemcpyAsync(dV, V, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dF, F, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dRaw, frac_raw, nf, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dRep, rep, sizeof(uint32_t) * n_rep, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dNear, near, sizeof(uint32_t) * nf, hipMemcpyHostToDevice, s), "upload");
  Timer tm(c, SHERD_M_SEGMENT, nf);
  if (c->prof.on) c->prof.begin("k_seg_setup", s);
  HIPCK(launch_seg_geom(dV, (int)nv, dF, n, dN, dA, dC, dFlags, s), "face geometry");
  // mesh/adjacency.rs:138-158: the rows sorted stably by key, consecutive equal keys paired.
  HIPCK(launch_seg_edge_keys(dF, n, (uint32_t)nv, dKey, dPos, s), "edges");
  int bits = 1;
  while (bits < 64 && ((unsigned __int128)1 << bits) <= (unsigned __int128)nv * nv) ++bits;
  HIPCK(rsort::sort_pairs<uint64_t>(dKey, dPos, dTk, dTv, dHist, m, 0, bits, s), "edge sort");
  HIPCK(hipMemsetAsync(dFlags + 4, 0, sizeof(uint32_t), s), "clear");
  HIPCK(launch_seg_pairs(dKey, dPos, m, n, dFa, dFb, dFlags + 4, s), "adjacency");
  if (c->prof.on) c->prof.end(s);
  uint32_t head[5];
  HIPCK(hipMemcpyAsync(head, dFlags, sizeof head, hipMemcpyDeviceToHost, s), "download");
  HIPCK(hipStreamSynchronize(s), "synchronize");
  if (head[0] & 1u) return fail(SHERD_E_ARG, "face index out of range");
  const int np = (int)head[4];

  DevGrid grid;
  HIPCK(grid_build(dC, n, thick * REFERENCE_T, &grid, s), "centroid grid");
  struct GridGuard {
    DevGrid *g;
    ~GridGuard() { grid_free(g); }
  } guard{&grid};
  const GridView gv = view_of(grid);

  // 3.4.4 the majority filter.
  if (c->prof.on) c->prof.begin("k_seg_majority", s);
  HIPCK(launch_seg_majority(gv, dC, dA, dRaw, n, dRep, (int)n_rep, dNear, thick * MAJORITY_T, dMaj, dFrac, dFlags, s),
        "majority");
  if (c->prof.on) c->prof.end(s);
  if (trace && trace->frac_majority)
    HIPCK(hipMemcpyAsync(trace->frac_majority, dFrac, nf, hipMemcpyDeviceToHost, s), "download");

  // 3.4.5-3.4.6 islands: fracture below 0.5 t^2 becomes shell, shell below 2 t^2 fracture.
  int ex;
  std::frexp(thick * thick, &ex);
  const SegDev d{c, n, np, dFa, dFb, dA, dPar, dFlags, dAcc, 64 - ex};
  if (c->prof.on) c->prof.begin("k_seg_islands", s);
  int rc;
  if ((rc = drop_small(d, dFrac, 1, MIN_FRACTURE_T2 * thick * thick))) return rc;
  if ((rc = drop_small(d, dFrac, 0, MIN_SHELL_T2 * thick * thick))) return rc;
  if (c->prof.on) c->prof.end(s);
  if (trace && trace->frac_islands)
    HIPCK(hipMemcpyAsync(trace->frac_islands, dFrac, nf, hipMemcpyDeviceToHost, s), "download");

  // 3.4.7 the shell's fixed reference normal, then the growth passes.
  const double cos_grow = std::cos(boundary_angle * (M_PI / 180.0));
  if (c->prof.on) c->prof.begin("k_seg_reference", s);
  HIPCK(launch_seg_reference(gv, dC, dA, dN, dFrac, n, dRep, (int)n_rep, dNear, thick * REFERENCE_T, cos_grow, dRefG,
                             dHasG, dRef, dHas, dElig, dFlags, s),
        "reference normals");
  if (c->prof.on) c->prof.end(s);
  if (trace && trace->reference)
    HIPCK(hipMemcpyAsync(trace->reference, dRef, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost, s), "download");
  if (trace && trace->has_ref) HIPCK(hipMemcpyAsync(trace->has_ref, dHas, nf, hipMemcpyDeviceToHost, s), "download");
  int passes = 0;
  for (int pass = 0; pass < MAX_GROWTH_PASSES; ++pass) {
    HIPCK(hipMemsetAsync(dFlags + 2, 0, 2 * sizeof(uint32_t), s), "clear");
    HIPCK(launch_grow_pass(dFa, dFb, np, n, dFrac, dCand, dElig, dFlags + 2, s), "growth");
    uint32_t g[2];
    HIPCK(hipMemcpyAsync(g, dFlags + 2, sizeof g, hipMemcpyDeviceToHost, s), "download");
    HIPCK(hipStreamSynchronize(s), "synchronize");
    if (!g[0]) break;
    ++passes;
    if (!g[1]) break;
  }

  // 3.4.8 one more fracture island pass.
  if ((rc = drop_small(d, dFrac, 1, MIN_FRACTURE_T2 * thick * thick))) return rc;
  HIPCK(hipMemcpyAsync(labels, dFrac, nf, hipMemcpyDeviceToHost, s), "download");
  uint32_t err = 0;
  HIPCK(hipMemcpyAsync(&err, dFlags, sizeof err, hipMemcpyDeviceToHost, s), "download");
  HIPCK(hipStreamSynchronize(s), "synchronize");
  tm.stop();
  c->prof.collect();
  if (err & 1u) return fail(SHERD_E_ARG, "a representative outside the faces");
  if (err & 2u) return fail(SHERD_E_LIMIT, "more faces within a representative's ball than the kernel holds");
  if (err & 4u) return fail(SHERD_E_ARG, "a face's representative outside the list");
  if (trace) trace->passes = passes;
  return SHERD_OK;
}

int sherd_smoothed_normals(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf,
                           const uint32_t *rep, size_t n_rep, const uint32_t *near, double radius, double *ns) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (nf == 0) return SHERD_OK;
  if (!V || !F || !rep || !near || !ns) return fail(SHERD_E_ARG, "null buffer");
  if (nf > (size_t)INT_MAX / 3 || nv > (size_t)UINT32_MAX || n_rep > (size_t)INT_MAX)
    return fail(SHERD_E_LIMIT, "mesh too large");
  if (n_rep == 0) return fail(SHERD_E_ARG, "no voxel representatives");
  if (!std::isfinite(radius) || radius < 0.0) return fail(SHERD_E_ARG, "radius must be finite and non-negative");
  const int n = (int)nf;
  hipStream_t s = c->stream;
  Arena ar;
  const size_t oV = ar.take<double>(3 * nv), oF = ar.take<uint32_t>(3 * nf), oN = ar.take<double>(3 * nf),
               oA = ar.take<double>(nf), oC = ar.take<double>(3 * nf), oRep = ar.take<uint32_t>(n_rep),
               oNear = ar.take<uint32_t>(nf), oRefG = ar.take<double>(3 * n_rep), oHasG = ar.take<uint8_t>(n_rep),
               oNs = ar.take<double>(3 * nf), oFlags = ar.take<uint32_t>(8);
  hipError_t e;
  char *base = (char *)c->scratch(SL_SEG, ar.used, &e);
  if (!base) return hip_fail(e, "scratch");
  double *dV = (double *)(base + oV), *dN = (double *)(base + oN), *dA = (double *)(base + oA),
         *dC = (double *)(base + oC), *dRefG = (double *)(base + oRefG), *dNs = (double *)(base + oNs);
  uint32_t *dF = (uint32_t *)(base + oF), *dRep = (uint32_t *)(base + oRep), *dNear = (uint32_t *)(base + oNear),
           *dFlags = (uint32_t *)(base + oFlags);
  uint8_t *dHasG = (uint8_t *)(base + oHasG);
  HIPCK(hipMemsetAsync(dFlags, 0, 8 * sizeof(uint32_t), s), "clear");
  HIPCK(hipMemcpyAsync(dV, V, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dF, F, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dRep, rep, sizeof(uint32_t) * n_rep, hipMemcpyHostToDevice, s), "upload");
  HIPCK(hipMemcpyAsync(dNear, near, sizeof(uint32_t) * nf, hipMemcpyHostToDevice, s), "upload");
  Timer tm(c, SHERD_M_SEGMENT, nf);
  HIPCK(launch_seg_geom(dV, (int)nv, dF, n, dN, dA, dC, dFlags, s), "face geometry");
  uint32_t bad = 0;
  HIPCK(hipMemcpyAsync(&bad, dFlags, sizeof bad, hipMemcpyDeviceToHost, s), "download");
  HIPCK(hipStreamSynchronize(s), "synchronize");
  if (bad & 1u) return fail(SHERD_E_ARG, "face index out of range");
  DevGrid grid;
  HIPCK(grid_build(dC, n, radius > 0.0 ? radius : 1.0, &grid, s), "centroid grid");
  struct GridGuard {
    DevGrid *g;
    ~GridGuard() { grid_free(g); }
  } guard{&grid};
  HIPCK(launch_seg_smoothed(view_of(grid), dC, dA, dN, n, dRep, (int)n_rep, dNear, radius, dRefG, dHasG, dNs, dFlags,
                            s),
        "smoothed normals");
  HIPCK(hipMemcpyAsync(ns, dNs, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost, s), "download");
  uint32_t err = 0;
  HIPCK(hipMemcpyAsync(&err, dFlags, sizeof err, hipMemcpyDeviceToHost, s), "download");
  HIPCK(hipStreamSynchronize(s), "synchronize");
  tm.stop();
  if (err & 1u) return fail(SHERD_E_ARG, "a representative outside the faces");
  if (err & 2u) return fail(SHERD_E_LIMIT, "more faces within a representative's ball than the kernel holds");
  if (err & 4u) return fail(SHERD_E_ARG, "a face's representative outside the list");
  return SHERD_OK;
}

}  // extern "C"
