// C ABI for R §7's ICP: cloud handles, one rung over a batch of cloud pairs and candidates, the
// whole ladder as one captured graph, and the per-call form of Executor::icp_rung. Host code only:
// descriptors, ordering, the workgroup-size autotuner and graph capture; the rungs run in icp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"

using namespace sherd;



namespace {

// Scratch slots of this file (the context's other entry points use slots below 21).
enum { SL_BLOB = 21, SL_CAND = 22, SL_T = 23, SL_OUT = 24, SL_MOVED = 25, SL_CORR = 26, SL_PART = 27, SL_SKIP = 28,
       SL_WIDE = 29 };

constexpr size_t MAX_POINTS = (size_t)ICP_SPAN * ICP_MAX_SPANS;
constexpr size_t MAX_RUNGS = 8;


__global__ void k_point_tris(const double *p, int n, float *V, uint32_t *F) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; ++k) {
    V[3 * (size_t)i + k] = (float)p[3 * (size_t)i + k];
    F[3 * (size_t)i + k] = (uint32_t)i;
  }
}

// After the build a point tree's leaf records need only their point: each record's first 24 bytes
// take the point's f64 coordinates (the f32 copies there served the build and nothing reads them
// again) and `face` keeps its index, so a leaf costs one fetch (cloud_nearest).
__global__ void k_point_records(Tri *t, int n, const double *p) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t j = t[s].face;
  double x = p[3 * (size_t)j], y = p[3 * (size_t)j + 1], z = p[3 * (size_t)j + 2];
  double *d = reinterpret_cast<double *>(t + s);
  d[0] = x;
  d[1] = y;
  d[2] = z;
}

}  // namespace

void sherd::free_cloud(sherd_cloud *cl) {
  if (cl->pts) (void)hipFree(cl->pts);
  if (cl->normals) (void)hipFree(cl->normals);
  scene_free(&cl->tree);
  delete cl;
}

hipError_t sherd::point_tree(Ctx *c, sherd_cloud *cl) {
  if (cl->n <= 0) return hipSuccess;
  size_t n = (size_t)cl->n;
  float *V = nullptr;
  uint32_t *F = nullptr;
  hipError_t e;
  do {
    if ((e = hipMalloc((void **)&V, sizeof(float) * 3 * n)) != hipSuccess) break;
    if ((e = hipMalloc((void **)&F, sizeof(uint32_t) * 3 * n)) != hipSuccess) break;
    hipLaunchKernelGGL(k_point_tris, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, cl->pts, (int)n, V,
                       F);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = scene_build(V, F, (int)n, &cl->tree, c->stream)) != hipSuccess) break;
    hipLaunchKernelGGL(k_point_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, cl->tree.tris,
                       (int)n, cl->pts);
    e = hipGetLastError();
  } while (0);
  if (V) (void)hipFree(V);
  if (F) (void)hipFree(F);
  return e;
}

namespace {

// Builds a cloud on the device: the f64 points (and normals), and an LBVH whose leaves are the
// points (degenerate triangles whose face index is the point's).
int make_cloud(Ctx *c, const double *pts, const double *normals, size_t n, sherd_cloud **out) {
  *out = nullptr;
  if (n > MAX_POINTS) return fail(SHERD_E_LIMIT, "cloud beyond 262144 points");
  sherd_cloud *cl = new sherd_cloud();
  cl->ctx = c;
  cl->n = (int)n;
  for (int k = 0; k < 3; ++k) {
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m += pts[3 * i + k];
    cl->centroid[k] = n ? m / (double)n : 0.0;
  }
  hipError_t e = hipSuccess;
  if (n > 0) {
    do {
      if ((e = hipMalloc((void **)&cl->pts, sizeof(double) * 3 * n)) != hipSuccess) break;
      if ((e = hipMemcpyAsync(cl->pts, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream)) !=
          hipSuccess)
        break;
      if (normals) {
        if ((e = hipMalloc((void **)&cl->normals, sizeof(double) * 3 * n)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(cl->normals, normals, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream)) !=
            hipSuccess)
          break;
      }
      e = point_tree(c, cl);
    } while (0);
  }
  if (e != hipSuccess) {
    free_cloud(cl);
    return hip_fail(e, "cloud build");
  }
  *out = cl;
  return SHERD_OK;
}

}  // namespace

int sherd::cloud_adopt(Ctx *c, double *dpts, double *dnormals, size_t n, bool tree, sherd_cloud **out) {
  *out = nullptr;
  sherd_cloud *cl = new sherd_cloud();
  cl->ctx = c;
  cl->n = (int)n;
  cl->pts = dpts;
  cl->normals = dnormals;
  if (n > MAX_POINTS) {
    free_cloud(cl);
    return fail(SHERD_E_LIMIT, "cloud beyond 262144 points");
  }
  std::vector<double> h(3 * n);
  hipError_t e = hipSuccess;
  if (n > 0) {
    e = hipMemcpyAsync(h.data(), dpts, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) {
    for (int k = 0; k < 3; ++k) {
      double m = 0.0;
      for (size_t i = 0; i < n; ++i) m += h[3 * i + k];
      cl->centroid[k] = n ? m / (double)n : 0.0;
    }
    if (tree) e = point_tree(c, cl);
  }
  if (e != hipSuccess) {
    free_cloud(cl);
    return hip_fail(e, "cloud build");
  }
  *out = cl;
  return SHERD_OK;
}

namespace {

void unmoved(const double *init, sherd_registration *r) {
  // Registration::unmoved (icp.rs:198-209).
  std::memcpy(r->transform, init, sizeof(double) * 16);
  r->fitness = r->inlier_rmse = 0.0;
  r->correspondences = 0;
  r->iterations = 0;
  r->converged = 0;
}

void to_registration(const double *o, sherd_registration *r) {
  std::memcpy(r->transform, o, sizeof(double) * 16);
  r->fitness = o[16];
  r->inlier_rmse = o[17];
  r->correspondences = (uint64_t)o[18];
  r->iterations = (uint32_t)o[19];
  r->converged = (uint32_t)o[20];
}

// One rung of a batch, laid out for the kernel. Candidates on a pair with an empty cloud are left
// out of `order` (the host answers them: they keep their pose).
struct RungPlan {
  std::vector<IcpPairDesc> pd;
  std::vector<uint32_t> order;
  std::vector<size_t> off;
  size_t points = 0;
  int max_spans = 1;
  int est = 0;
  int max_iter = 0;
  double mean_ns = 0.0;
  // The wide form's span table: per queue position its first span, then the total.
  std::vector<uint32_t> span_first;
};

// The forms a rung can run in: the persistent batched kernel at a workgroup size, or (1) the wide
// form. All give the same registrations to the bit; the autotuner picks by speed.
constexpr int WIDE = 1;

int plan_rung(Ctx *c, const sherd_icp_pair *pairs, size_t n_pairs, const uint32_t *cand_pair, size_t n_cand,
              int estimation, int max_iteration, RungPlan *pl) {
  pl->pd.resize(n_pairs);
  for (size_t p = 0; p < n_pairs; ++p) {
    const sherd_cloud *s = pairs[p].source, *t = pairs[p].target;
    if (!s || !t) return fail(SHERD_E_ARG, "pair without a cloud");
    if (s->ctx != c || t->ctx != c) return fail(SHERD_E_ARG, "cloud of another context");
    if (t->n > 0 && !t->tree.nodes) return fail(SHERD_E_ARG, "target cloud without a point tree");
    IcpPairDesc &d = pl->pd[p];
    std::memset(&d, 0, sizeof d);
    d.src = s->pts;
    d.ns = t->n > 0 ? s->n : 0;
    d.tp = t->pts;
    d.tn = estimation == SHERD_POINT_TO_PLANE ? t->normals : nullptr;
    if (t->n > 0) d.tree = view_of(t->tree);
    d.r = pairs[p].distance;
    for (int k = 0; k < 3; ++k) {
      d.cs[k] = s->centroid[k];
      d.ct[k] = t->centroid[k];
    }
  }
  // Point-to-plane without normals makes identity updates (icp.rs:460-462): the kernel's est 1
  // with a null normal pointer does that per pair.
  pl->est = estimation == SHERD_POINT_TO_PLANE ? 1 : 0;
  pl->max_iter = max_iteration;
  pl->off.assign(n_cand, 0);
  pl->order.clear();
  double sum = 0.0;
  for (size_t i = 0; i < n_cand; ++i) {
    if (cand_pair[i] >= n_pairs) return fail(SHERD_E_ARG, "candidate names a pair out of range");
    const IcpPairDesc &d = pl->pd[cand_pair[i]];
    if (d.ns == 0) continue;
    pl->off[i] = pl->points;
    pl->points += (size_t)d.ns;
    pl->order.push_back((uint32_t)i);
    pl->max_spans = std::max(pl->max_spans, (d.ns + ICP_SPAN - 1) / ICP_SPAN);
    sum += d.ns;
  }
  pl->mean_ns = pl->order.empty() ? 0.0 : sum / (double)pl->order.size();
  // Longest source first, so the persistent workgroups finish together (sorted on a compact copy of
  // the sizes: the descriptors are large).
  std::vector<std::pair<int, uint32_t>> keyed(pl->order.size());
  for (size_t k = 0; k < keyed.size(); ++k) keyed[k] = {-pl->pd[cand_pair[pl->order[k]]].ns, pl->order[k]};
  std::stable_sort(keyed.begin(), keyed.end(),
                   [](const std::pair<int, uint32_t> &a, const std::pair<int, uint32_t> &b) { return a.first < b.first; });
  pl->span_first.assign(1, 0u);
  for (size_t k = 0; k < keyed.size(); ++k) {
    pl->order[k] = keyed[k].second;
    pl->span_first.push_back(pl->span_first.back() + (uint32_t)((-keyed[k].first + ICP_SPAN - 1) / ICP_SPAN));
  }
  return SHERD_OK;
}

int log2_bucket(double x, int top) {
  int b = x > 1.0 ? (int)std::floor(std::log2(x)) : 0;
  return std::min(top, std::max(0, b));
}

// The autotuner's choice for a rung, by estimation, source size and candidate count.
int &tuned_block(Ctx *c, const RungPlan &pl) {
  return c->icp_tuned[pl.est][log2_bucket(pl.mean_ns, 39)][log2_bucket((double)pl.order.size(), 23)];
}

// Persistent workgroups for a rung: enough to fill the device, no more than candidates, and the
// span partials within 512 MiB.
int grid_for(Ctx *c, int block, size_t n, int max_spans) {
  size_t g = (size_t)std::max(1, c->cus) * (size_t)(2048 / block);
  g = std::min(g, std::max<size_t>(n, 1));
  size_t per = (size_t)max_spans * ICP_NV * sizeof(double);
  g = std::min(g, std::max<size_t>(1, ((size_t)512 << 20) / per));
  return (int)g;
}

// The device side of a batch: per rung its pair descriptors, candidate order, point offsets and
// queue counter, packed in one upload; shared by the rungs: candidate pairs, moved points,
// correspondences and span partials.
struct BatchDev {
  std::vector<IcpBatchView> v;
  std::vector<IcpWideView> w;  // the wide view of each rung (used by the wide ones)
  std::vector<IcpWideSplit> split;  // where each rung's queue splits into two chains
  // The context's pinned staging buffer: the blob and the candidate pairs at its front (read by
  // copies until the stream drains), then `extra` bytes from `pin_at` for the caller.
  char *pin = nullptr;
  size_t pin_at = 0;
};

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

int stage_batch(Ctx *c, const std::vector<RungPlan> &plans, const uint32_t *cand_pair, size_t n_cand, double *dT,
                double *dout, size_t extra, BatchDev *bd) {
  hipError_t e;
  size_t nr = plans.size();
  size_t max_points = 1;
  int max_spans = 1, max_grid = 1;
  for (const RungPlan &pl : plans) {
    max_points = std::max(max_points, pl.points);
    max_spans = std::max(max_spans, pl.max_spans);
    // The smallest workgroup the autotuner may try has the largest grid.
    max_grid = std::max(max_grid, grid_for(c, 64, pl.order.size(), pl.max_spans));
  }
  size_t wide_spans = 0;
  for (const RungPlan &pl : plans) wide_spans = std::max<size_t>(wide_spans, pl.span_first.back());
  auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
  constexpr int F = 5;  // blob fields per rung
  std::vector<size_t> at(nr * F);
  size_t bytes = 0;
  for (size_t r = 0; r < nr; ++r) {
    at[F * r] = bytes;
    bytes = al(bytes + sizeof(IcpPairDesc) * std::max<size_t>(plans[r].pd.size(), 1));
    at[F * r + 1] = bytes;
    bytes = al(bytes + sizeof(uint32_t) * std::max<size_t>(plans[r].order.size(), 1));
    at[F * r + 2] = bytes;
    bytes = al(bytes + sizeof(size_t) * n_cand);
    at[F * r + 3] = bytes;
    bytes = al(bytes + sizeof(uint32_t));
    at[F * r + 4] = bytes;
    bytes = al(bytes + sizeof(uint32_t) * plans[r].span_first.size());
  }
  const size_t at_cp = up256(bytes);
  bd->pin_at = up256(at_cp + sizeof(uint32_t) * n_cand);
  bd->pin = (char *)c->pinned(bd->pin_at + extra, &e);
  if (!bd->pin) return hip_fail(e, "pinned staging");
  char *blob = bd->pin;
  std::memset(blob, 0, bytes);
  std::memcpy(blob + at_cp, cand_pair, sizeof(uint32_t) * n_cand);
  for (size_t r = 0; r < nr; ++r) {
    const RungPlan &pl = plans[r];
    std::memcpy(&blob[at[F * r]], pl.pd.data(), sizeof(IcpPairDesc) * pl.pd.size());
    if (!pl.order.empty()) std::memcpy(&blob[at[F * r + 1]], pl.order.data(), sizeof(uint32_t) * pl.order.size());
    std::memcpy(&blob[at[F * r + 2]], pl.off.data(), sizeof(size_t) * n_cand);
    std::memcpy(&blob[at[F * r + 4]], pl.span_first.data(), sizeof(uint32_t) * pl.span_first.size());
  }
  char *dblob = (char *)c->scratch(SL_BLOB, bytes, &e);
  if (!dblob) return hip_fail(e, "scratch");
  uint32_t *dcp = (uint32_t *)c->scratch(SL_CAND, sizeof(uint32_t) * n_cand, &e);
  if (!dcp) return hip_fail(e, "scratch");
  double *dmoved = (double *)c->scratch(SL_MOVED, sizeof(double) * 3 * max_points, &e);
  if (!dmoved) return hip_fail(e, "scratch moved points");
  uint32_t *dcorr = (uint32_t *)c->scratch(SL_CORR, sizeof(uint32_t) * max_points, &e);
  if (!dcorr) return hip_fail(e, "scratch correspondences");
  size_t part_rows = std::max((size_t)max_spans * max_grid, wide_spans);
  double *dpart = (double *)c->scratch(SL_PART, sizeof(double) * ICP_NV * part_rows, &e);
  if (!dpart) return hip_fail(e, "scratch");
  IcpWideState *dst = (IcpWideState *)c->scratch(SL_WIDE, sizeof(IcpWideState) * n_cand, &e);
  if (!dst) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(dblob, blob, bytes, hipMemcpyHostToDevice, c->stream), "upload");
  HIPCK(hipMemcpyAsync(dcp, blob + at_cp, sizeof(uint32_t) * n_cand, hipMemcpyHostToDevice, c->stream), "upload");
  bd->v.assign(nr, IcpBatchView{});
  bd->w.assign(nr, IcpWideView{});
  bd->split.assign(nr, IcpWideSplit{});
  for (size_t r = 0; r < nr; ++r) {
    IcpBatchView &v = bd->v[r];
    v.pairs = (const IcpPairDesc *)(dblob + at[F * r]);
    v.cand_pair = dcp;
    v.order = (const uint32_t *)(dblob + at[F * r + 1]);
    v.off = (const size_t *)(dblob + at[F * r + 2]);
    v.next = (uint32_t *)(dblob + at[F * r + 3]);
    v.T_in = dT + 16 * n_cand * r;
    v.T_out = dT + 16 * n_cand * (r + 1);
    v.out = dout + 21 * n_cand * r;
    v.moved = dmoved;
    v.corr = dcorr;
    v.part = dpart;
    v.part_spans = max_spans;
    v.n_cand = (int)plans[r].order.size();
    v.est = plans[r].est;
    v.max_iter = plans[r].max_iter;
    IcpWideView &w = bd->w[r];
    w.b = v;
    w.span_first = (const uint32_t *)(dblob + at[F * r + 4]);
    w.n_spans = (int)plans[r].span_first.back();
    w.st = dst;
    w.part = dpart;
    w.k0 = w.s0 = 0;
    // Chains of about equal spans, as many as keep each at four spans per compute unit or more.
    const std::vector<uint32_t> &sf = plans[r].span_first;
    const size_t nq = sf.size() - 1;
    const uint32_t total = sf.back();
    IcpWideSplit &sp = bd->split[r];
    sp.n = 1;
    sp.k[1] = (int)nq;
    sp.s[1] = (int)total;
    const int parts = std::min<int>(ICP_CHAINS, (int)(total / (4u * (uint32_t)std::max(1, c->cus))));
    if (parts > 1 && nq >= (size_t)parts) {
      sp.n = parts;
      for (int j = 1; j < parts; ++j) {
        size_t k = std::lower_bound(sf.begin(), sf.end() - 1, (uint32_t)((uint64_t)total * j / parts)) - sf.begin();
        k = std::max<size_t>(k, (size_t)sp.k[j - 1] + 1);
        sp.k[j] = (int)std::min(k, nq - (size_t)(parts - j));
        sp.s[j] = (int)sf[sp.k[j]];
      }
      sp.k[parts] = (int)nq;
      sp.s[parts] = (int)total;
    }
  }
  return SHERD_OK;
}

// `timed`: each wide pass's correspondence kernel between profiler events (never in a capture).
hipError_t launch_rung(Ctx *c, const BatchDev &bd, size_t r, int block, int grid, bool timed) {
  if (block == WIDE)
    return launch_icp_wide(bd.w[r], c->stream, timed && c->prof.on ? &c->prof : nullptr, bd.split[r], c->side.side,
                           c->side.fork, c->side.join);
  return launch_icp_batch(bd.v[r], block, grid, c->stream);
}

// Candidates the kernel skips (an empty cloud) carry their pose into the next rung.
__global__ void k_carry(const double *T_in, double *T_out, const uint8_t *skip, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 16 * n) return;
  if (skip[i / 16]) T_out[i] = T_in[i];
}

// Runs rung r eagerly. Without a workgroup size yet for its bucket, the autotuner times the rung
// once per size (every size gives the same result, so the last run's output stands) and keeps the
// fastest.
int run_rung(Ctx *c, const BatchDev &bd, size_t r, const RungPlan &pl) {
  const IcpBatchView &v = bd.v[r];
  if (v.n_cand == 0) return SHERD_OK;
  int &tuned = tuned_block(c, pl);
  int block = c->icp_block ? c->icp_block : tuned;
  if (block) {
    if (c->prof.on && block != WIDE) c->prof.begin("k_icp_batch", c->stream);
    HIPCK(launch_rung(c, bd, r, block, grid_for(c, block, pl.order.size(), pl.max_spans), true), "icp kernel");
    if (c->prof.on && block != WIDE) c->prof.end(c->stream);
    return SHERD_OK;
  }
  // Fewer candidates than a quarter of the compute units leave the batched kernel at most one
  // workgroup per candidate on a nearly idle device: only its largest size is worth timing there.
  int sizes[8];
  int n = icp_block_sizes(sizes);
  if (pl.order.size() * 4 < (size_t)std::max(1, c->cus)) {
    sizes[0] = sizes[n - 1];
    n = 1;
  }
  sizes[n++] = WIDE;
  float best = 0.0f;
  int pick = 0;
  for (int k = 0; k < n; ++k) {
    int b = sizes[k];
    HIPCK(hipEventRecord(c->ev0, c->stream), "event");
    HIPCK(launch_rung(c, bd, r, b, b == WIDE ? 0 : grid_for(c, b, pl.order.size(), pl.max_spans), false), "icp kernel");
    HIPCK(hipEventRecord(c->ev1, c->stream), "event");
    HIPCK(hipEventSynchronize(c->ev1), "event");
    float ms = 0.0f;
    HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1), "event");
    if (!pick || ms < best) {
      pick = b;
      best = ms;
    }
  }
  tuned = pick;
  return SHERD_OK;
}

int check_rung(const sherd_icp_rung_spec &r) {
  if (r.estimation != SHERD_POINT_TO_POINT && r.estimation != SHERD_POINT_TO_PLANE)
    return fail(SHERD_E_ARG, "unknown estimation");
  if (r.max_iteration < 0) return fail(SHERD_E_ARG, "negative iteration cap");
  return SHERD_OK;
}

// The ladder: rungs in order on the device, each from the previous rung's poses. When every rung
// has a workgroup size and the batch repeats the previous call's launch signature (every kernel
// argument), the launches are captured once as a graph and replayed while the signature holds;
// otherwise the rungs run eagerly (and tune).
int run_ladder(Ctx *c, const sherd_icp_pair *pairs, size_t n_pairs, const sherd_icp_rung_spec *rungs, size_t n_rungs,
               const uint32_t *cand_pair, const double *inits, size_t n_cand, sherd_registration *out) {
  int rc;
  std::vector<RungPlan> plans(n_rungs);
  for (size_t r = 0; r < n_rungs; ++r) {
    if ((rc = check_rung(rungs[r]))) return rc;
    if ((rc = plan_rung(c, pairs + r * n_pairs, n_pairs, cand_pair, n_cand, rungs[r].estimation,
                        rungs[r].max_iteration, &plans[r])))
      return rc;
  }
  hipError_t e;
  double *dT = (double *)c->scratch(SL_T, sizeof(double) * 16 * n_cand * (n_rungs + 1), &e);
  if (!dT) return hip_fail(e, "scratch");
  double *dout = (double *)c->scratch(SL_OUT, sizeof(double) * 21 * n_cand * n_rungs, &e);
  if (!dout) return hip_fail(e, "scratch");
  // Per rung, which candidates the kernel skips (one byte each).
  uint8_t *dskip = (uint8_t *)c->scratch(SL_SKIP, n_cand * n_rungs, &e);
  if (!dskip) return hip_fail(e, "scratch");
  std::vector<uint8_t> skip(n_cand * n_rungs, 0);
  bool any_skip = false;
  for (size_t r = 0; r < n_rungs; ++r)
    for (size_t i = 0; i < n_cand; ++i)
      if (plans[r].pd[cand_pair[i]].ns == 0) {
        skip[r * n_cand + i] = 1;
        any_skip = true;
      }
  BatchDev bd;
  // Every return below leaves with the stream drained, so the copies queued from the pinned
  // staging buffer have finished before another call reuses it.
  StreamFence fence(c->stream);
  // After the blob in the staging buffer: the initial poses and skip bytes going up, the outputs
  // and poses coming down.
  const size_t a_init = 0, a_skip = up256(sizeof(double) * 16 * n_cand), a_out = a_skip + up256(skip.size()),
               a_T = a_out + up256(sizeof(double) * 21 * n_cand * n_rungs),
               extra = a_T + sizeof(double) * 16 * n_cand * (n_rungs + 1);
  if ((rc = stage_batch(c, plans, cand_pair, n_cand, dT, dout, extra, &bd))) return rc;
  char *hp = bd.pin + bd.pin_at;
  std::memcpy(hp + a_init, inits, sizeof(double) * 16 * n_cand);
  std::memcpy(hp + a_skip, skip.data(), skip.size());
  HIPCK(hipMemcpyAsync(dT, hp + a_init, sizeof(double) * 16 * n_cand, hipMemcpyHostToDevice, c->stream), "upload");
  HIPCK(hipMemcpyAsync(dskip, hp + a_skip, skip.size(), hipMemcpyHostToDevice, c->stream), "upload");
  bool tuned = true;
  std::vector<int> blocks(n_rungs, 0), grids(n_rungs, 0);
  for (size_t r = 0; r < n_rungs; ++r) {
    blocks[r] = c->icp_block ? c->icp_block : tuned_block(c, plans[r]);
    if (!blocks[r] && bd.v[r].n_cand > 0) tuned = false;
    if (blocks[r] > WIDE) grids[r] = grid_for(c, blocks[r], plans[r].order.size(), plans[r].max_spans);
  }
  int cb = (int)((16 * n_cand + 255) / 256);
  Timer tm(c, SHERD_M_ICP, n_cand * n_rungs);
  // The launch signature (every kernel argument of every rung): a batch whose signature the last
  // call also had is captured as a graph and replayed while it repeats; a one-off batch (the
  // refinement's walk issues one per depth) runs its launches directly.
  std::string sig;
  if (tuned) {
    for (size_t r = 0; r < n_rungs; ++r) {
      sig.append((const char *)&bd.v[r], sizeof(IcpBatchView));
      sig.append((const char *)&bd.w[r], sizeof(IcpWideView));
      sig.append((const char *)&blocks[r], sizeof(int));
      sig.append((const char *)&grids[r], sizeof(int));
    }
    sig.append((const char *)&dskip, sizeof dskip);
    sig.append((const char *)&any_skip, sizeof any_skip);
  }
  // (A replayed graph's kernels are not timed by the profiler's events; a kernel trace times them.)
  bool replay = tuned && ((c->ladder_exec && sig == c->ladder_sig) || sig == c->ladder_seen);
  c->ladder_seen = sig;
  if (!replay) {
    for (size_t r = 0; r < n_rungs; ++r) {
      if ((rc = run_rung(c, bd, r, plans[r]))) return rc;
      if (any_skip)
        hipLaunchKernelGGL(k_carry, dim3(cb), dim3(256), 0, c->stream, bd.v[r].T_in, bd.v[r].T_out,
                           dskip + r * n_cand, (int)n_cand);
    }
    HIPCK(hipGetLastError(), "carry");
  } else {
    if (!c->ladder_exec || sig != c->ladder_sig) {
      if (c->ladder_exec) {
        (void)hipGraphExecDestroy(c->ladder_exec);
        c->ladder_exec = nullptr;
      }
      hipGraph_t g = nullptr;
      HIPCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
      hipError_t le = hipSuccess;
      for (size_t r = 0; r < n_rungs && le == hipSuccess; ++r) {
        if (bd.v[r].n_cand > 0) le = launch_rung(c, bd, r, blocks[r], grids[r], false);
        if (le == hipSuccess && any_skip) {
          hipLaunchKernelGGL(k_carry, dim3(cb), dim3(256), 0, c->stream, bd.v[r].T_in, bd.v[r].T_out,
                             dskip + r * n_cand, (int)n_cand);
          le = hipGetLastError();
        }
      }
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le, "icp kernel capture");
      }
      if (ce != hipSuccess) return hip_fail(ce, "graph capture");
      hipGraphExec_t ex = nullptr;
      e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return hip_fail(e, "graph instantiate");
      c->ladder_exec = ex;
      c->ladder_sig = sig;
    }
    HIPCK(hipGraphLaunch(c->ladder_exec, c->stream), "graph launch");
  }
  tm.stop();
  const double *h = (const double *)(hp + a_out), *hT = (const double *)(hp + a_T);
  HIPCK(hipMemcpyAsync(hp + a_out, dout, sizeof(double) * 21 * n_cand * n_rungs, hipMemcpyDeviceToHost, c->stream),
        "download");
  HIPCK(hipMemcpyAsync(hp + a_T, dT, sizeof(double) * 16 * n_cand * (n_rungs + 1), hipMemcpyDeviceToHost, c->stream),
        "download");
  SYNC_RECEIPT(c);
  c->prof.collect();
  for (size_t i = 0; i < n_cand; ++i)
    for (size_t r = 0; r < n_rungs; ++r) {
      sherd_registration *o = &out[i * n_rungs + r];
      if (skip[r * n_cand + i])
        unmoved(&hT[16 * n_cand * r + 16 * i], o);
      else
        to_registration(&h[21 * n_cand * r + 21 * i], o);
    }
  return SHERD_OK;
}

}  // namespace

extern "C" {

int sherd_cloud_create(sherd_ctx *ctx, const double *pts, const double *normals, size_t n, sherd_cloud **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  LOCK(c);
  if (n > 0 && !pts) return fail(SHERD_E_ARG, "null points");
  return make_cloud(c, pts, normals, n, out);
}

void sherd_cloud_free(sherd_cloud *cloud) {
  if (!cloud) return;
  std::lock_guard<std::mutex> lock(cloud->ctx->mu);
  (void)hipSetDevice(cloud->ctx->device);
  (void)hipStreamSynchronize(cloud->ctx->stream);
  free_cloud(cloud);
}

int sherd_icp_block(sherd_ctx *ctx, int threads) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (threads != 0 && threads != 1 && threads != 64 && threads != 128 && threads != 256 && threads != 512)
    return fail(SHERD_E_ARG, "workgroup size must be 0, 1 (wide), 64, 128, 256 or 512");
  c->icp_block = threads;
  return SHERD_OK;
}

int sherd_icp_batch_ladder(sherd_ctx *ctx, const sherd_icp_pair *pairs, size_t n_pairs,
                           const sherd_icp_rung_spec *rungs, size_t n_rungs, const uint32_t *cand_pair,
                           const double *inits, size_t n_cand, sherd_registration *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (n_cand == 0 || n_rungs == 0) return SHERD_OK;
  if (!pairs || !rungs || !cand_pair || !inits || !out) return fail(SHERD_E_ARG, "null buffer");
  if (n_rungs > MAX_RUNGS) return fail(SHERD_E_LIMIT, "more than 8 rungs");
  if (n_pairs > (size_t)INT_MAX || n_cand > (size_t)INT_MAX / 16) return fail(SHERD_E_LIMIT, "batch too large");
  return run_ladder(c, pairs, n_pairs, rungs, n_rungs, cand_pair, inits, n_cand, out);
}

int sherd_icp_batch_rung(sherd_ctx *ctx, const sherd_icp_pair *pairs, size_t n_pairs, const uint32_t *cand_pair,
                         const double *inits, size_t n_cand, int estimation, int max_iteration,
                         sherd_registration *out) {
  sherd_icp_rung_spec r{estimation, max_iteration};
  return sherd_icp_batch_ladder(ctx, pairs, n_pairs, &r, 1, cand_pair, inits, n_cand, out);
}

int sherd_icp_rung(sherd_ctx *ctx, const double *src, size_t n_src, const double *tgt_p, const double *tgt_n,
                   size_t n_tgt, const double *inits, size_t k, int estimation, double distance, int max_iteration,
                   sherd_registration *out) {
  if (!ctx) return fail(SHERD_E_ARG, "null context");
  if (k == 0) return SHERD_OK;
  if (!inits || !out) return fail(SHERD_E_ARG, "null buffer");
  sherd_icp_rung_spec spec{estimation, max_iteration};
  int rc = check_rung(spec);
  if (rc) return rc;
  if (n_src == 0 || n_tgt == 0) {
    for (size_t i = 0; i < k; ++i) unmoved(inits + 16 * i, &out[i]);
    return SHERD_OK;
  }
  if (!src || !tgt_p) return fail(SHERD_E_ARG, "null buffer");
  sherd_cloud *s = nullptr, *t = nullptr;
  if ((rc = sherd_cloud_create(ctx, src, nullptr, n_src, &s))) return rc;
  if ((rc = sherd_cloud_create(ctx, tgt_p, tgt_n, n_tgt, &t))) {
    sherd_cloud_free(s);
    return rc;
  }
  sherd_icp_pair pair{s, t, distance};
  std::vector<uint32_t> cp(k, 0u);
  rc = sherd_icp_batch_rung(ctx, &pair, 1, cp.data(), inits, k, estimation, max_iteration, out);
  sherd_cloud_free(s);
  sherd_cloud_free(t);
  return rc;
}

}  // extern "C"
