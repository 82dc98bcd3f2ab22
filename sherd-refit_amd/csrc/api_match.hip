// C ABI for R §5.1-5.5 over many pairs (crates/sherd-core/src/matching/pair.rs:170-264): breakline
// frame handles, the hypotheses of a pair, the greedy suppression over a given walk, and the whole
// stage-1 hand-off for a batch of pairs (hypotheses, coarse scores, order, suppression, the
// breakline ladder, its re-score and the second suppression) with every pair in the same launches.
// Host code: descriptors, the few synchronisations that size the next step, and the ladder call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"
#include "sort.hpp"

using namespace sherd;

namespace {

// pair.rs:50-52 and ladder.rs:51-61.
constexpr uint32_t COARSE_ORDER_LIMIT = 5000;
constexpr double COARSE_FLOOR = 0.1;
constexpr double STAGE1_FLOOR = 0.05;
constexpr double STAGE1_RUNGS[2] = {0.2, 0.08};
constexpr int STAGE1_ITERATIONS = 20;

// A call's device buffers, freed on scope exit (the batch sizes vary too much for grow-only slots).
struct DevBufs {
  std::vector<void *> p;
  ~DevBufs() {
    for (void *x : p) (void)hipFree(x);
  }
  template <class T>
  T *get(size_t n, hipError_t *e) {
    void *x = nullptr;
    *e = hipMalloc(&x, sizeof(T) * std::max<size_t>(n, 1));
    if (*e != hipSuccess) return nullptr;
    p.push_back(x);
    return (T *)x;
  }
  template <class T>
  T *put(const T *h, size_t n, hipStream_t s, hipError_t *e) {
    T *d = get<T>(n, e);
    if (d && n) *e = hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s);
    return *e == hipSuccess ? d : nullptr;
  }
};

void free_frames(sherd_frames *fr) {
  for (double *x : {fr->p, fr->ns, fr->f, fr->t, fr->dih, fr->sub_p, fr->sub_n})
    if (x) (void)hipFree(x);
  if (fr->sub) (void)hipFree(fr->sub);
  grid_free(&fr->grid);
  delete fr;
}

// The frames' point grid, (re)built for queries of reach `reach`.
hipError_t frames_grid(Ctx *c, sherd_frames *fr, double reach) {
  if (fr->n == 0 || (fr->grid.n > 0 && fr->reach >= reach)) return hipSuccess;
  grid_free(&fr->grid);
  fr->reach = 0.0;
  hipError_t e = grid_build(fr->p, fr->n, reach, &fr->grid, c->stream);
  if (e == hipSuccess) fr->reach = reach;
  return e;
}

FramesView view(const sherd_frames *fr) {
  FramesView v;
  v.p = fr->p;
  v.ns = fr->ns;
  v.f = fr->f;
  v.t = fr->t;
  v.dih = fr->dih;
  v.sub_p = fr->sub_p;
  v.sub_n = fr->sub_n;
  v.n = fr->n;
  v.n_sub = fr->n_sub;
  std::memset(&v.grid, 0, sizeof v.grid);
  if (fr->n > 0) v.grid = view_of(fr->grid);
  return v;
}

int check_subset(const uint32_t *s, size_t n, int bound) {
  for (size_t k = 0; k < n; ++k)
    if (s[k] >= (uint32_t)bound) return fail(SHERD_E_ARG, "subset index out of range");
  return SHERD_OK;
}

// Hypotheses of the batch's pairs (device): per pair its first hypothesis and count.
struct HypSet {
  uint32_t total = 0;
  std::vector<uint32_t> first;  // n_pairs + 1
  uint32_t *pa = nullptr, *pb = nullptr, *hpair = nullptr, *dfirst = nullptr;
  double *pose = nullptr;
};

int build_hypotheses(Ctx *c, DevBufs &db, const StagePair *dsp, const std::vector<StagePair> &sp, HypSet *hs) {
  hipError_t e;
  size_t np = sp.size();
  std::vector<uint32_t> row_first(np + 1, 0);
  for (size_t p = 0; p < np; ++p) {
    size_t next = (size_t)row_first[p] + (size_t)sp[p].nia;
    if (next > (size_t)UINT32_MAX / 2) return fail(SHERD_E_LIMIT, "too many hypothesis rows");
    row_first[p + 1] = (uint32_t)next;
  }
  uint32_t rows = row_first[np];
  uint32_t *drf = db.put(row_first.data(), np + 1, c->stream, &e);
  if (!drf) return hip_fail(e, "upload");
  uint32_t *cnt = db.get<uint32_t>((size_t)rows + 1 + rsort::scan_parts((size_t)rows + 1), &e);
  if (!cnt) return hip_fail(e, "alloc");
  HIPCK(launch_hyp_count(dsp, drf, (int)np, rows, cnt, c->stream), "hypothesis count");
  HIPCK(rsort::scan_exclusive(cnt, (size_t)rows + 1, cnt + rows + 1, c->stream), "scan");
  std::vector<uint32_t> off((size_t)rows + 1);
  HIPCK(hipMemcpyAsync(off.data(), cnt, sizeof(uint32_t) * (rows + 1), hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  hs->total = off[rows];
  hs->first.assign(np + 1, 0);
  for (size_t p = 0; p <= np; ++p) hs->first[p] = off[row_first[p]];
  size_t H = hs->total;
  if (!(hs->pa = db.get<uint32_t>(H, &e)) || !(hs->pb = db.get<uint32_t>(H, &e)) ||
      !(hs->hpair = db.get<uint32_t>(H, &e)) || !(hs->pose = db.get<double>(12 * H, &e)))
    return hip_fail(e, "alloc hypotheses");
  if (!(hs->dfirst = db.put(hs->first.data(), np + 1, c->stream, &e))) return hip_fail(e, "upload");
  HIPCK(launch_hyp_build(dsp, drf, (int)np, rows, cnt, hs->total, hs->pa, hs->pb, hs->hpair, hs->pose, c->stream),
        "hypotheses");
  return SHERD_OK;
}

// Orders items (per pair, count descending, index ascending) and runs one suppression per pair.
int order_and_suppress(Ctx *c, DevBufs &db, const uint32_t *ipair, const uint32_t *dfirst,
                       const std::vector<uint32_t> &first, const uint32_t *counts, const double *pose,
                       const double *score, uint32_t n, const std::vector<double> &tol, uint32_t limit, int topk,
                       double floor, std::vector<uint32_t> *kept, std::vector<uint32_t> *n_kept) {
  hipError_t e;
  size_t np = first.size() - 1;
  uint64_t *keys = db.get<uint64_t>(n, &e), *tk = keys ? db.get<uint64_t>(n, &e) : nullptr;
  uint32_t *vals = tk ? db.get<uint32_t>(n, &e) : nullptr, *tv = vals ? db.get<uint32_t>(n, &e) : nullptr;
  uint32_t *hist = tv ? db.get<uint32_t>(order_hist_words(n), &e) : nullptr;
  uint32_t *dkept = hist ? db.get<uint32_t>(np * (size_t)std::max(topk, 1), &e) : nullptr;
  uint32_t *dn = dkept ? db.get<uint32_t>(np, &e) : nullptr;
  if (!dn) return hip_fail(e, "alloc order");
  HIPCK(launch_order(ipair, dfirst, counts, n, keys, vals, tk, tv, hist, c->stream), "order");
  std::vector<NmsProblem> pr(np);
  for (size_t p = 0; p < np; ++p) {
    NmsProblem &q = pr[p];
    q.order = vals + first[p];
    q.n = (int)std::min(first[p + 1] - first[p], limit);
    q.pose = pose;
    q.score = score;
    q.tol = tol[p];
    q.topk = topk;
    q.floor = floor;
    q.kept = dkept + p * (size_t)std::max(topk, 1);
    q.n_kept = dn + p;
  }
  NmsProblem *dpr = db.put(pr.data(), np, c->stream, &e);
  if (!dpr) return hip_fail(e, "upload");
  HIPCK(launch_nms(dpr, (int)np, c->stream), "suppression");
  kept->resize(np * (size_t)std::max(topk, 1));
  n_kept->resize(np);
  HIPCK(hipMemcpyAsync(kept->data(), dkept, sizeof(uint32_t) * kept->size(), hipMemcpyDeviceToHost, c->stream),
        "download");
  HIPCK(hipMemcpyAsync(n_kept->data(), dn, sizeof(uint32_t) * np, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  return SHERD_OK;
}

}  // namespace

extern "C" {

int sherd_frames_create(sherd_ctx *ctx, const double *p, const double *ns, const double *f, const double *t,
                        const double *dih, size_t n, const uint32_t *sub, size_t n_sub, sherd_frames **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  LOCK(c);
  if (n && (!p || !ns || !f || !t || !dih)) return fail(SHERD_E_ARG, "null frame buffer");
  if (n_sub && !sub) return fail(SHERD_E_ARG, "null subset");
  if (n > (size_t)INT_MAX / 3 || n_sub > (size_t)INT_MAX / 3) return fail(SHERD_E_LIMIT, "breakline too large");
  int rc = check_subset(sub, n_sub, (int)n);
  if (rc) return rc;
  std::vector<double> sp(3 * n_sub), sn(3 * n_sub);
  for (size_t k = 0; k < n_sub; ++k)
    for (int a = 0; a < 3; ++a) {
      sp[3 * k + a] = p[3 * (size_t)sub[k] + a];
      sn[3 * k + a] = ns[3 * (size_t)sub[k] + a];
    }
  sherd_frames *fr = new sherd_frames();
  fr->ctx = c;
  fr->n = (int)n;
  fr->n_sub = (int)n_sub;
  hipError_t e = hipSuccess;
  auto up = [&](const double *h, size_t m, double **d) {
    if (e != hipSuccess) return;
    if ((e = hipMalloc((void **)d, sizeof(double) * std::max<size_t>(m, 1))) != hipSuccess) return;
    if (m) e = hipMemcpyAsync(*d, h, sizeof(double) * m, hipMemcpyHostToDevice, c->stream);
  };
  up(p, 3 * n, &fr->p);
  up(ns, 3 * n, &fr->ns);
  up(f, 3 * n, &fr->f);
  up(t, 3 * n, &fr->t);
  up(dih, n, &fr->dih);
  up(sp.data(), 3 * n_sub, &fr->sub_p);
  up(sn.data(), 3 * n_sub, &fr->sub_n);
  if (e == hipSuccess) e = hipMalloc((void **)&fr->sub, sizeof(uint32_t) * std::max<size_t>(n_sub, 1));
  if (e == hipSuccess && n_sub)
    e = hipMemcpyAsync(fr->sub, sub, sizeof(uint32_t) * n_sub, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free_frames(fr);
    return hip_fail(e, "frames upload");
  }
  *out = fr;
  return SHERD_OK;
}

void sherd_frames_free(sherd_frames *fr) {
  if (!fr) return;
  std::lock_guard<std::mutex> lock(fr->ctx->mu);
  (void)hipSetDevice(fr->ctx->device);
  (void)hipStreamSynchronize(fr->ctx->stream);
  free_frames(fr);
}

int sherd_hypotheses(sherd_ctx *ctx, const sherd_frames *a, const sherd_frames *b, const uint32_t *ia, size_t nia,
                     const uint32_t *ib, size_t nib, double dihedral_tol, size_t capacity, uint32_t *pa, uint32_t *pb,
                     double *poses, size_t *n) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!a || !b || !n || (nia && !ia) || (nib && !ib)) return fail(SHERD_E_ARG, "null argument");
  LOCK(c);
  if (a->ctx != c || b->ctx != c) return fail(SHERD_E_ARG, "frames of another context");
  int rc;
  if ((rc = check_subset(ia, nia, a->n)) || (rc = check_subset(ib, nib, b->n))) return rc;
  if (nia > (size_t)INT_MAX || nib > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "subset too large");
  DevBufs db;
  hipError_t e;
  StagePair s;
  std::memset(&s, 0, sizeof s);
  s.a = view(a);
  s.b = view(b);
  s.ia = db.put(ia, nia, c->stream, &e);
  s.ib = s.ia ? db.put(ib, nib, c->stream, &e) : nullptr;
  if (!s.ib) return hip_fail(e, "upload");
  s.nia = (int)nia;
  s.nib = (int)nib;
  s.dihedral_tol = dihedral_tol;
  std::vector<StagePair> sp{s};
  StagePair *dsp = db.put(sp.data(), 1, c->stream, &e);
  if (!dsp) return hip_fail(e, "upload");
  HypSet hs;
  if ((rc = build_hypotheses(c, db, dsp, sp, &hs))) return rc;
  *n = hs.total;
  if (hs.total > capacity) return fail(SHERD_E_LIMIT, "more hypotheses than the output holds");
  if (hs.total) {
    if (!pa || !pb || !poses) return fail(SHERD_E_ARG, "null output");
    HIPCK(hipMemcpyAsync(pa, hs.pa, sizeof(uint32_t) * hs.total, hipMemcpyDeviceToHost, c->stream), "download");
    HIPCK(hipMemcpyAsync(pb, hs.pb, sizeof(uint32_t) * hs.total, hipMemcpyDeviceToHost, c->stream), "download");
    HIPCK(hipMemcpyAsync(poses, hs.pose, sizeof(double) * 12 * hs.total, hipMemcpyDeviceToHost, c->stream),
          "download");
  }
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  return SHERD_OK;
}

int sherd_nms(sherd_ctx *ctx, const uint32_t *order, size_t n_order, const double *poses, const double *score,
              size_t n_items, double trans_tol, size_t topk, double floor, uint32_t *kept, size_t *n_kept) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!n_kept || (n_order && (!order || !poses || !score || !kept))) return fail(SHERD_E_ARG, "null argument");
  LOCK(c);
  if (n_order > (size_t)INT_MAX || topk > (size_t)INT_MAX / 2) return fail(SHERD_E_LIMIT, "walk too long");
  for (size_t k = 0; k < n_order; ++k)
    if (order[k] >= n_items) return fail(SHERD_E_ARG, "walk names an item out of range");
  *n_kept = 0;
  if (n_order == 0) return SHERD_OK;
  DevBufs db;
  hipError_t e;
  NmsProblem q;
  q.order = db.put(order, n_order, c->stream, &e);
  q.pose = q.order ? db.put(poses, 12 * n_items, c->stream, &e) : nullptr;
  q.score = q.pose ? db.put(score, n_items, c->stream, &e) : nullptr;
  // At most min(topk, walk) poses are kept, plus the one a topk of zero still keeps.
  size_t cap = std::min(n_order, topk + 1);
  q.kept = q.score ? db.get<uint32_t>(cap, &e) : nullptr;
  q.n_kept = q.kept ? db.get<uint32_t>(1, &e) : nullptr;
  if (!q.n_kept) return hip_fail(e, "upload");
  q.n = (int)n_order;
  q.tol = trans_tol;
  q.topk = (int)topk;
  q.floor = floor;
  NmsProblem *dq = db.put(&q, 1, c->stream, &e);
  if (!dq) return hip_fail(e, "upload");
  HIPCK(launch_nms(dq, 1, c->stream), "suppression");
  uint32_t nk = 0;
  HIPCK(hipMemcpyAsync(&nk, q.n_kept, sizeof nk, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  if (nk) HIPCK(hipMemcpy(kept, q.kept, sizeof(uint32_t) * nk, hipMemcpyDeviceToHost), "download");
  *n_kept = nk;
  return SHERD_OK;
}

int sherd_match_stage1(sherd_ctx *ctx, const sherd_match_pair *pairs, size_t n_pairs, size_t stage1, size_t stage2,
                       double dihedral_tol, sherd_match_result *res, uint32_t *kept1, double *s1_T, double *s1_score,
                       uint32_t *kept2) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  CALL_SCOPE(c);
  if (n_pairs == 0) return SHERD_OK;
  if (!pairs || !res || !kept1 || !s1_T || !s1_score || !kept2) return fail(SHERD_E_ARG, "null buffer");
  if (n_pairs > 65535) return fail(SHERD_E_LIMIT, "more than 65535 pairs in one batch");
  if (stage1 == 0 || stage1 > 4096 || stage2 == 0 || stage2 > stage1)
    return fail(SHERD_E_ARG, "stage sizes must satisfy 0 < stage2 <= stage1 <= 4096");
  const size_t np = n_pairs;
  std::vector<uint32_t> k1, n1, k2, n2;
  std::vector<double> hpose;  // kept hypotheses' poses, 12 each
  std::vector<uint32_t> cand_pair;
  std::vector<size_t> cand_first(np + 1, 0);
  DevBufs db;
  hipError_t e;
  int rc;
  std::vector<double> tol(np);
  {
    RELOCK(c);
    // Every target grid at the largest reach any pair asks of it, before any descriptor reads it.
    std::vector<std::pair<sherd_frames *, double>> reach;
    for (size_t p = 0; p < np; ++p) {
      const sherd_match_pair &m = pairs[p];
      if (!m.a || !m.b || !m.s1_source || !m.s1_target) return fail(SHERD_E_ARG, "pair without frames or clouds");
      if (m.a->ctx != c || m.b->ctx != c) return fail(SHERD_E_ARG, "frames of another context");
      if (m.n_probe && (!m.probe_p || !m.probe_n)) return fail(SHERD_E_ARG, "null probe");
      // Agreement counts are ordered in 16 bits.
      if (m.n_probe > 65535 || m.b->n_sub > 65535 || m.nia > (size_t)INT_MAX || m.nib > (size_t)INT_MAX)
        return fail(SHERD_E_LIMIT, "probe or subset beyond 65535 points");
      if ((m.ia && (rc = check_subset(m.ia, m.nia, m.a->n))) || (m.ib && (rc = check_subset(m.ib, m.nib, m.b->n))))
        return rc;
      sherd_frames *fa = const_cast<sherd_frames *>(m.a);
      double r = std::fmax(m.scales[2], m.scales[3]);
      auto it = std::find_if(reach.begin(), reach.end(), [&](const std::pair<sherd_frames *, double> &x) {
        return x.first == fa;
      });
      if (it == reach.end())
        reach.emplace_back(fa, r);
      else
        it->second = std::fmax(it->second, r);
    }
    for (auto &x : reach)
      if (x.second > 0.0) HIPCK(frames_grid(c, x.first, x.second), "frames grid");
    std::vector<StagePair> sp(np);
    for (size_t p = 0; p < np; ++p) {
      const sherd_match_pair &m = pairs[p];
      StagePair &s = sp[p];
      std::memset(&s, 0, sizeof s);
      s.a = view(m.a);
      s.b = view(m.b);
      s.ia = m.ia ? db.put(m.ia, m.nia, c->stream, &e) : m.a->sub;
      s.ib = m.ib ? db.put(m.ib, m.nib, c->stream, &e) : m.b->sub;
      if (!s.ia || !s.ib) return hip_fail(e, "upload");
      s.nia = m.ia ? (int)m.nia : m.a->n_sub;
      s.nib = m.ib ? (int)m.nib : m.b->n_sub;
      s.probe_p = db.put(m.probe_p, 3 * m.n_probe, c->stream, &e);
      s.probe_n = s.probe_p ? db.put(m.probe_n, 3 * m.n_probe, c->stream, &e) : nullptr;
      if (!s.probe_n) return hip_fail(e, "upload");
      s.n_probe = (int)m.n_probe;
      s.dihedral_tol = dihedral_tol;
      s.coarse_r = m.scales[2];
      s.stage1_r = m.scales[3];
      tol[p] = m.scales[10];
    }
    StagePair *dsp = db.put(sp.data(), np, c->stream, &e);
    if (!dsp) return hip_fail(e, "upload");
    // R §5.1-5.3: hypotheses, coarse scores, order and suppression.
    HypSet hs;
    if ((rc = build_hypotheses(c, db, dsp, sp, &hs))) return rc;
    uint32_t *cnt = db.get<uint32_t>(hs.total, &e);
    double *score = cnt ? db.get<double>(hs.total, &e) : nullptr;
    if (!score) return hip_fail(e, "alloc");
    HIPCK(launch_stage_coarse(dsp, hs.hpair, hs.pose, hs.total, 0, cnt, score, c->stream), "coarse");
    if ((rc = order_and_suppress(c, db, hs.hpair, hs.dfirst, hs.first, cnt, hs.pose, score, hs.total, tol,
                                 COARSE_ORDER_LIMIT, (int)stage1, COARSE_FLOOR, &k1, &n1)))
      return rc;
    for (size_t p = 0; p < np; ++p) {
      res[p].n_hyp = hs.first[p + 1] - hs.first[p];
      res[p].n_kept1 = n1[p];
      cand_first[p + 1] = cand_first[p] + n1[p];
    }
    size_t nc = cand_first[np];
    hpose.resize(12 * nc);
    cand_pair.resize(nc);
    std::vector<uint32_t> flat(nc);
    for (size_t p = 0; p < np; ++p)
      for (uint32_t k = 0; k < n1[p]; ++k) {
        flat[cand_first[p] + k] = k1[p * stage1 + k];
        cand_pair[cand_first[p] + k] = (uint32_t)p;
        kept1[p * stage1 + k] = k1[p * stage1 + k] - hs.first[p];
      }
    // The kept hypotheses' poses, gathered on the device.
    if (nc) {
      uint32_t *dflat = db.put(flat.data(), nc, c->stream, &e);
      double *dg = dflat ? db.get<double>(12 * nc, &e) : nullptr;
      if (!dg) return hip_fail(e, "upload");
      HIPCK(launch_gather12(hs.pose, dflat, (uint32_t)nc, dg, c->stream), "gather");
      HIPCK(hipMemcpyAsync(hpose.data(), dg, sizeof(double) * 12 * nc, hipMemcpyDeviceToHost, c->stream), "download");
    }
    HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  }
  // R §5.4: the breakline ladder over every kept hypothesis of every pair.
  size_t nc = cand_first[np];
  std::vector<double> T(16 * nc, 0.0);
  std::vector<double> s1(nc, 0.0);
  if (nc) {
    std::vector<double> inits(16 * nc);
    for (size_t i = 0; i < nc; ++i) {
      const double *h = &hpose[12 * i];
      double *m = &inits[16 * i];
      for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) m[4 * r + q] = h[3 * r + q];
        m[4 * r + 3] = h[9 + r];
      }
      m[12] = m[13] = m[14] = 0.0;
      m[15] = 1.0;
    }
    std::vector<sherd_icp_pair> ip(2 * np);
    for (int r = 0; r < 2; ++r)
      for (size_t p = 0; p < np; ++p)
        ip[r * np + p] = sherd_icp_pair{pairs[p].s1_source, pairs[p].s1_target,
                                        STAGE1_RUNGS[r] * pairs[p].scales[0] * pairs[p].scales[11]};
    sherd_icp_rung_spec rungs[2] = {{SHERD_POINT_TO_POINT, STAGE1_ITERATIONS}, {SHERD_POINT_TO_POINT, STAGE1_ITERATIONS}};
    std::vector<sherd_registration> reg(2 * nc);
    if ((rc = sherd_icp_batch_ladder(ctx, ip.data(), np, rungs, 2, cand_pair.data(), inits.data(), nc, reg.data())))
      return rc;
    for (size_t i = 0; i < nc; ++i) std::memcpy(&T[16 * i], reg[2 * i + 1].transform, sizeof(double) * 16);
  }
  {
    RELOCK(c);
    std::vector<StagePair> sp(np);
    for (size_t p = 0; p < np; ++p) {
      StagePair &s = sp[p];
      std::memset(&s, 0, sizeof s);
      s.a = view(pairs[p].a);
      s.b = view(pairs[p].b);
      s.stage1_r = pairs[p].scales[3];
    }
    StagePair *dsp = db.put(sp.data(), np, c->stream, &e);
    if (!dsp) return hip_fail(e, "upload");
    std::vector<uint32_t> first(np + 1);
    for (size_t p = 0; p <= np; ++p) first[p] = (uint32_t)cand_first[p];
    double *dT = db.put(T.data(), 16 * nc, c->stream, &e);
    double *pose = dT ? db.get<double>(12 * nc, &e) : nullptr;
    uint32_t *dcp = pose ? db.put(cand_pair.data(), nc, c->stream, &e) : nullptr;
    uint32_t *dfirst = dcp ? db.put(first.data(), np + 1, c->stream, &e) : nullptr;
    uint32_t *cnt = dfirst ? db.get<uint32_t>(nc, &e) : nullptr;
    double *score = cnt ? db.get<double>(nc, &e) : nullptr;
    if (!score) return hip_fail(e, "upload");
    HIPCK(launch_pose12(dT, (uint32_t)nc, pose, c->stream), "poses");
    // R §5.4's re-score (ladder.rs:177-186), then R §5.5's suppression over the stage-1 poses.
    HIPCK(launch_stage_coarse(dsp, dcp, pose, (uint32_t)nc, 1, cnt, score, c->stream), "re-score");
    if ((rc = order_and_suppress(c, db, dcp, dfirst, first, cnt, pose, score, (uint32_t)nc, tol, UINT32_MAX,
                                 (int)stage2, STAGE1_FLOOR, &k2, &n2)))
      return rc;
    if (nc) HIPCK(hipMemcpyAsync(s1.data(), score, sizeof(double) * nc, hipMemcpyDeviceToHost, c->stream), "download");
    SYNC_RECEIPT(c);
  }
  for (size_t p = 0; p < np; ++p) {
    res[p].n_kept2 = n2[p];
    res[p].pad = 0;
    double best = 0.0;
    for (uint32_t k = 0; k < n1[p]; ++k) {
      size_t i = cand_first[p] + k;
      std::memcpy(&s1_T[16 * (p * stage1 + k)], &T[16 * i], sizeof(double) * 16);
      s1_score[p * stage1 + k] = s1[i];
      best = k == 0 ? s1[i] : std::fmax(best, s1[i]);
    }
    res[p].best1 = best;
    for (uint32_t k = 0; k < n2[p]; ++k) kept2[p * stage2 + k] = k2[p * stage2 + k] - (uint32_t)cand_first[p];
  }
  return SHERD_OK;
}

}  // extern "C"
