// C ABI for the R §6 verification: fragment handles and the batched verify entry point.
// Host code only: descriptors, uploads, chunking and argument checks; the statistics run in
// verify.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstring>
#include <limits>
#include <utility>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"

using namespace sherd;

namespace {

// Copies n x 3 f64 to a fresh device allocation (nullptr for n == 0).
hipError_t upload_new(const double *h, size_t n, double **d, hipStream_t s) {
  *d = nullptr;
  if (n == 0) return hipSuccess;
  hipError_t e = hipMalloc((void **)d, sizeof(double) * 3 * n);
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(*d, h, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s);
}

void free_fragment(sherd_fragment *f) {
  double *ptrs[] = {f->pf, f->s, f->brk_p, f->brk_ns, f->margin_p, f->margin_n, f->pf_blk, f->s_blk};
  for (double *p : ptrs)
    if (p) (void)hipFree(p);
  grid_free(&f->margin_grid);
  grid_free(&f->brk_grid);
  delete f;
}

// The points of a sample set in Morton order of their own bounding box (10 bits per axis), so
// that the lanes of a wave query neighbouring places: their traversals share nodes, and the
// wave-coherent parity walk visits little more than one ray's nodes. Only sets whose statistics
// are order-free (counts, extremes, key multisets: Pf and S) are reordered.
std::vector<double> morton_ordered(const double *h, size_t n) {
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::numeric_limits<double>::infinity();
    hi[k] = -std::numeric_limits<double>::infinity();
  }
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      double v = h[3 * i + k];
      if (v < lo[k]) lo[k] = v;
      if (v > hi[k]) hi[k] = v;
    }
  double scale[3];
  for (int k = 0; k < 3; ++k) scale[k] = hi[k] > lo[k] ? 1023.0 / (hi[k] - lo[k]) : 0.0;
  auto spread = [](uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  };
  std::vector<std::pair<uint32_t, uint32_t>> order(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
      double q = (h[3 * i + k] - lo[k]) * scale[k];
      uint32_t b = q >= 1023.0 ? 1023u : (q > 0.0 ? (uint32_t)q : 0u);  // NaN lands on 0
      code |= spread(b) << (2 - k);
    }
    order[i] = {code, (uint32_t)i};
  }
  std::sort(order.begin(), order.end());
  std::vector<double> out(3 * n);
  for (size_t j = 0; j < n; ++j) std::memcpy(&out[3 * j], h + 3 * (size_t)order[j].second, 3 * sizeof(double));
  return out;
}

// The bounding sphere of every block of `per` consecutive points (x, y, z, r): the centre of the
// block's box and the largest distance from it to a point of the block, stepped up two ulps so the
// radius is never below the exact one.
std::vector<double> block_spheres(const std::vector<double> &p, size_t n, size_t per = 64) {
  size_t nb = (n + per - 1) / per;
  std::vector<double> out(4 * nb);
  for (size_t b = 0; b < nb; ++b) {
    size_t s = per * b, e = std::min(n, s + per);
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::numeric_limits<double>::infinity();
      hi[k] = -std::numeric_limits<double>::infinity();
    }
    for (size_t i = s; i < e; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], p[3 * i + k]);
        hi[k] = std::max(hi[k], p[3 * i + k]);
      }
    double c[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    double r2 = 0.0;
    for (size_t i = s; i < e; ++i) {
      double dx = p[3 * i] - c[0], dy = p[3 * i + 1] - c[1], dz = p[3 * i + 2] - c[2];
      r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
    }
    double r = std::sqrt(r2);
    r = std::nextafter(std::nextafter(r, INFINITY), INFINITY);
    out[4 * b] = c[0];
    out[4 * b + 1] = c[1];
    out[4 * b + 2] = c[2];
    out[4 * b + 3] = std::isfinite(r) ? r : INFINITY;
  }
  return out;
}

hipError_t upload_doubles(const std::vector<double> &h, double **d, hipStream_t s) {
  *d = nullptr;
  if (h.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void **)d, sizeof(double) * h.size());
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(*d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s);
}

uint32_t pow2_at_least(uint64_t x) {
  uint32_t c = 2;
  while (c < x) c <<= 1;
  return c;
}

}  // namespace

extern "C" {

int sherd_fragment_create(sherd_ctx *ctx, const sherd_scene *fracture, const sherd_scene *mesh, int watertight,
                          double frac_area, const double *pf, size_t n_pf, const double *s, size_t n_s,
                          const double *brk_p, const double *brk_ns, size_t n_brk, const double *margin_p,
                          const double *margin_n, size_t n_margin, sherd_fragment **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  LOCK(c);
  if (!fracture || fracture->ctx != c || (mesh && mesh->ctx != c)) return fail(SHERD_E_ARG, "scene of another context");
  if ((n_pf && !pf) || (n_s && !s) || (n_brk && (!brk_p || !brk_ns)) || (n_margin && (!margin_p || !margin_n)))
    return fail(SHERD_E_ARG, "null sample array");
  const size_t LIM = (size_t)1 << 28;
  if (n_pf > LIM || n_s > LIM || n_brk > LIM || n_margin > LIM) return fail(SHERD_E_LIMIT, "sample set too large");
  if (mesh) {
    // R §6.4 against this mesh: its parity rays read the parity grids, its closest approach
    // prunes with the distance bounds.
    DevScene *md = &const_cast<sherd_scene *>(mesh)->d;
    hipError_t e = scene_parity_grids(md, c->stream);
    if (e == hipErrorInvalidValue) return fail(SHERD_E_LIMIT, "parity grid beyond 2^31 entries");
    if (e != hipSuccess) return hip_fail(e, "parity grids");
    if ((e = scene_distance_bounds(md, c->stream)) != hipSuccess) return hip_fail(e, "distance bounds");
  }
  sherd_fragment *f = new sherd_fragment();
  f->ctx = c;
  f->frac = fracture;
  f->mesh = mesh;
  f->watertight = watertight ? 1 : 0;
  f->frac_area = frac_area;
  f->n_pf = (int)n_pf;
  f->n_s = (int)n_s;
  f->n_brk = (int)n_brk;
  f->n_margin = (int)n_margin;
  std::vector<double> pf_sorted = morton_ordered(pf, n_pf), s_sorted = morton_ordered(s, n_s);
  hipError_t e = upload_new(pf_sorted.data(), n_pf, &f->pf, c->stream);
  if (e == hipSuccess) e = upload_new(s_sorted.data(), n_s, &f->s, c->stream);
  if (e == hipSuccess) e = upload_new(brk_p, n_brk, &f->brk_p, c->stream);
  if (e == hipSuccess) e = upload_new(brk_ns, n_brk, &f->brk_ns, c->stream);
  if (e == hipSuccess) e = upload_new(margin_p, n_margin, &f->margin_p, c->stream);
  if (e == hipSuccess) e = upload_new(margin_n, n_margin, &f->margin_n, c->stream);
  std::vector<double> pf_blk = block_spheres(pf_sorted, n_pf), s_blk = block_spheres(s_sorted, n_s);
  if (n_pf) std::memcpy(f->pf_all, block_spheres(pf_sorted, n_pf, n_pf).data(), sizeof f->pf_all);
  if (n_s) std::memcpy(f->s_all, block_spheres(s_sorted, n_s, n_s).data(), sizeof f->s_all);
  if (n_brk)
    std::memcpy(f->brk_all, block_spheres(std::vector<double>(brk_p, brk_p + 3 * n_brk), n_brk, n_brk).data(),
                sizeof f->brk_all);
  if (n_margin)
    std::memcpy(f->margin_all,
                block_spheres(std::vector<double>(margin_p, margin_p + 3 * n_margin), n_margin, n_margin).data(),
                sizeof f->margin_all);
  if (e == hipSuccess) e = upload_doubles(pf_blk, &f->pf_blk, c->stream);
  if (e == hipSuccess) e = upload_doubles(s_blk, &f->s_blk, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free_fragment(f);
    return hip_fail(e, "fragment upload");
  }
  *out = f;
  return SHERD_OK;
}

void sherd_fragment_free(sherd_fragment *frag) {
  if (!frag) return;
  std::lock_guard<std::mutex> lock(frag->ctx->mu);
  (void)hipSetDevice(frag->ctx->device);
  (void)hipStreamSynchronize(frag->ctx->stream);
  free_fragment(frag);
}

int sherd_verify_batch(sherd_ctx *ctx, const sherd_verify_pair *pairs, size_t n_pairs, const uint32_t *cand_pair,
                       const double *cand_T, size_t n_cand, int full, int retest, double *scores,
                       uint8_t *accepted) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return fail(SHERD_E_ARG, "null context");
  LOCK(c);
  if (n_cand == 0) return SHERD_OK;
  if (!pairs || !cand_pair || !cand_T || !scores || !accepted) return fail(SHERD_E_ARG, "null buffer");
  if (n_pairs > (size_t)INT_MAX || n_cand > (size_t)INT_MAX / 4) return fail(SHERD_E_LIMIT, "batch too large");
  // Fragments of the batch, and the reach each one's margin grid must cover (its largest `near`
  // as side A, verify.rs:456-462) and its breakline grid (its largest `seam` as side B).
  std::map<const sherd_fragment *, int> fidx;
  std::vector<sherd_fragment *> frags;
  std::vector<double> need, need_brk;
  for (size_t p = 0; p < n_pairs; ++p) {
    const sherd_fragment *ab[2] = {pairs[p].a, pairs[p].b};
    for (int k = 0; k < 2; ++k) {
      if (!ab[k] || ab[k]->ctx != c) return fail(SHERD_E_ARG, "fragment missing or of another context");
      if (!fidx.count(ab[k])) {
        fidx[ab[k]] = (int)frags.size();
        frags.push_back(const_cast<sherd_fragment *>(ab[k]));
        need.push_back(0.0);
        need_brk.push_back(0.0);
      }
    }
    double nr = pairs[p].scales[8];
    int ia = fidx[pairs[p].a];
    if (nr > need[ia]) need[ia] = nr;
    double sr = pairs[p].scales[7];
    int ib = fidx[pairs[p].b];
    if (sr > need_brk[ib]) need_brk[ib] = sr;
  }
  for (size_t i = 0; i < n_cand; ++i)
    if (cand_pair[i] >= n_pairs) return fail(SHERD_E_ARG, "candidate names a pair out of range");
  int max_pf = 0, max_s = 0, max_margin = 0, max_brk = 0;
  std::vector<FragDesc> fd(frags.size());
  for (size_t i = 0; i < frags.size(); ++i) {
    sherd_fragment *f = frags[i];
    if (f->n_margin > 0 && need[i] > 0.0 && (f->margin_grid.pts == nullptr || f->margin_reach < need[i])) {
      grid_free(&f->margin_grid);
      HIPCK(grid_build(f->margin_p, f->n_margin, need[i], &f->margin_grid, c->stream), "margin grid");
      f->margin_reach = need[i];
    }
    // The breakline grid is queried at A's points taken into B's frame (T^-1 p), which differ from
    // the exact preimage by rounding: the reach carries a relative margin far above it.
    double brk_need = need_brk[i] * (1.0 + 1e-6);
    if (f->n_brk > 0 && need_brk[i] > 0.0 && (f->brk_grid.pts == nullptr || f->brk_reach < brk_need)) {
      grid_free(&f->brk_grid);
      HIPCK(grid_build(f->brk_p, f->n_brk, brk_need, &f->brk_grid, c->stream), "breakline grid");
      f->brk_reach = brk_need;
    }
    FragDesc &d = fd[i];
    std::memset(&d, 0, sizeof d);
    d.frac = view_of(f->frac->d);
    d.has_mesh = f->mesh != nullptr;
    if (f->mesh) d.mesh = view_of(f->mesh->d);
    d.watertight = f->watertight;
    d.frac_area = f->frac_area;
    d.pf = f->pf;
    d.s = f->s;
    d.pf_blk = f->pf_blk;
    d.s_blk = f->s_blk;
    std::memcpy(d.pf_all, f->pf_all, sizeof d.pf_all);
    std::memcpy(d.s_all, f->s_all, sizeof d.s_all);
    std::memcpy(d.brk_all, f->brk_all, sizeof d.brk_all);
    std::memcpy(d.margin_all, f->margin_all, sizeof d.margin_all);
    d.brk_p = f->brk_p;
    d.brk_ns = f->brk_ns;
    d.margin_p = f->margin_p;
    d.margin_n = f->margin_n;
    d.n_pf = f->n_pf;
    d.n_s = f->n_s;
    d.n_brk = f->n_brk;
    d.n_margin = f->n_margin;
    d.margin = view_of(f->margin_grid);
    d.brk = view_of(f->brk_grid);
    max_pf = std::max(max_pf, f->n_pf);
    max_s = std::max(max_s, f->n_s);
    max_margin = std::max(max_margin, f->n_margin);
    max_brk = std::max(max_brk, f->n_brk);
  }
  std::vector<PairDesc> pd(n_pairs);
  for (size_t p = 0; p < n_pairs; ++p) {
    const sherd_fragment *a = pairs[p].a, *b = pairs[p].b;
    pd[p].a = fidx[a];
    pd[p].b = fidx[b];
    pd[p].pen_ok = (a->mesh && b->mesh && a->watertight && b->watertight) ? 1 : 0;
    for (int k = 0; k < 12; ++k) pd[p].sc[k] = pairs[p].scales[k];
  }
  hipError_t e;
  // Fragment and pair descriptors go up together from the front of the pinned staging buffer; the
  // chunks stage after them (stage0), so a chunk never writes what a copy may still be reading.
  auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t fd_bytes = up256(sizeof(FragDesc) * fd.size()), stage0 = fd_bytes + up256(sizeof(PairDesc) * pd.size());
  char *dfp = (char *)c->scratch(10, stage0, &e);
  if (!dfp) return hip_fail(e, "scratch");
  char *h0 = (char *)c->pinned(stage0, &e);
  if (!h0) return hip_fail(e, "pinned staging");
  std::memcpy(h0, fd.data(), sizeof(FragDesc) * fd.size());
  std::memcpy(h0 + fd_bytes, pd.data(), sizeof(PairDesc) * pd.size());
  HIPCK(hipMemcpyAsync(dfp, h0, stage0, hipMemcpyHostToDevice, c->stream), "upload");
  FragDesc *dfd = (FragDesc *)dfp;
  PairDesc *dpd = (PairDesc *)(dfp + fd_bytes);
  Timer tm(c, SHERD_M_VERIFY, n_cand);
  // Chunks bounded by the key and voxel-table scratch they need: up to 2 GiB of keys and 2 GiB of
  // closest-approach queue per chunk, so a collection of thousands of candidates runs as one chunk
  // (one pass of every kernel) on a 288 GB device.
  const size_t KEY_BUDGET = (size_t)1 << 28;
  const size_t HARD_BUDGET = (size_t)1 << 28;  // closest-approach queue entries
  // Workgroups per (candidate, side) task of the per-point kernels; a chunk keeps its grid in int.
  const size_t per_task = (size_t)std::max({max_pf, max_s, max_margin, max_brk, 1}) / 256 + 1;
  // Candidates per chunk: a grid's y extent (65535) and x extent in int.
  const size_t MAX_TASK_CAND = std::min<size_t>(65535, (size_t)INT_MAX / (2 * per_task));
  size_t begin = 0;
  uint64_t rays = 0;  // parity rays in the unit of SHERD_M_VERIFY_RAYS
  std::vector<uint32_t> task, qfrag, first;
  std::vector<CandDesc> cd;
  while (begin < n_cand) {
    cd.clear();
    size_t keys = 0, table = 0;
    size_t end = begin;
    while (end < n_cand) {
      const PairDesc &p = pd[cand_pair[end]];
      const FragDesc &A = fd[p.a], &B = fd[p.b];
      size_t need_keys = (size_t)A.n_pf + B.n_pf + 2 * (size_t)B.n_margin;
      if (end > begin && (keys + need_keys > KEY_BUDGET || end - begin >= MAX_TASK_CAND ||
                          (end - begin + 1) * 2 * (size_t)max_s > HARD_BUDGET))
        break;
      CandDesc x;
      std::memset(&x, 0, sizeof x);
      x.pair = (int)cand_pair[end];
      x.seg[0] = keys;
      x.seg[1] = keys + A.n_pf;
      x.seg[2] = keys + A.n_pf + B.n_pf;
      x.seg[3] = x.seg[2] + B.n_margin;
      keys += need_keys;
      uint32_t cap = pow2_at_least(2 * (uint64_t)(A.n_brk > 0 ? A.n_brk : 1));
      x.seam_mask = cap - 1;
      x.seam_off = table;
      table += cap;
      std::memcpy(x.T, cand_T + 16 * end, sizeof x.T);
      cd.push_back(x);
      ++end;
    }
    size_t nc = cd.size();
    // (candidate, side) tasks grouped by the fragment they query (side 0 queries B, side 1 A), in
    // task order within a fragment: a counting sort over the fragment indices.
    task.resize(2 * nc);
    qfrag.resize(2 * nc);
    first.assign(frags.size() + 1, 0);
    for (size_t t = 0; t < 2 * nc; ++t) {
      const PairDesc &p = pd[cd[t >> 1].pair];
      qfrag[t] = (uint32_t)((t & 1u) ? p.a : p.b);
      ++first[qfrag[t] + 1];
    }
    for (size_t f = 0; f < frags.size(); ++f) first[f + 1] += first[f];
    for (size_t t = 0; t < 2 * nc; ++t) task[first[qfrag[t]]++] = (uint32_t)t;
    // What goes up (descriptors, then tasks) and what comes back (accumulators, scores, verdicts)
    // each sit in one device piece, so one pinned copy moves each way.
    const size_t up_cd = 0, up_task = up256(sizeof(CandDesc) * nc), up_bytes = up_task + sizeof(uint32_t) * 2 * nc;
    const size_t dn_acc = 0, dn_sc = up256(sizeof(CandAcc) * nc), dn_ac = dn_sc + up256(sizeof(double) * VS_COUNT * nc),
                 dn_bytes = dn_ac + nc;
    char *dup = (char *)c->scratch(12, up_bytes, &e);
    if (!dup) return hip_fail(e, "scratch");
    char *ddn = (char *)c->scratch(13, dn_bytes, &e);
    if (!ddn) return hip_fail(e, "scratch");
    // (Growing the buffer waits for the stream first, so the descriptors' copy has finished.)
    char *hpin = (char *)c->pinned(stage0 + std::max(up_bytes, dn_bytes), &e);
    if (!hpin) return hip_fail(e, "pinned staging");
    char *hbuf = hpin + stage0;
    CandDesc *dcd = (CandDesc *)(dup + up_cd);
    uint32_t *dtask = (uint32_t *)(dup + up_task);
    CandAcc *dacc = (CandAcc *)(ddn + dn_acc);
    double *dsc = (double *)(ddn + dn_sc);
    uint8_t *dac = (uint8_t *)(ddn + dn_ac);
    uint64_t *dkeys = (uint64_t *)c->scratch(14, sizeof(uint64_t) * (keys ? keys : 1), &e);
    if (!dkeys) return hip_fail(e, "scratch keys");
    uint64_t *dtab = (uint64_t *)c->scratch(15, sizeof(uint64_t) * table, &e);
    if (!dtab) return hip_fail(e, "scratch seam tables");
    uint64_t *dhard = (uint64_t *)c->scratch(19, sizeof(uint64_t) * 2 * nc * (size_t)std::max(max_s, 1), &e);
    if (!dhard) return hip_fail(e, "scratch");
    uint64_t *dhard_a = (uint64_t *)c->scratch(18, sizeof(uint64_t) * 2 * nc * (size_t)std::max(max_s, 1), &e);
    if (!dhard_a) return hip_fail(e, "scratch");
    uint32_t *dhard_n = (uint32_t *)c->scratch(20, sizeof(uint32_t) * 8, &e);
    if (!dhard_n) return hip_fail(e, "scratch");
    // The in-box samples' list and the inside samples': at most every sample of every (candidate,
    // side) task.
    uint32_t *dboxq = (uint32_t *)c->scratch(9, sizeof(uint32_t) * 2 * nc * (size_t)std::max(max_s, 1), &e);
    if (!dboxq) return hip_fail(e, "scratch");
    uint64_t *dboxw = (uint64_t *)c->scratch(11, sizeof(uint64_t) * 2 * nc * (((size_t)std::max(max_s, 1) + 63) / 64), &e);
    if (!dboxw) return hip_fail(e, "scratch");
    uint32_t *dgstart = (uint32_t *)c->scratch(21, sizeof(uint32_t) * (2 * (size_t)nc + 1), &e);
    if (!dgstart) return hip_fail(e, "scratch");
    uint64_t *ddeep = (uint64_t *)c->scratch(16, sizeof(uint64_t) * 2 * nc * (size_t)std::max(max_s, 1), &e);
    if (!ddeep) return hip_fail(e, "scratch");
    float *ddeep_t = (float *)c->scratch(17, sizeof(float) * 2 * nc * (size_t)std::max(max_s, 1), &e);
    if (!ddeep_t) return hip_fail(e, "scratch");
    // The staging buffer is free: the previous chunk's download was waited for.
    std::memcpy(hbuf + up_cd, cd.data(), sizeof(CandDesc) * nc);
    std::memcpy(hbuf + up_task, task.data(), sizeof(uint32_t) * 2 * nc);
    HIPCK(hipMemcpyAsync(dup, hbuf, up_bytes, hipMemcpyHostToDevice, c->stream), "upload");
    VerifyBatch vb;
    vb.frag = dfd;
    vb.pair = dpd;
    vb.cand = dcd;
    vb.acc = dacc;
    vb.keys = dkeys;
    vb.seam_table = dtab;
    vb.scores = dsc;
    vb.accepted = dac;
    vb.task = dtask;
    vb.hard = dhard;
    vb.hard_a = dhard_a;
    vb.hard_n = dhard_n;
    vb.boxq = dboxq;
    vb.boxw = dboxw;
    vb.gstart = dgstart;
    vb.box_cap = std::max(max_s, 1);
    vb.deep = ddeep;
    vb.deep_t = ddeep_t;
    vb.n_cand = (int)nc;
    vb.full = full ? 1 : 0;
    vb.retest = retest ? 1 : 0;
    HIPCK(launch_verify(vb, max_pf, max_s, max_margin, max_brk, c->stream, &c->side, &c->prof), "verify kernels");
    // Stream order puts the download after the upload has read the staging buffer.
    HIPCK(hipMemcpyAsync(hbuf, ddn, dn_bytes, hipMemcpyDeviceToHost, c->stream), "download");
    SYNC_RECEIPT(c);
    c->prof.collect();
    const CandAcc *hacc = (const CandAcc *)(hbuf + dn_acc);
    for (size_t i = 0; i < nc; ++i) {
      if (hacc[i].flags & FLAG_VOXEL_RANGE) return fail(SHERD_E_LIMIT, "a seam voxel coordinate exceeds 2^20");
      if (hacc[i].flags & FLAG_GAVE_UP) return fail(SHERD_E_LIMIT, "a closest-approach query exceeded its step budget");
    }
    for (size_t i = 0; i < nc; ++i) rays += 3 * ((uint64_t)hacc[i].in_box[0] + hacc[i].in_box[1]);
    std::memcpy(scores + VS_COUNT * begin, hbuf + dn_sc, sizeof(double) * VS_COUNT * nc);
    std::memcpy(accepted + begin, hbuf + dn_ac, nc);
    begin = end;
  }
  tm.stop();
  if (full) {
    c->stats[SHERD_M_VERIFY_RAYS].calls += 1;
    c->stats[SHERD_M_VERIFY_RAYS].items += rays;
  }
  return SHERD_OK;
}

}  // extern "C"
