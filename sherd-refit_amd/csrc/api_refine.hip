// C ABI for R §9, the full-resolution refinement (crates/sherd-core/src/refine.rs): fracture clouds
// of the original meshes, their placement through a pose, and the spanning walk over the assembly's
// joins with the two point-to-plane rungs of every join on the device. Host code: validation, the
// capped draw (a partial Fisher-Yates over ChaCha8, as crates/sherd-core/src/rng.rs draws it), the
// walk's order and its batching.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"

using namespace sherd;

namespace {

// Scratch slots of this file.
enum { SL_SEL = 30, SL_POS = 31, SL_FN = 32, SL_NSCR = 33, SL_IDX = 34, SL_VPTS = 36, SL_VNRM = 37, SL_VGOOD = 38,
       SL_VFRAC = 39 };

constexpr int MAX_RUNG_POINTS = ICP_SPAN * ICP_MAX_SPANS;

// ChaCha with 8 rounds, key from a u64 by PCG32 (rand_core's SeedableRng::seed_from_u64), 64-bit
// block counter from 0, stream 0; words in block order (rand_chacha 0.10 ChaCha8Rng, the
// generator crates/sherd-core/src/rng.rs:96-98 seeds).
class ChaCha8 {
 public:
  explicit ChaCha8(uint64_t state) {
    for (int k = 0; k < 8; ++k) {
      state = state * 6364136223846793005ull + 11634580027462260723ull;
      uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
      uint32_t rot = (uint32_t)(state >> 59);
      key_[k] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
  }
  uint32_t next_u32() {
    if (pos_ == 16) refill();
    return buf_[pos_++];
  }

 private:
  static uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
  static void quarter(uint32_t *s, int a, int b, int c, int d) {
    s[a] += s[b];
    s[d] = rotl(s[d] ^ s[a], 16);
    s[c] += s[d];
    s[b] = rotl(s[b] ^ s[c], 12);
    s[a] += s[b];
    s[d] = rotl(s[d] ^ s[a], 8);
    s[c] += s[d];
    s[b] = rotl(s[b] ^ s[c], 7);
  }
  void refill() {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int k = 0; k < 8; ++k) in[4 + k] = key_[k];
    in[12] = (uint32_t)counter_;
    in[13] = (uint32_t)(counter_ >> 32);
    in[14] = in[15] = 0u;
    uint32_t x[16];
    std::memcpy(x, in, sizeof x);
    for (int r = 0; r < 4; ++r) {
      quarter(x, 0, 4, 8, 12);
      quarter(x, 1, 5, 9, 13);
      quarter(x, 2, 6, 10, 14);
      quarter(x, 3, 7, 11, 15);
      quarter(x, 0, 5, 10, 15);
      quarter(x, 1, 6, 11, 12);
      quarter(x, 2, 7, 8, 13);
      quarter(x, 3, 4, 9, 14);
    }
    for (int k = 0; k < 16; ++k) buf_[k] = x[k] + in[k];
    ++counter_;
    pos_ = 0;
  }
  uint32_t key_[8];
  uint64_t counter_ = 0;
  uint32_t buf_[16];
  int pos_ = 16;
};

// rng.rs:140-156 uniform_below: Lemire's multiply-shift with rejection.
uint32_t uniform_below(ChaCha8 &g, uint32_t n) {
  uint64_t m = (uint64_t)g.next_u32() * n;
  uint32_t low = (uint32_t)m;
  if (low < n) {
    uint32_t threshold = (0u - n) % n;
    while (low < threshold) {
      m = (uint64_t)g.next_u32() * n;
      low = (uint32_t)m;
    }
  }
  return (uint32_t)(m >> 32);
}

// refine.rs:186-192 cap_selection over rng.rs:127-137 without_replacement.
void cap_selection(std::vector<uint32_t> &idx, size_t max_points, uint64_t seed) {
  size_t n = idx.size();
  if (n <= max_points) return;
  ChaCha8 g(seed);
  std::vector<uint32_t> pool(n);
  for (size_t i = 0; i < n; ++i) pool[i] = (uint32_t)i;
  for (size_t i = 0; i < max_points; ++i) {
    size_t j = i + uniform_below(g, (uint32_t)(n - i));
    std::swap(pool[i], pool[j]);
  }
  std::vector<uint32_t> out(max_points);
  for (size_t i = 0; i < max_points; ++i) out[i] = idx[pool[i]];
  idx.swap(out);
}

int check_mesh(const double *V, size_t nv, const uint32_t *F, size_t nf) {
  if ((nv && !V) || (nf && !F)) return fail(SHERD_E_ARG, "null mesh buffer");
  if (nv > (size_t)INT_MAX / 3 || nf > (size_t)INT_MAX / 3) return fail(SHERD_E_LIMIT, "mesh too large");
  for (size_t k = 0; k < 3 * nf; ++k)
    if (F[k] >= nv) return fail(SHERD_E_ARG, "face index out of range");
  return SHERD_OK;
}

// The device mesh of a call: vertices and faces uploaded, freed on scope exit.
struct DevMesh {
  double *V = nullptr;
  uint32_t *F = nullptr;
  ~DevMesh() {
    if (V) (void)hipFree(V);
    if (F) (void)hipFree(F);
  }
  hipError_t upload(Ctx *c, const double *hV, size_t nv, const uint32_t *hF, size_t nf) {
    hipError_t e = hipMalloc((void **)&V, sizeof(double) * 3 * std::max<size_t>(nv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&F, sizeof(uint32_t) * 3 * std::max<size_t>(nf, 1));
    if (e == hipSuccess && nv) e = hipMemcpyAsync(V, hV, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && nf) e = hipMemcpyAsync(F, hF, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, c->stream);
    return e;
  }
};

// The cloud of the vertices idx of a mesh already on the device: points, vertex normals, no tree.
int vertex_cloud(Ctx *c, const DevMesh &m, size_t nv, size_t nf, std::vector<uint32_t> idx, sherd_cloud **out) {
  size_t n = idx.size();
  if (n > (size_t)MAX_RUNG_POINTS) return fail(SHERD_E_LIMIT, "cloud beyond 262144 points");
  hipError_t e;
  uint32_t *didx = (uint32_t *)c->scratch(SL_IDX, sizeof(uint32_t) * std::max<size_t>(n, 1), &e);
  if (!didx) return hip_fail(e, "scratch");
  double *fn = (double *)c->scratch(SL_FN, sizeof(double) * 3 * std::max<size_t>(nf, 1), &e);
  if (!fn) return hip_fail(e, "scratch");
  uint32_t *ns = (uint32_t *)c->scratch(SL_NSCR, sizeof(uint32_t) * normals_scratch_words((int)nv, (int)nf), &e);
  if (!ns) return hip_fail(e, "scratch");
  double *pts = nullptr, *nrm = nullptr;
  e = hipMalloc((void **)&pts, sizeof(double) * 3 * std::max<size_t>(n, 1));
  if (e == hipSuccess) e = hipMalloc((void **)&nrm, sizeof(double) * 3 * std::max<size_t>(n, 1));
  if (e == hipSuccess && n)
    e = hipMemcpyAsync(didx, idx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = launch_vertex_cloud(m.V, (int)nv, m.F, (int)nf, didx, (int)n, fn, ns, pts, nrm, c->stream);
  if (e != hipSuccess) {
    if (pts) (void)hipFree(pts);
    if (nrm) (void)hipFree(nrm);
    return hip_fail(e, "vertex cloud");
  }
  int rc = cloud_adopt(c, pts, nrm, n, false, out);
  if (rc == SHERD_OK) (*out)->idx.swap(idx);
  return rc;
}

// refine.rs:163-182 select_candidates on the device: the kept vertices in ascending order.
int select_vertices(Ctx *c, const DevMesh &m, size_t nv, const double *centroids, size_t nc, const uint8_t *fracture,
                    double radius, std::vector<uint32_t> *idx) {
  idx->clear();
  if (nv == 0 || nc == 0) return SHERD_OK;
  hipError_t e;
  sherd_cloud cc;  // the centroids' point tree (a temporary: its arrays are freed below)
  cc.ctx = c;
  cc.n = (int)nc;
  uint8_t *dfrac = nullptr;
  uint8_t *sel = (uint8_t *)c->scratch(SL_SEL, nv, &e);
  uint32_t *pos = sel ? (uint32_t *)c->scratch(SL_POS, sizeof(uint32_t) * (nv + select_scan_words((int)nv)), &e)
                      : nullptr;
  uint32_t last[2] = {0, 0};
  uint8_t last_sel = 0;
  if (sel && pos) {
    do {
      if ((e = hipMalloc((void **)&cc.pts, sizeof(double) * 3 * nc)) != hipSuccess) break;
      if ((e = hipMalloc((void **)&dfrac, nc)) != hipSuccess) break;
      if ((e = hipMemcpyAsync(cc.pts, centroids, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, c->stream)) !=
          hipSuccess)
        break;
      if ((e = hipMemcpyAsync(dfrac, fracture, nc, hipMemcpyHostToDevice, c->stream)) != hipSuccess) break;
      if ((e = point_tree(c, &cc)) != hipSuccess) break;
      if ((e = launch_select(m.V, (int)nv, view_of(cc.tree), cc.pts, dfrac, radius, sel, pos, pos + nv, c->stream)) !=
          hipSuccess)
        break;
      if ((e = hipMemcpyAsync(&last[0], pos + nv - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream)) !=
          hipSuccess)
        break;
      if ((e = hipMemcpyAsync(&last_sel, sel + nv - 1, 1, hipMemcpyDeviceToHost, c->stream)) != hipSuccess) break;
      if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) break;
      size_t kept = (size_t)last[0] + last_sel;
      idx->resize(kept);
      if (kept) {
        uint32_t *didx = (uint32_t *)c->scratch(SL_IDX, sizeof(uint32_t) * kept, &e);
        if (!didx) break;
        if ((e = launch_compact(sel, pos, (int)nv, didx, c->stream)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(idx->data(), didx, sizeof(uint32_t) * kept, hipMemcpyDeviceToHost, c->stream)) !=
            hipSuccess)
          break;
        e = hipStreamSynchronize(c->stream);
      }
    } while (0);
  }
  scene_free(&cc.tree);
  if (cc.pts) (void)hipFree(cc.pts);
  if (dfrac) (void)hipFree(dfrac);
  if (e != hipSuccess) return hip_fail(e, "vertex selection");
  return SHERD_OK;
}

// One join of the walk: its endpoints and its depth below the group's first fragment.
struct Step {
  uint32_t fixed, moving;
  int level;
  bool clouds;
};

// refine.rs:270-303's walk, which depends on `used` order and group membership only: per group of
// two or more, from its first fragment, repeatedly the first remaining join with exactly one
// endpoint placed. A fragment is reached once, from a fragment already final, so the joins of one
// depth are independent of each other and can be registered together.
std::vector<Step> walk(const uint32_t *members, const size_t *sizes, size_t n_groups, const uint32_t *used,
                       size_t n_used, const sherd_refine_piece *pieces, size_t n_pieces) {
  std::vector<Step> out;
  std::vector<int> level(n_pieces, -1);
  size_t at = 0;
  for (size_t g = 0; g < n_groups; ++g) {
    const uint32_t *grp = members + at;
    size_t gn = sizes[g];
    at += gn;
    if (gn < 2) continue;
    auto in_group = [&](uint32_t x) { return std::find(grp, grp + gn, x) != grp + gn; };
    std::vector<uint32_t> done{grp[0]};
    level[grp[0]] = 0;
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    for (size_t k = 0; k < n_used; ++k)
      if (in_group(used[2 * k]) && in_group(used[2 * k + 1])) edges.emplace_back(used[2 * k], used[2 * k + 1]);
    auto placed = [&](uint32_t x) { return std::find(done.begin(), done.end(), x) != done.end(); };
    for (;;) {
      auto it = std::find_if(edges.begin(), edges.end(),
                             [&](const std::pair<uint32_t, uint32_t> &e) { return placed(e.first) != placed(e.second); });
      if (it == edges.end()) break;
      uint32_t a = it->first, b = it->second;
      edges.erase(it);
      uint32_t fixed = placed(a) ? a : b, moving = placed(a) ? b : a;
      level[moving] = level[fixed] + 1;
      out.push_back(Step{fixed, moving, level[moving], pieces[fixed].cloud && pieces[moving].cloud});
      done.push_back(moving);
    }
  }
  return out;
}

void mul4(const double *a, const double *b, double *c) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = a[4 * i] * b[j];
      for (int k = 1; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
      c[4 * i + j] = s;
    }
}

}  // namespace

extern "C" {

int sherd_fracture_cloud(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf,
                         const double *centroids, const uint8_t *fracture, size_t n_working_faces, double thick,
                         double res, size_t max_points, uint64_t seed, sherd_cloud **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  LOCK(c);
  int rc = check_mesh(V, nv, F, nf);
  if (rc) return rc;
  if (n_working_faces && (!centroids || !fracture)) return fail(SHERD_E_ARG, "null working-mesh buffer");
  if (max_points > (size_t)MAX_RUNG_POINTS) return fail(SHERD_E_LIMIT, "cap beyond 262144 points");
  DevMesh m;
  hipError_t e = m.upload(c, V, nv, F, nf);
  if (e != hipSuccess) return hip_fail(e, "upload");
  // refine.rs:167: radius = max(0.15 t, 1.5 res).
  double radius = std::fmax(0.15 * thick, 1.5 * res);
  std::vector<uint32_t> idx;
  if ((rc = select_vertices(c, m, nv, centroids, n_working_faces, fracture, radius, &idx))) return rc;
  cap_selection(idx, max_points, seed);
  return vertex_cloud(c, m, nv, nf, std::move(idx), out);
}

int sherd_cloud_from_indices(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf,
                             const uint32_t *idx, size_t n, sherd_cloud **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  LOCK(c);
  int rc = check_mesh(V, nv, F, nf);
  if (rc) return rc;
  if (n && !idx) return fail(SHERD_E_ARG, "null index buffer");
  for (size_t k = 0; k < n; ++k)
    if (idx[k] >= nv) return fail(SHERD_E_ARG, "vertex index out of range");
  DevMesh m;
  hipError_t e = m.upload(c, V, nv, F, nf);
  if (e != hipSuccess) return hip_fail(e, "upload");
  return vertex_cloud(c, m, nv, nf, std::vector<uint32_t>(idx, idx + n), out);
}

int sherd_vertex_votes(sherd_ctx *ctx, const sherd_scene *scene, const double *V, size_t nv, const uint32_t *F,
                       size_t nf, double thick, int retest, uint8_t *good, uint8_t *frac) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  int rc = check_mesh(V, nv, F, nf);
  if (rc) return rc;
  if (!good && nv) return fail(SHERD_E_ARG, "null output");
  if (nv > (size_t)INT_MAX / 3 || nf > (size_t)INT_MAX / 3) return fail(SHERD_E_LIMIT, "mesh too large");
  if (nv == 0) return SHERD_OK;
  if (!(thick > 0.0) || !std::isfinite(thick)) return fail(SHERD_E_ARG, "thickness must be positive");
  DevMesh m;
  hipError_t e = m.upload(c, V, nv, F, nf);
  if (e != hipSuccess) return hip_fail(e, "upload");
  uint32_t *didx = (uint32_t *)c->scratch(SL_IDX, sizeof(uint32_t) * nv, &e);
  if (!didx) return hip_fail(e, "scratch");
  double *fn = (double *)c->scratch(SL_FN, sizeof(double) * 3 * std::max<size_t>(nf, 1), &e);
  if (!fn) return hip_fail(e, "scratch");
  uint32_t *ns = (uint32_t *)c->scratch(SL_NSCR, sizeof(uint32_t) * normals_scratch_words((int)nv, (int)nf), &e);
  if (!ns) return hip_fail(e, "scratch");
  double *pts = (double *)c->scratch(SL_VPTS, sizeof(double) * 3 * nv, &e);
  if (!pts) return hip_fail(e, "scratch");
  double *nrm = (double *)c->scratch(SL_VNRM, sizeof(double) * 3 * nv, &e);
  if (!nrm) return hip_fail(e, "scratch");
  uint8_t *dgood = (uint8_t *)c->scratch(SL_VGOOD, nv, &e);
  if (!dgood) return hip_fail(e, "scratch");
  uint8_t *dfrac = nullptr;
  if (frac && nf) {
    dfrac = (uint8_t *)c->scratch(SL_VFRAC, nf, &e);
    if (!dfrac) return hip_fail(e, "scratch");
  }
  Timer tm(c, SHERD_M_SEGMENT, nv);
  HIPCK(launch_iota(didx, (int)nv, c->stream), "iota");
  HIPCK(launch_vertex_cloud(m.V, (int)nv, m.F, (int)nf, didx, (int)nv, fn, ns, pts, nrm, c->stream), "vertex normals");
  if (c->prof.on) c->prof.begin("k_vertex_votes", c->stream);
  HIPCK(launch_vertex_votes(view_of(scene->d), m.V, (int)nv, m.F, (int)nf, nrm, thick, retest, dgood, dfrac,
                            c->stream),
        "vertex rays");
  if (c->prof.on) c->prof.end(c->stream);
  tm.stop();
  HIPCK(hipMemcpyAsync(good, dgood, nv, hipMemcpyDeviceToHost, c->stream), "download");
  if (dfrac) HIPCK(hipMemcpyAsync(frac, dfrac, nf, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  c->prof.collect();
  return SHERD_OK;
}

size_t sherd_cloud_size(const sherd_cloud *cloud) { return cloud ? (size_t)cloud->n : 0; }

int sherd_cloud_read(const sherd_cloud *cloud, uint32_t *idx, double *points, double *normals) {
  if (!cloud) return fail(SHERD_E_ARG, "null cloud");
  Ctx *c = cloud->ctx;
  LOCK(c);
  size_t n = (size_t)cloud->n;
  if (idx) {
    if (cloud->idx.size() != n) return fail(SHERD_E_ARG, "the cloud has no vertex indices");
    std::memcpy(idx, cloud->idx.data(), sizeof(uint32_t) * n);
  }
  if (normals && n && !cloud->normals) return fail(SHERD_E_ARG, "the cloud has no normals");
  hipError_t e = hipSuccess;
  if (points && n) e = hipMemcpyAsync(points, cloud->pts, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && normals && n)
    e = hipMemcpyAsync(normals, cloud->normals, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "download");
  return SHERD_OK;
}

int sherd_cloud_placed(const sherd_cloud *cloud, const double T[16], int with_tree, sherd_cloud **out) {
  if (!cloud || !T || !out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  Ctx *c = cloud->ctx;
  LOCK(c);
  size_t n = (size_t)cloud->n;
  double *pts = nullptr, *nrm = nullptr;
  hipError_t e = hipMalloc((void **)&pts, sizeof(double) * 3 * std::max<size_t>(n, 1));
  if (e == hipSuccess && cloud->normals) e = hipMalloc((void **)&nrm, sizeof(double) * 3 * std::max<size_t>(n, 1));
  Pose P;
  std::memcpy(P.m, T, sizeof P.m);
  if (e == hipSuccess) e = launch_place(cloud->pts, cloud->normals, (int)n, P, pts, nrm, c->stream);
  if (e != hipSuccess) {
    if (pts) (void)hipFree(pts);
    if (nrm) (void)hipFree(nrm);
    return hip_fail(e, "placement");
  }
  int rc = cloud_adopt(c, pts, nrm, n, with_tree != 0, out);
  if (rc == SHERD_OK) (*out)->idx = cloud->idx;
  return rc;
}

int sherd_refine_joins(sherd_ctx *ctx, const sherd_refine_piece *pieces, size_t n_pieces, const double *poses,
                       const uint32_t *members, const size_t *sizes, size_t n_groups, const uint32_t *used,
                       size_t n_used, double *poses_out, sherd_refined_join *joins, size_t *n_joins) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return fail(SHERD_E_ARG, "null context");
  if (!n_joins || (n_pieces && (!pieces || !poses || !poses_out)) || (n_groups && (!members || !sizes)) ||
      (n_used && !used))
    return fail(SHERD_E_ARG, "null buffer");
  size_t total = 0;
  for (size_t g = 0; g < n_groups; ++g) total += sizes[g];
  for (size_t k = 0; k < total; ++k)
    if (members[k] >= n_pieces) return fail(SHERD_E_ARG, "a group names a fragment out of range");
  for (size_t k = 0; k < 2 * n_used; ++k)
    if (used[k] >= n_pieces) return fail(SHERD_E_ARG, "a join names a fragment out of range");
  for (size_t p = 0; p < n_pieces; ++p)
    if (pieces[p].cloud && pieces[p].cloud->ctx != c) return fail(SHERD_E_ARG, "cloud of another context");
  std::vector<Step> steps = walk(members, sizes, n_groups, used, n_used, pieces, n_pieces);
  std::memcpy(poses_out, poses, sizeof(double) * 16 * n_pieces);
  int depth = 0;
  for (const Step &s : steps) depth = std::max(depth, s.level);
  // Results in walk order; a join without a cloud on either side leaves its pose alone.
  std::vector<int> slot(steps.size(), -1);
  size_t nj = 0;
  for (size_t k = 0; k < steps.size(); ++k)
    if (steps[k].clouds) slot[k] = (int)nj++;
  std::vector<sherd_refined_join> res(nj);
  const double K[2] = {0.05, 0.02};  // refine.rs:66 RUNGS
  const int ITERATIONS = 40;         // refine.rs:68
  int rc = SHERD_OK;
  for (int L = 1; L <= depth && rc == SHERD_OK; ++L) {
    std::vector<size_t> batch;
    for (size_t k = 0; k < steps.size(); ++k)
      if (steps[k].level == L && steps[k].clouds) batch.push_back(k);
    if (batch.empty()) continue;
    size_t nb = batch.size();
    std::vector<sherd_cloud *> made;
    std::vector<sherd_icp_pair> pairs(2 * nb);
    for (size_t b = 0; b < nb && rc == SHERD_OK; ++b) {
      const Step &s = steps[batch[b]];
      const sherd_refine_piece &pf = pieces[s.fixed], &pm = pieces[s.moving];
      // refine.rs:318-326: the pair's own scales, and the two clouds where the walk has them.
      double t = std::fmin(pf.thick, pm.thick), r = std::fmax(pf.res, pm.res);
      double icp = std::fmax(0.04 * t, 0.6 * r) / (0.04 * t);  // scales.rs:86
      sherd_cloud *src = nullptr, *tgt = nullptr;
      rc = sherd_cloud_placed(pm.cloud, poses + 16 * (size_t)s.moving, 0, &src);
      if (rc == SHERD_OK) made.push_back(src);
      if (rc == SHERD_OK) rc = sherd_cloud_placed(pf.cloud, poses_out + 16 * (size_t)s.fixed, 1, &tgt);
      if (rc == SHERD_OK) made.push_back(tgt);
      for (int q = 0; q < 2; ++q) {
        pairs[q * nb + b] = sherd_icp_pair{src, tgt, K[q] * t * icp};  // scales.rs:100 icp_dist
        res[slot[batch[b]]].dist[q] = K[q] * t * icp;
      }
      res[slot[batch[b]]].fixed = s.fixed;
      res[slot[batch[b]]].moving = s.moving;
    }
    if (rc == SHERD_OK) {
      std::vector<uint32_t> cp(nb);
      std::vector<double> init(16 * nb, 0.0);
      for (size_t b = 0; b < nb; ++b) {
        cp[b] = (uint32_t)b;
        for (int d = 0; d < 4; ++d) init[16 * b + 5 * d] = 1.0;
      }
      sherd_icp_rung_spec rungs[2] = {{SHERD_POINT_TO_PLANE, ITERATIONS}, {SHERD_POINT_TO_PLANE, ITERATIONS}};
      std::vector<sherd_registration> reg(2 * nb);
      rc = sherd_icp_batch_ladder(ctx, pairs.data(), nb, rungs, 2, cp.data(), init.data(), nb, reg.data());
      for (size_t b = 0; b < nb && rc == SHERD_OK; ++b) {
        const Step &s = steps[batch[b]];
        sherd_refined_join &j = res[slot[batch[b]]];
        for (int q = 0; q < 2; ++q) std::memcpy(j.rungs[q], reg[2 * b + q].transform, sizeof(double) * 16);
        double t = std::fmin(pieces[s.fixed].thick, pieces[s.moving].thick);
        j.fitness = reg[2 * b + 1].fitness;
        j.rmse_t = reg[2 * b + 1].inlier_rmse / t;
        double nt[16];
        mul4(j.rungs[1], poses_out + 16 * (size_t)s.moving, nt);
        std::memcpy(poses_out + 16 * (size_t)s.moving, nt, sizeof nt);
      }
    }
    for (sherd_cloud *m : made) sherd_cloud_free(m);
  }
  if (rc != SHERD_OK) return rc;
  if (joins) std::memcpy(joins, res.data(), sizeof(sherd_refined_join) * nj);
  *n_joins = nj;
  return SHERD_OK;
}

}  // extern "C"
