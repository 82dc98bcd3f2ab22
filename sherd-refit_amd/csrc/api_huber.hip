// C ABI for the robust seam energy: a batch of poses with frozen correspondences, its energy and
// se(3) gradient, and the batched LM refinement (huber.hip). Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"

using namespace sherd;

struct sherd_huber {
  Ctx *ctx = nullptr;
  int k = 0;
  size_t points = 0;
  std::vector<size_t> off;
  // Every device array below lives in one block (`block`, `block_n` bytes), carved in this order.
  void *block = nullptr;
  size_t block_n = 0;
  IcpPairDesc *pairs = nullptr;
  uint32_t *cand_pair = nullptr;
  size_t *doff = nullptr;
  double *T0 = nullptr;
  int32_t *corr = nullptr;
  int32_t *lm_corr = nullptr;
  double *xi = nullptr;
  double *E = nullptr;
  double *G = nullptr;
  double *T_out = nullptr;
  uint32_t *steps = nullptr;
  LmState *lm = nullptr;
  int max_ns = 0;
  HuberView view() const {
    HuberView v;
    v.pairs = pairs;
    v.cand_pair = cand_pair;
    v.off = doff;
    v.T0 = T0;
    v.corr = corr;
    v.lm_corr = lm_corr;
    v.lm = lm;
    v.k = k;
    v.max_ns = max_ns;
    return v;
  }
};

namespace {

// Returns a batch's block to its context as the spare (keeping the larger of the two blocks), or
// frees it, and deletes the batch.
void free_huber(sherd_huber *h) {
  Ctx *c = h->ctx;
  if (h->block) {
    if (!c->huber_spare.p || c->huber_spare.n < h->block_n) {
      if (c->huber_spare.p) (void)hipFree(c->huber_spare.p);
      c->huber_spare.p = h->block;
      c->huber_spare.n = h->block_n;
    } else {
      (void)hipFree(h->block);
    }
  }
  delete h;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// The byte offsets of a batch's arrays in its block, in the struct's order; the last entry is the
// block's size. The first four (pairs, candidate pairs, offsets, base poses) are the upload.
void huber_layout(size_t n_pairs, size_t k, size_t points, size_t at[13]) {
  const size_t sz[12] = {sizeof(IcpPairDesc) * n_pairs, sizeof(uint32_t) * k, sizeof(size_t) * k,
                         sizeof(double) * 16 * k,       sizeof(int32_t) * points, sizeof(int32_t) * points,
                         sizeof(double) * 6 * k,        sizeof(double) * k,   sizeof(double) * 6 * k,
                         sizeof(double) * 16 * k,       sizeof(uint32_t) * k, sizeof(LmState) * k};
  size_t b = 0;
  for (int i = 0; i < 12; ++i) {
    at[i] = b;
    b += up256(sz[i] ? sz[i] : 1);
  }
  at[12] = b;
}

}  // namespace

extern "C" {

int sherd_huber_create(sherd_ctx *ctx, const sherd_icp_pair *pairs, size_t n_pairs, const uint32_t *cand_pair,
                       const double *T0, size_t k, sherd_huber **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !out) return fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  LOCK(c);
  if (k == 0 || !pairs || !cand_pair || !T0) return fail(SHERD_E_ARG, "empty batch or null buffer");
  if (k > (size_t)INT_MAX || n_pairs > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "batch too large");
  std::vector<IcpPairDesc> pd(n_pairs);
  for (size_t p = 0; p < n_pairs; ++p) {
    const sherd_cloud *s = pairs[p].source, *t = pairs[p].target;
    if (!s || !t || s->ctx != c || t->ctx != c) return fail(SHERD_E_ARG, "pair without a cloud of this context");
    if (t->n > 0 && !t->normals) return fail(SHERD_E_ARG, "the point-to-plane energy needs target normals");
    IcpPairDesc &d = pd[p];
    std::memset(&d, 0, sizeof d);
    d.src = s->pts;
    d.ns = t->n > 0 ? s->n : 0;
    d.tp = t->pts;
    d.tn = t->normals;
    if (t->n > 0) d.tree = view_of(t->tree);
    d.r = pairs[p].distance;
    for (int a = 0; a < 3; ++a) {
      d.cs[a] = s->centroid[a];
      d.ct[a] = t->centroid[a];
    }
  }
  sherd_huber *h = new sherd_huber();
  h->ctx = c;
  h->k = (int)k;
  h->off.resize(k);
  for (size_t i = 0; i < k; ++i) {
    if (cand_pair[i] >= n_pairs) {
      delete h;
      return fail(SHERD_E_ARG, "pose names a pair out of range");
    }
    h->off[i] = h->points;
    h->points += (size_t)pd[cand_pair[i]].ns;
    h->max_ns = std::max(h->max_ns, pd[cand_pair[i]].ns);
  }
  size_t at[13];
  huber_layout(n_pairs, k, h->points, at);
  hipError_t e = hipSuccess;
  do {
    // The context's spare block when it is large enough, else a new one.
    if (c->huber_spare.p && c->huber_spare.n >= at[12]) {
      h->block = c->huber_spare.p;
      h->block_n = c->huber_spare.n;
      c->huber_spare.p = nullptr;
      c->huber_spare.n = 0;
    } else {
      if (take_fault(c, SHERD_FAULT_NOMEM)) {
        e = hipErrorOutOfMemory;
        break;
      }
      if ((e = hipMalloc(&h->block, at[12])) != hipSuccess) {
        h->block = nullptr;
        break;
      }
      h->block_n = at[12];
    }
    char *b = (char *)h->block;
    h->pairs = (IcpPairDesc *)(b + at[0]);
    h->cand_pair = (uint32_t *)(b + at[1]);
    h->doff = (size_t *)(b + at[2]);
    h->T0 = (double *)(b + at[3]);
    h->corr = (int32_t *)(b + at[4]);
    h->lm_corr = (int32_t *)(b + at[5]);
    h->xi = (double *)(b + at[6]);
    h->E = (double *)(b + at[7]);
    h->G = (double *)(b + at[8]);
    h->T_out = (double *)(b + at[9]);
    h->steps = (uint32_t *)(b + at[10]);
    h->lm = (LmState *)(b + at[11]);
    // The four inputs go up in one copy from the pinned staging buffer, laid out as in the block.
    char *hp = (char *)c->pinned(at[4], &e);
    if (!hp) break;
    std::memcpy(hp + at[0], pd.data(), sizeof(IcpPairDesc) * n_pairs);
    std::memcpy(hp + at[1], cand_pair, sizeof(uint32_t) * k);
    std::memcpy(hp + at[2], h->off.data(), sizeof(size_t) * k);
    std::memcpy(hp + at[3], T0, sizeof(double) * 16 * k);
    if ((e = hipMemcpyAsync(b, hp, at[4], hipMemcpyHostToDevice, c->stream)) != hipSuccess) break;
    if ((e = launch_correspond(h->view(), h->T0, h->corr, c->stream)) != hipSuccess) break;
    e = hipStreamSynchronize(c->stream);
  } while (0);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(c->stream);
    free_huber(h);
    return hip_fail(e, "huber batch");
  }
  *out = h;
  return SHERD_OK;
}

void sherd_huber_free(sherd_huber *h) {
  if (!h) return;
  std::lock_guard<std::mutex> lock(h->ctx->mu);
  (void)hipSetDevice(h->ctx->device);
  (void)hipStreamSynchronize(h->ctx->stream);
  free_huber(h);
}

int sherd_huber_correspondences(sherd_huber *h, int32_t *corr, size_t n) {
  if (!h || (!corr && n)) return fail(SHERD_E_ARG, "null argument");
  Ctx *c = h->ctx;
  LOCK(c);
  if (n != h->points) return fail(SHERD_E_ARG, "buffer size differs from the batch's source points");
  HIPCK(hipMemcpyAsync(corr, h->corr, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  return SHERD_OK;
}

int sherd_huber_eval(sherd_huber *h, const double *xi, double delta, int precision, double *energy, double *grad) {
  if (!h || !xi || !energy) return fail(SHERD_E_ARG, "null argument");
  if (precision != 32 && precision != 64) return fail(SHERD_E_ARG, "precision must be 32 or 64");
  if (!(delta > 0.0)) return fail(SHERD_E_ARG, "delta must be positive");
  Ctx *c = h->ctx;
  LOCK(c);
  size_t k = (size_t)h->k;
  HIPCK(hipMemcpyAsync(h->xi, xi, sizeof(double) * 6 * k, hipMemcpyHostToDevice, c->stream), "upload");
  Timer tm(c, SHERD_M_HUBER, k);
  if (c->prof.on) c->prof.begin(grad ? "k_huber_grad" : "k_huber_energy", c->stream);
  HIPCK(launch_huber_eval(h->view(), h->xi, delta, precision == 64, h->E, grad ? h->G : nullptr, c->stream),
        "huber kernel");
  if (c->prof.on) c->prof.end(c->stream);
  tm.stop();
  HIPCK(hipMemcpyAsync(energy, h->E, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream), "download");
  if (grad) HIPCK(hipMemcpyAsync(grad, h->G, sizeof(double) * 6 * k, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  c->prof.collect();
  return SHERD_OK;
}

int sherd_huber_lm(sherd_huber *h, double delta, int iterations, int precision, double *T_out, double *energy,
                   uint32_t *steps) {
  if (!h || !T_out || !energy) return fail(SHERD_E_ARG, "null argument");
  if (precision != 32 && precision != 64) return fail(SHERD_E_ARG, "precision must be 32 or 64");
  if (!(delta > 0.0) || iterations < 0) return fail(SHERD_E_ARG, "bad delta or iteration count");
  Ctx *c = h->ctx;
  LOCK(c);
  size_t k = (size_t)h->k;
  Timer tm(c, SHERD_M_HUBER, k);
  if (c->prof.on) c->prof.begin("k_huber_lm", c->stream);
  HIPCK(launch_huber_lm(h->view(), delta, iterations, precision == 64, h->T_out, h->E, h->steps, c->stream),
        "lm kernel");
  if (c->prof.on) c->prof.end(c->stream);
  tm.stop();
  HIPCK(hipMemcpyAsync(T_out, h->T_out, sizeof(double) * 16 * k, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipMemcpyAsync(energy, h->E, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream), "download");
  if (steps) HIPCK(hipMemcpyAsync(steps, h->steps, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, c->stream), "download");
  SYNC_RECEIPT(c);
  c->prof.collect();
  return SHERD_OK;
}

}  // extern "C"
