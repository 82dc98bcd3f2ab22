// The C ABI (include/sherd_hip.h): contexts, handles, argument checks, transfers, counters.
//
// Host code only: every computation is a device kernel launched from the subsystem translation
// units; this file moves caller buffers to and from the device and reports failures as statuses
// (a failed call computes nothing on the CPU; the caller delegates, as GpuExecutor does in
// crates/sherd-gpu/src/executor.rs:425-487).
#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sherd_hip.h"
#include "api_internal.hpp"
#include "kernels.hpp"

using namespace sherd;

namespace {
thread_local std::string g_error;
thread_local uint64_t g_fails = 0;  // failures set on this thread, and the last one's status
thread_local int g_last = SHERD_OK;

__global__ void k_receipt(uint32_t *word, uint32_t value) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *word = value;
}
}  // namespace

namespace sherd {

int fail(int code, const std::string &msg) {
  g_error = msg;
  g_fails += 1;
  g_last = code;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  return fail((e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SHERD_E_NOMEM : SHERD_E_HIP, m);
}

CallScope::CallScope(Ctx *c) : c_(c), n0_(g_fails) { c_->health[Ctx::H_CALLS] += 1; }

CallScope::~CallScope() {
  if (g_fails == n0_) return;
  c_->health[Ctx::H_FAILED] += 1;
  switch (g_last) {
    case SHERD_E_ARG: c_->health[Ctx::H_ARG] += 1; break;
    case SHERD_E_NOMEM: c_->health[Ctx::H_NOMEM] += 1; break;
    case SHERD_E_LIMIT: c_->health[Ctx::H_LIMIT] += 1; break;
    case SHERD_E_CORRUPT: c_->health[Ctx::H_CORRUPT] += 1; break;
    default: c_->health[Ctx::H_HOST] += 1; break;
  }
}

bool take_fault(Ctx *c, int kind) {
  if (c->fault_kind != kind) return false;
  if (c->fault_after > 0) {
    c->fault_after -= 1;
    return false;
  }
  c->fault_kind = SHERD_FAULT_NONE;
  c->health[Ctx::H_INJECTED] += 1;
  return true;
}

int sync_receipt(Ctx *c) {
  c->nonce += 1;
  uint32_t want = c->nonce + 1u;
  // A fired corrupt fault leaves the word as a stale readback would hold it.
  uint32_t value = take_fault(c, SHERD_FAULT_CORRUPT) ? c->nonce : want;
  *c->receipt_host = 0u;
  hipLaunchKernelGGL(k_receipt, dim3(1), dim3(64), 0, c->stream, c->receipt_dev, value);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(c->receipt_host, c->receipt_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "receipt");
  if (*c->receipt_host != want)
    return fail(SHERD_E_CORRUPT, "the receipt came back as " + std::to_string(*c->receipt_host) + ", not " +
                                     std::to_string(want));
  return SHERD_OK;
}

void *Ctx::pinned(size_t bytes, hipError_t *err) {
  *err = hipSuccess;
  if (host.n >= bytes && host.p) return host.p;
  if (host.p) {
    (void)hipStreamSynchronize(stream);
    (void)hipHostFree(host.p);
    host.p = nullptr;
    host.n = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes + bytes / 4;
  *err = hipHostMalloc(&host.p, want, hipHostMallocDefault);
  if (*err != hipSuccess) {
    host.p = nullptr;
    return nullptr;
  }
  host.n = want;
  return host.p;
}

void *Ctx::scratch(int slot, size_t bytes, hipError_t *err) {
  *err = hipSuccess;
  if (take_fault(this, SHERD_FAULT_NOMEM)) {
    *err = hipErrorOutOfMemory;
    return nullptr;
  }
  Buf &b = bufs[slot];
  if (b.n >= bytes && b.p) return b.p;
  if (b.p) {
    (void)hipStreamSynchronize(stream);
    (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  want += want / 4;
  *err = hipMalloc(&b.p, want);
  if (*err != hipSuccess) {
    b.p = nullptr;
    return nullptr;
  }
  b.n = want;
  return b.p;
}

Timer::Timer(Ctx *c, int method, uint64_t items) : c_(c), m_(method) {
  c_->stats[m_].calls += 1;
  c_->stats[m_].items += items;
  (void)hipEventRecord(c_->ev0, c_->stream);
}

void Timer::stop() {
  if (done_) return;
  done_ = true;
  (void)hipEventRecord(c_->ev1, c_->stream);
  if (hipEventSynchronize(c_->ev1) == hipSuccess) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c_->ev0, c_->ev1) == hipSuccess) c_->stats[m_].device_ns += (uint64_t)(ms * 1e6);
  }
}

int KernelProfiler::slot(const char *k) {
  for (int i = 0; i < n; ++i)
    if (std::strcmp(name[i], k) == 0) return i;
  if (n >= MAXK) return MAXK - 1;
  name[n] = k;
  return n++;
}

void KernelProfiler::begin(const char *k, hipStream_t s) {
  if (npending >= MAXK) return;
  int i = slot(k);
  if (!ev[2 * npending]) {
    (void)hipEventCreate(&ev[2 * npending]);
    (void)hipEventCreate(&ev[2 * npending + 1]);
  }
  pending[npending] = i;
  (void)hipEventRecord(ev[2 * npending], s);
}

void KernelProfiler::end(hipStream_t s) {
  if (npending >= MAXK) return;
  (void)hipEventRecord(ev[2 * npending + 1], s);
  ++npending;
}

void KernelProfiler::collect() {
  for (int p = 0; p < npending; ++p) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, ev[2 * p], ev[2 * p + 1]) == hipSuccess) {
      ms[pending[p]] += t;
      launches[pending[p]] += 1;
    }
  }
  npending = 0;
}

}  // namespace sherd

namespace {

// ---- the no-allocation form of bounded_distance -------------------------------------------------
//
// `scratch` is the only device memory sherd_bounded_distance_noalloc uses beyond the scene: the
// moved-from points (n x 3 f64), the per-point f32 distances, their f64 widening, and the running
// minimum's bits. Its size is fixed here; each device that opens a context holds one copy until the
// process ends, and every call on that device shares it.
constexpr size_t NOALLOC_MAX_POINTS = (size_t)1 << 18;
constexpr size_t NA_PTS = 0;                                                // f64 x 3n
constexpr size_t NA_DIST = NA_PTS + sizeof(double) * 3 * NOALLOC_MAX_POINTS;  // f32 x n
constexpr size_t NA_WIDE = NA_DIST + sizeof(float) * NOALLOC_MAX_POINTS;      // f64 x n
constexpr size_t NA_BEST = NA_WIDE + sizeof(double) * NOALLOC_MAX_POINTS;     // u32 (own 256 B)
constexpr size_t NA_BYTES = NA_BEST + 256;

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// The device address of this device's `scratch` (call with the device current).
hipError_t noalloc_address(void **ptr) {
  *ptr = nullptr;
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

__global__ void k_set_u32(unsigned int *p, unsigned int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ void k_widen(const float *__restrict__ in, double *__restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)in[i];
}

}  // namespace

extern "C" {

const char *sherd_name(void) { return "hip"; }
const char *sherd_last_error(void) { return g_error.c_str(); }

int sherd_health_read(sherd_ctx *ctx, sherd_health *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !out) return fail(SHERD_E_ARG, "null argument");
  static_assert(sizeof(sherd_health) == sizeof(uint64_t) * Ctx::H_COUNT, "sherd_health's fields");
  uint64_t *o = &out->calls;  // sherd_health is H_COUNT consecutive uint64_t in Ctx's order
  for (int k = 0; k < Ctx::H_COUNT; ++k) o[k] = c->health[k].load();
  return SHERD_OK;
}

int sherd_fault_inject(sherd_ctx *ctx, int kind, int after) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return fail(SHERD_E_ARG, "null context");
  if (kind < SHERD_FAULT_NONE || kind > SHERD_FAULT_CORRUPT || after < 0) return fail(SHERD_E_ARG, "bad fault");
  std::lock_guard<std::mutex> lock(c->mu);
  c->fault_kind = kind;
  c->fault_after = after;
  return SHERD_OK;
}

int sherd_open(int device, sherd_ctx **out) {
  if (!out) return fail(SHERD_E_ARG, "null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(SHERD_E_NODEV, "no HIP device");
  if (device < 0 || device >= count) return fail(SHERD_E_NODEV, "device index out of range");
  hipDeviceProp_t prop;
  HIPCK(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(SHERD_E_NODEV, std::string("built for gfx950, found ") + prop.gcnArchName);
  HIPCK(hipSetDevice(device), "hipSetDevice");
  Ctx *c = new Ctx();
  c->device = device;
  c->name = std::string(prop.name) + " (" + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) + " CUs)";
  c->cus = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipStreamCreateWithFlags(&c->side.side[k], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.early, hipEventDisableTiming);
  for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&c->side.join[k], hipEventDisableTiming);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "stream/event creation");
  }
  // The no-allocation entry point's array: each device holds its own copy at its own address.
  if ((e = noalloc_address(&c->noalloc)) != hipSuccess) {
    sherd_close(reinterpret_cast<sherd_ctx *>(c));
    return hip_fail(e, "hipGetSymbolAddress(scratch)");
  }
  e = hipMalloc(&c->receipt_dev, 256);
  if (e == hipSuccess) e = hipHostMalloc(&c->receipt_host, 256, hipHostMallocDefault);
  if (e != hipSuccess) {
    sherd_close(reinterpret_cast<sherd_ctx *>(c));
    return hip_fail(e, "receipt buffers");
  }
  *out = reinterpret_cast<sherd_ctx *>(c);
  return SHERD_OK;
}

void sherd_close(sherd_ctx *ctx) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &b : c->bufs)
    if (b.p) (void)hipFree(b.p);
  if (c->host.p) (void)hipHostFree(c->host.p);
  if (c->huber_spare.p) (void)hipFree(c->huber_spare.p);
  if (c->receipt_dev) (void)hipFree(c->receipt_dev);
  if (c->receipt_host) (void)hipHostFree(c->receipt_host);
  if (c->ladder_exec) (void)hipGraphExecDestroy(c->ladder_exec);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  for (int k = 0; k < 3; ++k) {
    if (c->side.side[k]) {
      (void)hipStreamSynchronize(c->side.side[k]);
      (void)hipStreamDestroy(c->side.side[k]);
    }
    if (c->side.join[k]) (void)hipEventDestroy(c->side.join[k]);
  }
  if (c->side.fork) (void)hipEventDestroy(c->side.fork);
  if (c->side.early) (void)hipEventDestroy(c->side.early);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int sherd_device_name(sherd_ctx *ctx, char *buf, size_t len) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !buf || len == 0) return fail(SHERD_E_ARG, "null argument");
  std::snprintf(buf, len, "%s", c->name.c_str());
  return SHERD_OK;
}

int sherd_stats(sherd_ctx *ctx, int method, sherd_method_stats *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !out || method < 0 || method >= SHERD_M_COUNT) return fail(SHERD_E_ARG, "bad argument");
  std::lock_guard<std::mutex> lock(c->mu);
  *out = c->stats[method];
  return SHERD_OK;
}

// ---- scenes --------------------------------------------------------------------------------------

int sherd_scene_build(sherd_ctx *ctx, const float *V, size_t nv, const uint32_t *F, size_t nf,
                      const uint8_t *mask, sherd_scene **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(SHERD_E_ARG, "null output");
  *out = nullptr;
  LOCK(c);
  if (!V || !F || nf == 0 || nv == 0) return fail(SHERD_E_ARG, "a scene needs vertices and at least one face");
  if (nf > (size_t)(INT_MAX / 2) || nv > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "mesh larger than 2^30 faces");
  std::vector<uint32_t> kept;
  kept.reserve(3 * nf);
  for (size_t i = 0; i < nf; ++i) {
    if (mask && !mask[i]) continue;
    for (int k = 0; k < 3; ++k) {
      if (F[3 * i + k] >= nv) return fail(SHERD_E_ARG, "face index out of range");
      kept.push_back(F[3 * i + k]);
    }
  }
  if (kept.empty()) {  // bvh.rs:71-76: the first face
    for (int k = 0; k < 3; ++k) {
      if (F[k] >= nv) return fail(SHERD_E_ARG, "face index out of range");
      kept.push_back(F[k]);
    }
  }
  int m = (int)(kept.size() / 3);
  float *dV = nullptr;
  uint32_t *dF = nullptr;
  HIPCK(hipMalloc(&dV, sizeof(float) * 3 * nv), "hipMalloc vertices");
  hipError_t e = hipMalloc(&dF, sizeof(uint32_t) * kept.size());
  if (e != hipSuccess) {
    (void)hipFree(dV);
    return hip_fail(e, "hipMalloc faces");
  }
  sherd_scene *s = new sherd_scene();
  s->ctx = c;
  e = hipMemcpyAsync(dV, V, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dF, kept.data(), sizeof(uint32_t) * kept.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    Timer tm(c, SHERD_M_SCENE_BUILD, (uint64_t)m);
    e = scene_build(dV, dF, m, &s->d, c->stream);
    tm.stop();
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dV);
  (void)hipFree(dF);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "scene build");
  }
  *out = s;
  return SHERD_OK;
}

void sherd_scene_free(sherd_scene *scene) {
  if (!scene) return;
  std::lock_guard<std::mutex> lock(scene->ctx->mu);
  (void)hipSetDevice(scene->ctx->device);
  (void)hipStreamSynchronize(scene->ctx->stream);
  scene_free(&scene->d);
  delete scene;
}

int sherd_scene_info(const sherd_scene *scene, size_t *faces, float lo[3], float hi[3]) {
  if (!scene) return fail(SHERD_E_ARG, "null scene");
  if (faces) *faces = (size_t)scene->d.n;
  for (int k = 0; k < 3; ++k) {
    if (lo) lo[k] = scene->d.lo[k];
    if (hi) hi[k] = scene->d.hi[k];
  }
  return SHERD_OK;
}

// ---- executor methods over a scene ---------------------------------------------------------------

int sherd_first_hit(sherd_ctx *ctx, const sherd_scene *scene, const float *orig, const float *dir, size_t n,
                    int retest, float *t, uint32_t *prim) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  if (n == 0) return SHERD_OK;
  if (!orig || !dir || !t || !prim) return fail(SHERD_E_ARG, "null buffer");
  if (n > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "more than 2^31 rays");
  hipError_t e;
  float *d_o = (float *)c->scratch(0, sizeof(float) * 3 * n, &e);
  if (!d_o) return hip_fail(e, "scratch");
  float *d_d = (float *)c->scratch(1, sizeof(float) * 3 * n, &e);
  if (!d_d) return hip_fail(e, "scratch");
  float *d_t = (float *)c->scratch(2, sizeof(float) * n, &e);
  if (!d_t) return hip_fail(e, "scratch");
  uint32_t *d_p = (uint32_t *)c->scratch(3, sizeof(uint32_t) * n, &e);
  if (!d_p) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(d_o, orig, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream), "upload");
  HIPCK(hipMemcpyAsync(d_d, dir, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream), "upload");
  Timer tm(c, SHERD_M_FIRST_HIT, n);
  HIPCK(launch_first_hit(view_of(scene->d), d_o, d_d, (int)n, retest, d_t, d_p, c->stream), "first_hit kernel");
  tm.stop();
  HIPCK(hipMemcpyAsync(t, d_t, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipMemcpyAsync(prim, d_p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream), "download");
  SYNC_RECEIPT(c);
  return SHERD_OK;
}

static bool pose_ok(const double *T, Pose *p) {
  if (!T) return false;
  for (int i = 0; i < 16; ++i) p->m[i] = T[i];
  return true;
}

int sherd_bounded_distance(sherd_ctx *ctx, const sherd_scene *scene, const double *pts, size_t n,
                           const double T[16], double max_dist, int reduce, double *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  Pose P;
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  if (!pose_ok(T, &P) || !out || (n && !pts)) return fail(SHERD_E_ARG, "null buffer");
  if (reduce != 0 && reduce != 1) return fail(SHERD_E_ARG, "reduce must be 0 (All) or 1 (Min)");
  if (n > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "more than 2^31 points");
  float window = (float)max_dist;
  if (reduce == 1) {
    if (n == 0) {
      out[0] = (double)window;
      return SHERD_OK;
    }
  } else if (n == 0) {
    return SHERD_OK;
  }
  hipError_t e;
  double *d_p = (double *)c->scratch(0, sizeof(double) * 3 * n, &e);
  if (!d_p) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(d_p, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream), "upload");
  if (reduce == 0) {
    float *d_o = (float *)c->scratch(1, sizeof(float) * n, &e);
    if (!d_o) return hip_fail(e, "scratch");
    Timer tm(c, SHERD_M_BOUNDED_DISTANCE, n);
    HIPCK(launch_bounded(view_of(scene->d), d_p, (int)n, P, window, d_o, c->stream), "bounded kernel");
    tm.stop();
    std::vector<float> h(n);
    HIPCK(hipMemcpyAsync(h.data(), d_o, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "download");
    SYNC_RECEIPT(c);
    for (size_t i = 0; i < n; ++i) out[i] = (double)h[i];
    return SHERD_OK;
  }
  unsigned int *d_b = (unsigned int *)c->scratch(1, sizeof(unsigned int), &e);
  if (!d_b) return hip_fail(e, "scratch");
  // The window starts at max_dist (f32); a negative or NaN window answers itself (bvh.rs:124).
  unsigned int init = __builtin_bit_cast(unsigned int, window);
  if (!(window > 0.0f)) {
    out[0] = (double)window;
    return SHERD_OK;
  }
  HIPCK(hipMemcpyAsync(d_b, &init, sizeof init, hipMemcpyHostToDevice, c->stream), "upload");
  Timer tm(c, SHERD_M_BOUNDED_DISTANCE, n);
  HIPCK(launch_bounded_min(view_of(scene->d), d_p, (int)n, P, d_b, c->stream), "bounded min kernel");
  tm.stop();
  unsigned int bits = 0;
  HIPCK(hipMemcpyAsync(&bits, d_b, sizeof bits, hipMemcpyDeviceToHost, c->stream), "download");
  SYNC_RECEIPT(c);
  out[0] = (double)__builtin_bit_cast(float, bits);
  return SHERD_OK;
}

int sherd_bounded_distance_noalloc(sherd_ctx *ctx, const sherd_scene *scene, const double *pts, size_t n,
                                   const double T[16], double max_dist, int reduce, double *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  Pose P;
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  if (!pose_ok(T, &P) || !out || (n && !pts)) return fail(SHERD_E_ARG, "null buffer");
  if (reduce != 0 && reduce != 1) return fail(SHERD_E_ARG, "reduce must be 0 (All) or 1 (Min)");
  if (n > NOALLOC_MAX_POINTS)
    return fail(SHERD_E_LIMIT, "sherd_bounded_distance_noalloc takes at most 262144 points");
  if (!c->noalloc) return fail(SHERD_E_HIP, "no scratch address for this device");
  float window = (float)max_dist;
  if (n == 0) {
    if (reduce == 1) out[0] = (double)window;
    return SHERD_OK;
  }
  if (reduce == 1 && !(window > 0.0f)) {  // bvh.rs:124: a negative or NaN window answers itself
    out[0] = (double)window;
    return SHERD_OK;
  }
  char *base = (char *)c->noalloc;
  double *d_p = (double *)(base + NA_PTS);
  HIPCK(hipMemcpyAsync(d_p, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream), "upload");
  Timer tm(c, SHERD_M_BOUNDED_DISTANCE, n);
  if (reduce == 0) {
    float *d_o = (float *)(base + NA_DIST);
    double *d_w = (double *)(base + NA_WIDE);
    HIPCK(launch_bounded(view_of(scene->d), d_p, (int)n, P, window, d_o, c->stream), "bounded kernel");
    hipLaunchKernelGGL(k_widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d_o, d_w, (int)n);
    HIPCK(hipGetLastError(), "widen kernel");
    tm.stop();
    HIPCK(hipMemcpyAsync(out, d_w, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream), "download");
    HIPCK(hipStreamSynchronize(c->stream), "synchronize");
    return SHERD_OK;
  }
  // The running minimum starts at the window's bits (set by a kernel, not a memset, so the call
  // could sit in a captured graph), every call afresh: the array keeps the last call's value.
  unsigned int *d_b = (unsigned int *)(base + NA_BEST);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c->stream, d_b, __builtin_bit_cast(unsigned int, window));
  HIPCK(hipGetLastError(), "window kernel");
  HIPCK(launch_bounded_min(view_of(scene->d), d_p, (int)n, P, d_b, c->stream), "bounded min kernel");
  tm.stop();
  unsigned int bits = 0;
  HIPCK(hipMemcpyAsync(&bits, d_b, sizeof bits, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  out[0] = (double)__builtin_bit_cast(float, bits);
  return SHERD_OK;
}

int sherd_inside(sherd_ctx *ctx, const sherd_scene *scene, const double *pts, size_t n, const double T[16],
                 int retest, sherd_inside_outcome *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  Pose P;
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  if (n == 0) return SHERD_OK;
  if (!pose_ok(T, &P) || !out || !pts) return fail(SHERD_E_ARG, "null buffer");
  if (n > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "more than 2^31 points");
  hipError_t e = scene_parity_grids(&const_cast<sherd_scene *>(scene)->d, c->stream);
  if (e == hipErrorInvalidValue) return fail(SHERD_E_LIMIT, "parity grid beyond 2^31 entries");
  if (e != hipSuccess) return hip_fail(e, "parity grids");
  double *d_p = (double *)c->scratch(0, sizeof(double) * 3 * n, &e);
  if (!d_p) return hip_fail(e, "scratch");
  float *d_depth = (float *)c->scratch(1, sizeof(float) * n, &e);
  if (!d_depth) return hip_fail(e, "scratch");
  uint32_t *d_in = (uint32_t *)c->scratch(2, sizeof(uint32_t) * n, &e);
  if (!d_in) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(d_p, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream), "upload");
  Timer tm(c, SHERD_M_INSIDE, n);
  HIPCK(launch_inside(view_of(scene->d), d_p, (int)n, P, retest, d_depth, d_in, c->stream), "inside kernel");
  tm.stop();
  std::vector<float> hd(n);
  std::vector<uint32_t> hi(n);
  HIPCK(hipMemcpyAsync(hd.data(), d_depth, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipMemcpyAsync(hi.data(), d_in, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream), "download");
  SYNC_RECEIPT(c);
  for (size_t i = 0; i < n; ++i) {
    out[i].depth = hd[i];
    out[i].inside = hi[i];
  }
  return SHERD_OK;
}

}  // extern "C"

// ---- nearest neighbours, coarse score, ICP rung ---------------------------------------------------

namespace {

// Uploads n x 3 f64 into scratch slot `slot`.
double *upload3(Ctx *c, int slot, const double *h, size_t n, hipError_t *e) {
  double *d = (double *)c->scratch(slot, sizeof(double) * 3 * (n ? n : 1), e);
  if (!d) return nullptr;
  if (n) *e = hipMemcpyAsync(d, h, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream);
  return *e == hipSuccess ? d : nullptr;
}

struct GridGuard {
  DevGrid g;
  ~GridGuard() { grid_free(&g); }
};

}  // namespace

extern "C" {

int sherd_nearest_below(sherd_ctx *ctx, const double *cloud, size_t n, const double *q, size_t m, double bound,
                        int64_t *idx, double *dist) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (m == 0) return SHERD_OK;
  if (!q || !idx || !dist || (n && !cloud)) return fail(SHERD_E_ARG, "null buffer");
  if (n > (size_t)INT_MAX / 2 || m > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "cloud or query set too large");
  if (n == 0 || !(bound > 0.0)) {
    for (size_t i = 0; i < m; ++i) {
      idx[i] = -1;
      dist[i] = __builtin_inf();
    }
    return SHERD_OK;
  }
  hipError_t e = hipSuccess;
  double *dc = upload3(c, 0, cloud, n, &e);
  if (!dc) return hip_fail(e, "upload cloud");
  double *dq = upload3(c, 1, q, m, &e);
  if (!dq) return hip_fail(e, "upload queries");
  int64_t *di = (int64_t *)c->scratch(2, sizeof(int64_t) * m, &e);
  if (!di) return hip_fail(e, "scratch");
  double *dd = (double *)c->scratch(3, sizeof(double) * m, &e);
  if (!dd) return hip_fail(e, "scratch");
  GridGuard gg;
  HIPCK(grid_build(dc, (int)n, bound, &gg.g, c->stream), "grid build");
  Timer tm(c, SHERD_M_NEAREST, m);
  HIPCK(launch_nearest(view_of(gg.g), dq, (int)m, bound, di, dd, c->stream), "nearest kernel");
  tm.stop();
  HIPCK(hipMemcpyAsync(idx, di, sizeof(int64_t) * m, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipMemcpyAsync(dist, dd, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream), "download");
  SYNC_RECEIPT(c);
  return SHERD_OK;
}

int sherd_coarse_scores(sherd_ctx *ctx, const double *target_p, const double *target_n, size_t n_target,
                        const double *probe_p, const double *probe_n, size_t n_probe, const double *poses,
                        size_t k, double radius, double normal_agree, uint32_t *counts, double *scores) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (k == 0) return SHERD_OK;
  if (!poses || !scores) return fail(SHERD_E_ARG, "null buffer");
  if (k > (size_t)INT_MAX / 64 || n_probe > (size_t)INT_MAX || n_target > (size_t)INT_MAX / 2)
    return fail(SHERD_E_LIMIT, "batch too large");
  if (n_probe == 0 || n_target == 0 || !(radius > 0.0)) {
    // cpu.rs:47-55: an empty probe scores zero; no target or no radius finds no neighbour.
    for (size_t h = 0; h < k; ++h) {
      scores[h] = 0.0;
      if (counts) counts[h] = 0;
    }
    return SHERD_OK;
  }
  if (!target_p || !target_n || !probe_p || !probe_n) return fail(SHERD_E_ARG, "null buffer");
  hipError_t e = hipSuccess;
  double *dtp = upload3(c, 0, target_p, n_target, &e);
  if (!dtp) return hip_fail(e, "upload target");
  double *dtn = upload3(c, 1, target_n, n_target, &e);
  if (!dtn) return hip_fail(e, "upload target normals");
  double *dpp = upload3(c, 2, probe_p, n_probe, &e);
  if (!dpp) return hip_fail(e, "upload probe");
  double *dpn = upload3(c, 3, probe_n, n_probe, &e);
  if (!dpn) return hip_fail(e, "upload probe normals");
  double *dposes = (double *)c->scratch(4, sizeof(double) * 12 * k, &e);
  if (!dposes) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(dposes, poses, sizeof(double) * 12 * k, hipMemcpyHostToDevice, c->stream), "upload poses");
  uint32_t *dcnt = (uint32_t *)c->scratch(5, sizeof(uint32_t) * k, &e);
  if (!dcnt) return hip_fail(e, "scratch");
  double *dsc = (double *)c->scratch(6, sizeof(double) * k, &e);
  if (!dsc) return hip_fail(e, "scratch");
  GridGuard gg;
  HIPCK(grid_build(dtp, (int)n_target, radius, &gg.g, c->stream), "grid build");
  Timer tm(c, SHERD_M_COARSE, k);
  HIPCK(launch_coarse(view_of(gg.g), dtn, dpp, dpn, (int)n_probe, dposes, (int)k, radius, normal_agree, dcnt, dsc,
                      c->stream),
        "coarse kernel");
  tm.stop();
  if (counts) HIPCK(hipMemcpyAsync(counts, dcnt, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipMemcpyAsync(scores, dsc, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream), "download");
  SYNC_RECEIPT(c);
  return SHERD_OK;
}

}  // extern "C"

extern "C" {

int sherd_cone_votes(sherd_ctx *ctx, const sherd_scene *scene, const double *V, size_t nv, const uint32_t *F,
                     size_t nf, const int64_t *faces, const double *ns, size_t n_faces, double thick, int retest,
                     uint8_t *good) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  if (n_faces == 0) return SHERD_OK;
  if (!V || !F || !ns || !good) return fail(SHERD_E_ARG, "null buffer");
  if (!faces && n_faces != nf) return fail(SHERD_E_ARG, "without a face list, n_faces must be nf");
  if (nf > (size_t)INT_MAX / 3 || n_faces > (size_t)INT_MAX || nv > (size_t)INT_MAX / 3)
    return fail(SHERD_E_LIMIT, "mesh too large");
  if ((size_t)scene->d.n != nf) return fail(SHERD_E_ARG, "the scene must be built over every face of F");
  for (size_t i = 0; i < 3 * nf; ++i)
    if (F[i] >= nv) return fail(SHERD_E_ARG, "face index out of range");
  if (faces)
    for (size_t i = 0; i < n_faces; ++i)
      if (faces[i] < 0 || (size_t)faces[i] >= nf) return fail(SHERD_E_ARG, "listed face out of range");
  hipError_t e = hipSuccess;
  double *dV = upload3(c, 0, V, nv, &e);
  if (!dV) return hip_fail(e, "upload vertices");
  uint32_t *dF = (uint32_t *)c->scratch(1, sizeof(uint32_t) * 3 * nf, &e);
  if (!dF) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(dF, F, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, c->stream), "upload faces");
  int64_t *dfaces = nullptr;
  if (faces) {
    dfaces = (int64_t *)c->scratch(2, sizeof(int64_t) * n_faces, &e);
    if (!dfaces) return hip_fail(e, "scratch");
    HIPCK(hipMemcpyAsync(dfaces, faces, sizeof(int64_t) * n_faces, hipMemcpyHostToDevice, c->stream), "upload");
  }
  double *dns = upload3(c, 3, ns, n_faces, &e);
  if (!dns) return hip_fail(e, "upload normals");
  uint8_t *dg = (uint8_t *)c->scratch(4, n_faces, &e);
  if (!dg) return hip_fail(e, "scratch");
  ConeConst cc;
  const double PI = 3.14159265358979323846;
  double angle = 15.0 * (PI / 180.0);
  cc.cos_a = std::cos(angle);
  cc.sin_a = std::sin(angle);
  for (int k = 1; k < 7; ++k) {
    double phi = 2.0 * PI * (double)(k - 1) / 6.0;
    cc.cp[k - 1] = std::cos(phi);
    cc.sp[k - 1] = std::sin(phi);
  }
  cc.min_hit = (float)(0.1 * thick);
  cc.thick = (float)thick;
  double *dfn = (double *)c->scratch(5, sizeof(double) * 3 * nf, &e);
  if (!dfn) return hip_fail(e, "scratch");
  Timer tm(c, SHERD_M_CONE, 7 * (uint64_t)n_faces);
  if (c->prof.on) c->prof.begin("k_cone", c->stream);
  HIPCK(launch_cone(view_of(scene->d), dV, dF, (int)nf, dfaces, dns, (int)n_faces, cc, retest, dg, dfn, c->stream),
        "cone kernel");
  if (c->prof.on) c->prof.end(c->stream);
  tm.stop();
  HIPCK(hipMemcpyAsync(good, dg, n_faces, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  c->prof.collect();
  return SHERD_OK;
}

}  // extern "C"

namespace {

constexpr size_t THICK_RAYS = 300000;  // thickness.rs RAYS
constexpr uint32_t THICK_MIN_HITS = 100;

// thickness_from_hits's decisions over the device's two lists (thickness.rs:150-160): too few
// valid hits is None; the filtered mode when it has enough hits, else the plain one.
// The estimate's device state (counts, modes, the face-index flag, the selection's counts), zeroed
// on the context's stream.
ThickState *thick_state(Ctx *c, hipError_t *e) {
  ThickState *st = (ThickState *)c->scratch(8, sizeof(ThickState), e);
  if (!st) return nullptr;
  if ((*e = hipMemsetAsync(st, 0, sizeof(ThickState), c->stream)) != hipSuccess) return nullptr;
  return st;
}

int thick_modes(Ctx *c, ThickState *st, const double *dN, int nf, const uint32_t *didx, int stride, const float *dt,
                const uint32_t *dp, int m, float *t, float *thick_mode, int *estimated) {
  hipError_t e;
  float *ok_d = (float *)c->scratch(6, sizeof(float) * (size_t)(m > 0 ? m : 1), &e);
  if (!ok_d) return hip_fail(e, "scratch");
  float *far_d = (float *)c->scratch(7, sizeof(float) * (size_t)(m > 0 ? m : 1), &e);
  if (!far_d) return hip_fail(e, "scratch");
  if (c->prof.on) c->prof.begin("k_thick_modes", c->stream);
  HIPCK(launch_thick_modes(dN, nf, didx, stride, dt, dp, m, ok_d, far_d, st, c->stream), "modes");
  if (c->prof.on) c->prof.end(c->stream);
  struct {
    uint32_t counts[2];
    float modes[2];
    uint32_t bad_face;
  } h;
  static_assert(sizeof h <= offsetof(ThickState, sel), "the state's head");
  HIPCK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, c->stream), "download");
  HIPCK(hipStreamSynchronize(c->stream), "synchronize");
  if (h.bad_face) return fail(SHERD_E_ARG, "face index out of range");
  *estimated = h.counts[0] >= THICK_MIN_HITS ? 1 : 0;
  if (*estimated) {
    *thick_mode = h.modes[0];
    *t = h.counts[1] >= THICK_MIN_HITS ? h.modes[1] : h.modes[0];
  }
  return SHERD_OK;
}

}  // namespace

extern "C" {

int sherd_thickness(sherd_ctx *ctx, const sherd_scene *scene, const double *V, size_t nv, const uint32_t *F,
                    size_t nf, float *t, float *thick_mode, int *estimated) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (!scene || scene->ctx != c) return fail(SHERD_E_ARG, "scene of another context");
  if (!V || !F || !t || !thick_mode || !estimated) return fail(SHERD_E_ARG, "null buffer");
  *estimated = 0;
  if (nf == 0) return SHERD_OK;
  if (nf > (size_t)INT_MAX / 3 || nv > (size_t)INT_MAX / 3) return fail(SHERD_E_LIMIT, "mesh too large");
  if ((size_t)scene->d.n != nf) return fail(SHERD_E_ARG, "the scene must be built over every face of F");
  hipError_t e = hipSuccess;
  // Face indices are checked on the device (k_thick_normals raises the state's flag; the call then
  // fails as an out-of-range index would on the host).
  ThickState *st = thick_state(c, &e);
  if (!st) return hip_fail(e, "scratch");
  double *dV = upload3(c, 0, V, nv, &e);
  if (!dV) return hip_fail(e, "upload vertices");
  uint32_t *dF = (uint32_t *)c->scratch(1, sizeof(uint32_t) * 3 * nf, &e);
  if (!dF) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(dF, F, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, c->stream), "upload faces");
  double *dN = (double *)c->scratch(2, sizeof(double) * 3 * nf, &e);
  if (!dN) return hip_fail(e, "scratch");
  // R §3.2's stride: every face up to THICK_RAYS, else every ceil(nf / THICK_RAYS)-th.
  const int stride = (int)((nf + THICK_RAYS - 1) / THICK_RAYS);
  const int m = (int)((nf + (size_t)stride - 1) / (size_t)stride);
  float *dt = (float *)c->scratch(3, sizeof(float) * (size_t)m, &e);
  if (!dt) return hip_fail(e, "scratch");
  uint32_t *dp = (uint32_t *)c->scratch(4, sizeof(uint32_t) * (size_t)m, &e);
  if (!dp) return hip_fail(e, "scratch");
  Timer tm(c, SHERD_M_FIRST_HIT, (uint64_t)m);
  if (c->prof.on) c->prof.begin("k_thick_rays", c->stream);
  HIPCK(launch_thick_normals(dV, (int)nv, dF, (int)nf, dN, &st->bad_face, c->stream), "normals");
  HIPCK(launch_thick_rays(view_of(scene->d), dV, (int)nv, dF, dN, stride, m, dt, dp, c->stream), "rays");
  if (c->prof.on) c->prof.end(c->stream);
  tm.stop();
  int rc = thick_modes(c, st, dN, (int)nf, nullptr, stride, dt, dp, m, t, thick_mode, estimated);
  c->prof.collect();
  return rc;
}

int sherd_thickness_from_hits(sherd_ctx *ctx, const double *normals, size_t nf, const uint32_t *idx,
                              const float *t_hit, const uint32_t *prim, size_t n_rays, float *t,
                              float *thick_mode, int *estimated) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (!t || !thick_mode || !estimated) return fail(SHERD_E_ARG, "null buffer");
  *estimated = 0;
  if (n_rays == 0) return SHERD_OK;
  if (!normals || !idx || !t_hit || !prim) return fail(SHERD_E_ARG, "null buffer");
  if (nf > (size_t)INT_MAX / 3 || n_rays > (size_t)INT_MAX) return fail(SHERD_E_LIMIT, "too many faces or rays");
  for (size_t r = 0; r < n_rays; ++r)
    if (idx[r] >= nf) return fail(SHERD_E_ARG, "a ray's face is out of range");
  hipError_t e = hipSuccess;
  double *dN = upload3(c, 0, normals, nf, &e);
  if (!dN) return hip_fail(e, "upload normals");
  uint32_t *didx = (uint32_t *)c->scratch(1, sizeof(uint32_t) * n_rays, &e);
  if (!didx) return hip_fail(e, "scratch");
  float *dt = (float *)c->scratch(3, sizeof(float) * n_rays, &e);
  if (!dt) return hip_fail(e, "scratch");
  uint32_t *dp = (uint32_t *)c->scratch(4, sizeof(uint32_t) * n_rays, &e);
  if (!dp) return hip_fail(e, "scratch");
  HIPCK(hipMemcpyAsync(didx, idx, sizeof(uint32_t) * n_rays, hipMemcpyHostToDevice, c->stream), "upload");
  HIPCK(hipMemcpyAsync(dt, t_hit, sizeof(float) * n_rays, hipMemcpyHostToDevice, c->stream), "upload");
  HIPCK(hipMemcpyAsync(dp, prim, sizeof(uint32_t) * n_rays, hipMemcpyHostToDevice, c->stream), "upload");
  ThickState *st = thick_state(c, &e);
  if (!st) return hip_fail(e, "scratch");
  return thick_modes(c, st, dN, (int)nf, didx, 1, dt, dp, (int)n_rays, t, thick_mode, estimated);
}

}  // extern "C"

extern "C" {

int sherd_profile(sherd_ctx *ctx, int on) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  c->prof.on = on != 0;
  for (int i = 0; i < KernelProfiler::MAXK; ++i) {
    c->prof.ms[i] = 0.0;
    c->prof.launches[i] = 0;
  }
  return SHERD_OK;
}

int sherd_profile_read(sherd_ctx *ctx, char *buf, size_t len) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  LOCK(c);
  if (!buf || len == 0) return fail(SHERD_E_ARG, "null buffer");
  std::string out = "{";
  for (int i = 0; i < c->prof.n; ++i) {
    char item[160];
    std::snprintf(item, sizeof item, "%s\"%s\": {\"ms\": %.6f, \"launches\": %llu}", i ? ", " : "", c->prof.name[i],
                  c->prof.ms[i], (unsigned long long)c->prof.launches[i]);
    out += item;
  }
  out += "}";
  if (out.size() + 1 > len) return fail(SHERD_E_LIMIT, "buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return SHERD_OK;
}

}  // extern "C"

