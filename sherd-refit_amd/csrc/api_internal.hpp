// Internal host structures behind the C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sherd_hip.h"
#include "kernels.hpp"

namespace sherd {

struct Ctx {
  int device = 0;
  int cus = 0;
  std::string name;
  hipStream_t stream = nullptr;
  SideStreams side;  // verification's independent sections
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex mu;
  sherd_method_stats stats[SHERD_M_COUNT] = {};
  struct Buf {
    void *p = nullptr;
    size_t n = 0;
  };
  Buf bufs[40];
  KernelProfiler prof;
  // Batched ICP: the workgroup size (0 = autotuned, 1 = the wide form), the autotuner's choices
  // by estimation, source-size bucket and candidate-count bucket, and the captured ladder graph
  // with the launch signature it was captured for.
  int icp_block = 0;
  int icp_tuned[2][40][24] = {};
  hipGraphExec_t ladder_exec = nullptr;
  std::string ladder_sig;   // the captured graph's signature
  std::string ladder_seen;  // the last call's signature

  // A grow-only device buffer per slot, reused across calls (the context is serialised).
  void *scratch(int slot, size_t bytes, hipError_t *err);
  // A grow-only pinned host buffer for uploads and downloads that must not go through pageable
  // staging; its previous contents are not kept when it grows.
  Buf host;
  void *pinned(size_t bytes, hipError_t *err);
  // The device block of the last freed Huber batch, kept for the next one (a batch is created and
  // freed per LM call; this spares a hipMalloc and a hipFree per call).
  Buf huber_spare;
  // This device's copy of the file-scope `scratch` array of the no-allocation entry point (its
  // address from hipGetSymbolAddress, taken at open with the device current).
  void *noalloc = nullptr;
  // The failure contract: counters, an armed fault, and the receipt word (device and pinned host).
  // sherd_health's fields in order (atomic: an entry point counts outside its locked sections).
  enum { H_CALLS, H_FAILED, H_ARG, H_HOST, H_NOMEM, H_LIMIT, H_CORRUPT, H_INJECTED, H_COUNT };
  std::atomic<uint64_t> health[H_COUNT] = {};
  int fault_kind = SHERD_FAULT_NONE;
  int fault_after = 0;
  uint32_t nonce = 0;
  uint32_t *receipt_dev = nullptr;
  uint32_t *receipt_host = nullptr;
};

// Whether the armed fault of `kind` fires at this opportunity (counting the armed countdown down).
bool take_fault(Ctx *c, int kind);
// Queues the receipt kernel (the context's next nonce + 1, or the stale nonce when a corrupt fault
// fires) and its readback behind everything on the stream, synchronises, and checks it.
int sync_receipt(Ctx *c);

// Counts a call and, when it fails, its status (any status a failure helper set on this thread
// during the call).
class CallScope {
 public:
  explicit CallScope(Ctx *c);
  ~CallScope();
  CallScope(const CallScope &) = delete;
  CallScope &operator=(const CallScope &) = delete;

 private:
  Ctx *c_;
  uint64_t n0_;
};

// Accumulates one method's call count, item count and device time (HIP events on the stream).
class Timer {
 public:
  Timer(Ctx *c, int method, uint64_t items);
  ~Timer() { stop(); }
  void stop();

 private:
  Ctx *c_;
  int m_;
  bool done_ = false;
};

int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

// Synchronises a stream when it goes out of scope. Declared after the host buffers that queued
// copies read from or write to, it keeps them alive until the stream has drained, on every return
// path (an early error return included).
struct StreamFence {
  hipStream_t s;
  explicit StreamFence(hipStream_t stream) : s(stream) {}
  StreamFence(const StreamFence &) = delete;
  StreamFence &operator=(const StreamFence &) = delete;
  ~StreamFence() { (void)hipStreamSynchronize(s); }
};

}  // namespace sherd

// An entry point's prologue: the context's lock, its device, and a fence that drains the context's
// stream before the lock is released, so no copy the call queued outlives the call (the caller's
// buffers and the pinned staging buffer are free to change once it returns).
#define LOCK(ctx)                                                                              \
  if (!(ctx)) return ::sherd::fail(SHERD_E_ARG, "null context");                               \
  std::lock_guard<std::mutex> lock_((ctx)->mu);                                                \
  ::sherd::StreamFence fence_((ctx)->stream);                                                  \
  ::sherd::CallScope scope_(ctx);                                                              \
  if (hipSetDevice((ctx)->device) != hipSuccess) return ::sherd::hip_fail(hipGetLastError(), "hipSetDevice"); \
  if (::sherd::take_fault((ctx), SHERD_FAULT_HOST)) return ::sherd::fail(SHERD_E_HIP, "injected host error")

// For an entry point whose work spans several locked sections (RELOCK): counts the call and offers
// the host fault once, up front.
#define CALL_SCOPE(ctx)                                                                        \
  if (!(ctx)) return ::sherd::fail(SHERD_E_ARG, "null context");                               \
  ::sherd::CallScope scope_(ctx);                                                              \
  {                                                                                            \
    std::lock_guard<std::mutex> l_((ctx)->mu);                                                 \
    if (::sherd::take_fault((ctx), SHERD_FAULT_HOST)) return ::sherd::fail(SHERD_E_HIP, "injected host error"); \
  }

// LOCK for a section of an entry point that opened with CALL_SCOPE: the same lock, fence and
// device, without counting the call again or offering a host fault.
#define RELOCK(ctx)                                                                            \
  std::lock_guard<std::mutex> lock_((ctx)->mu);                                                \
  ::sherd::StreamFence fence_((ctx)->stream);                                                  \
  if (hipSetDevice((ctx)->device) != hipSuccess) return ::sherd::hip_fail(hipGetLastError(), "hipSetDevice")

// An entry point's last synchronisation, with its receipt checked.
#define SYNC_RECEIPT(ctx)                                     \
  do {                                                        \
    int r_ = ::sherd::sync_receipt(ctx);                      \
    if (r_ != SHERD_OK) return r_;                            \
  } while (0)

#define HIPCK(call, what)                                    \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return ::sherd::hip_fail(e_, what); \
  } while (0)

struct sherd_scene {
  sherd::Ctx *ctx = nullptr;
  sherd::DevScene d;
};

struct sherd_cloud {
  sherd::Ctx *ctx = nullptr;
  double *pts = nullptr;
  double *normals = nullptr;
  int n = 0;
  double centroid[3] = {0, 0, 0};
  sherd::DevScene tree;         // point LBVH (each leaf one point); absent on a source-only cloud
  std::vector<uint32_t> idx;    // a fracture cloud's vertex indices, in summation order
};

struct sherd_frames {
  sherd::Ctx *ctx = nullptr;
  double *p = nullptr, *ns = nullptr, *f = nullptr, *t = nullptr, *dih = nullptr;
  double *sub_p = nullptr, *sub_n = nullptr;  // the subset's points and shell normals
  uint32_t *sub = nullptr;
  int n = 0, n_sub = 0;
  sherd::DevGrid grid;  // over p, rebuilt when a batch needs a larger reach
  double reach = 0.0;
};

namespace sherd {
// A cloud over device arrays the cloud takes over (freed with it): the centroid from a copy of the
// points (summed in index order, as sherd_cloud_create does), and the point LBVH when `tree`.
// Synchronises the context's stream. On failure the arrays are freed too.
int cloud_adopt(Ctx *c, double *dpts, double *dnormals, size_t n, bool tree, sherd_cloud **out);
// The LBVH over a cloud's device points (each leaf a point); synchronises the stream.
hipError_t point_tree(Ctx *c, sherd_cloud *cl);
void free_cloud(sherd_cloud *cl);
}  // namespace sherd

struct sherd_fragment {
  sherd::Ctx *ctx = nullptr;
  const sherd_scene *frac = nullptr;
  const sherd_scene *mesh = nullptr;
  int watertight = 0;
  double frac_area = 0.0;
  double *pf = nullptr, *s = nullptr, *brk_p = nullptr, *brk_ns = nullptr, *margin_p = nullptr, *margin_n = nullptr;
  double *pf_blk = nullptr, *s_blk = nullptr;  // bounding spheres of Pf's and S's 64-point blocks
  double pf_all[4] = {0, 0, 0, -1}, s_all[4] = {0, 0, 0, -1};  // and of each whole set
  double brk_all[4] = {0, 0, 0, -1}, margin_all[4] = {0, 0, 0, -1};
  int n_pf = 0, n_s = 0, n_brk = 0, n_margin = 0;
  sherd::DevGrid margin_grid;  // rebuilt when a batch needs a larger reach
  double margin_reach = 0.0;
  sherd::DevGrid brk_grid;     // over brk_p in the fragment's frame, likewise
  double brk_reach = 0.0;
};
