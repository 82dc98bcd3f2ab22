// The self-test that gates backend selection (crates/sherd-gpu/src/selftest.rs:179 SelfTest::run,
// read by crates/sherd-backend/src/lib.rs:230-247): four checks on an open context.
//
//   limits       what the library's kernels assume of the device
//   reduction    1e7 f32 terms summed in one fixed order on the device (256 workgroups, each a
//                lane-strided accumulate and a 256 -> 1 tree in LDS, then one workgroup over the
//                partials) against the same order on the host, to the bit (selftest.rs:405-460)
//   bounded nn   sherd_nearest_below's grid and kernel against a host grid search of the same
//                rule (nearest strictly within the bound, ties to the lower index)
//   throughput   that batch's kernel wall time against the host search on every hardware thread
//
// The host halves are the checks' own references, as the reference's mirrors are; nothing here
// answers a computation in place of the device.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "api_internal.hpp"

using namespace sherd;

namespace {

constexpr int RED_LANES = 256;
constexpr int RED_GROUPS = 256;
constexpr size_t RED_TERMS = 10000000;
constexpr int NN_POINTS = 6000;
constexpr int NN_POSES = 64;
constexpr double NN_RADIUS = 0.03;
constexpr double NN_MAX_DELTA = 1e-6;
constexpr double NN_MAX_DIFFERING_SHARE = 1e-5;

// The terms of selftest.rs:405-418: a xorshift64 stream, 24 bits into [-1, 1), every 997th term
// scaled by 1e5 so that the sum's rounding depends on the order.
std::vector<float> reduce_terms(size_t n) {
  std::vector<float> out(n);
  uint64_t s = 0x9e3779b97f4a7c15ull;
  for (size_t i = 0; i < n; ++i) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    float unit = (float)(s >> 40) * (1.0f / 16777216.0f);
    float v = unit * 2.0f - 1.0f;
    out[i] = (i % 997 == 0) ? v * 1e5f : v;
  }
  return out;
}

// One pass: `groups` partials; group g sums terms [g*block, min((g+1)*block, n)) lane-strided,
// then halves 256 -> 1.
std::vector<float> reduce_pass_host(const std::vector<float> &x, int groups) {
  size_t n = x.size();
  size_t block = (n + groups - 1) / groups;
  std::vector<float> out(groups);
  float lanes[RED_LANES];
  for (int g = 0; g < groups; ++g) {
    size_t start = (size_t)g * block, end = std::min(start + block, n);
    for (int l = 0; l < RED_LANES; ++l) {
      float acc = 0.0f;
      for (size_t i = start + l; i < end; i += RED_LANES) acc += x[i];
      lanes[l] = acc;
    }
    for (int stride = RED_LANES / 2; stride > 0; stride /= 2)
      for (int l = 0; l < stride; ++l) lanes[l] += lanes[l + stride];
    out[g] = lanes[0];
  }
  return out;
}

__global__ void __launch_bounds__(RED_LANES) k_reduce_pass(const float *x, size_t n, size_t block, float *out) {
  __shared__ float lanes[RED_LANES];
  size_t start = (size_t)blockIdx.x * block;
  size_t end = start + block < n ? start + block : n;
  float acc = 0.0f;
  for (size_t i = start + threadIdx.x; i < end; i += RED_LANES) acc += x[i];
  lanes[threadIdx.x] = acc;
  for (int stride = RED_LANES / 2; stride > 0; stride /= 2) {
    __syncthreads();
    if ((int)threadIdx.x < stride) lanes[threadIdx.x] += lanes[threadIdx.x + stride];
  }
  if (threadIdx.x == 0) out[blockIdx.x] = lanes[0];
}

uint32_t bits_of(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return b;
}

// Two clouds on a warped unit sheet and small rigid poses (a rotation about z within 0.025 rad, a
// translation within 0.01 per axis), from one xorshift stream.
struct NnCase {
  std::vector<double> target, queries;
};

NnCase nn_case() {
  uint64_t s = 0x2545f4914f6cdd1dull;
  auto next = [&s]() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  };
  auto sheet = [](double u, double v, double *p) {
    p[0] = u;
    p[1] = v;
    p[2] = 0.15 * std::sin(u * 6.3) * std::cos(v * 4.1);
  };
  NnCase c;
  c.target.resize(3 * NN_POINTS);
  std::vector<double> src(3 * NN_POINTS);
  for (int i = 0; i < NN_POINTS; ++i) {
    double u = next(), v = next();
    sheet(u, v, &c.target[3 * i]);
  }
  for (int i = 0; i < NN_POINTS; ++i) {
    double u = next(), v = next();
    sheet(u, v, &src[3 * i]);
  }
  c.queries.resize(3 * (size_t)NN_POINTS * NN_POSES);
  for (int k = 0; k < NN_POSES; ++k) {
    double a = (next() - 0.5) * 0.05, sa = std::sin(a), ca = std::cos(a);
    double t[3] = {(next() - 0.5) * 0.02, (next() - 0.5) * 0.02, (next() - 0.5) * 0.02};
    for (int i = 0; i < NN_POINTS; ++i) {
      const double *p = &src[3 * i];
      double *q = &c.queries[3 * ((size_t)k * NN_POINTS + i)];
      q[0] = ca * p[0] - sa * p[1] + t[0];
      q[1] = sa * p[0] + ca * p[1] + t[1];
      q[2] = p[2] + t[2];
    }
  }
  return c;
}

// The host grid search: cells of side `bound`, the 27 around the query's, the nearest point with
// sqrt(d2) < bound, d2 = (dx*dx + dy*dy) + dz*dz, ties to the lower index.
struct HostGrid {
  double r;
  std::unordered_map<uint64_t, std::vector<int>> cells;
  const double *p;
  static uint64_t key(int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)(x + (1 << 20)) << 42) | ((uint64_t)(y + (1 << 20)) << 21) | (uint64_t)(z + (1 << 20));
  }
  HostGrid(const double *pts, int n, double bound) : r(bound), p(pts) {
    for (int i = 0; i < n; ++i)
      cells[key((int64_t)std::floor(pts[3 * i] / r), (int64_t)std::floor(pts[3 * i + 1] / r),
                (int64_t)std::floor(pts[3 * i + 2] / r))]
          .push_back(i);
  }
  void nearest(const double *q, int64_t *idx, double *dist) const {
    int64_t cx = (int64_t)std::floor(q[0] / r), cy = (int64_t)std::floor(q[1] / r), cz = (int64_t)std::floor(q[2] / r);
    int64_t best = -1;
    double bd = INFINITY;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          auto it = cells.find(key(cx + dx, cy + dy, cz + dz));
          if (it == cells.end()) continue;
          for (int i : it->second) {
            double ex = p[3 * i] - q[0], ey = p[3 * i + 1] - q[1], ez = p[3 * i + 2] - q[2];
            double d2 = (ex * ex + ey * ey) + ez * ez;
            if (std::sqrt(d2) < r && (d2 < bd || (d2 == bd && i < best))) {
              bd = d2;
              best = i;
            }
          }
        }
    *idx = best;
    *dist = best < 0 ? INFINITY : std::sqrt(bd);
  }
};

void set_check(sherd_check *c, const char *name, bool ok, const std::string &detail) {
  std::snprintf(c->name, sizeof c->name, "%s", name);
  c->passed = ok ? 1 : 0;
  std::snprintf(c->detail, sizeof c->detail, "%s", detail.c_str());
}

}  // namespace

extern "C" int sherd_selftest(sherd_ctx *ctx, sherd_selftest_report *out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(SHERD_E_ARG, "null report");
  std::memset(out, 0, sizeof *out);
  char buf[256];
  // 1. Limits.
  {
    if (!c) return fail(SHERD_E_ARG, "null context");
    hipDeviceProp_t prop;
    HIPCK(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
    bool ok = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 && prop.warpSize == 64 &&
              prop.sharedMemPerBlock >= 65536 && prop.maxThreadsPerBlock >= 1024;
    std::snprintf(buf, sizeof buf, "%s, %d CUs, wave %d, %zu B LDS per workgroup, %d lanes per workgroup",
                  prop.gcnArchName, prop.multiProcessorCount, prop.warpSize, (size_t)prop.sharedMemPerBlock,
                  prop.maxThreadsPerBlock);
    set_check(&out->checks[0], "limits", ok, buf);
  }
  std::vector<float> terms = reduce_terms(RED_TERMS);
  std::vector<float> host_partials = reduce_pass_host(terms, RED_GROUPS);
  float want = reduce_pass_host(host_partials, 1)[0];
  NnCase nc = nn_case();
  const size_t m = (size_t)NN_POINTS * NN_POSES;
  std::vector<int64_t> di_h(m);
  std::vector<double> dd_h(m);
  double gpu_ms = 0.0;
  {
    LOCK(c);
    // 2. The reduction, on the device.
    hipError_t e;
    float *dx = (float *)c->scratch(0, sizeof(float) * RED_TERMS, &e);
    if (!dx) return hip_fail(e, "scratch");
    float *dp = (float *)c->scratch(1, sizeof(float) * (RED_GROUPS + 1), &e);
    if (!dp) return hip_fail(e, "scratch");
    HIPCK(hipMemcpyAsync(dx, terms.data(), sizeof(float) * RED_TERMS, hipMemcpyHostToDevice, c->stream), "upload");
    hipLaunchKernelGGL(k_reduce_pass, dim3(RED_GROUPS), dim3(RED_LANES), 0, c->stream, dx, RED_TERMS,
                       (RED_TERMS + RED_GROUPS - 1) / RED_GROUPS, dp);
    HIPCK(hipGetLastError(), "reduce kernel");
    hipLaunchKernelGGL(k_reduce_pass, dim3(1), dim3(RED_LANES), 0, c->stream, dp, (size_t)RED_GROUPS,
                       (size_t)RED_GROUPS, dp + RED_GROUPS);
    HIPCK(hipGetLastError(), "reduce kernel");
    float got = 0.0f;
    HIPCK(hipMemcpyAsync(&got, dp + RED_GROUPS, sizeof got, hipMemcpyDeviceToHost, c->stream), "download");
    HIPCK(hipStreamSynchronize(c->stream), "synchronize");
    out->reduce_device_bits = bits_of(got);
    out->reduce_host_bits = bits_of(want);
    std::snprintf(buf, sizeof buf, "%zu terms, %d workgroups: gpu %.9g (0x%08x), cpu %.9g (0x%08x)", RED_TERMS,
                  RED_GROUPS, (double)got, bits_of(got), (double)want, bits_of(want));
    set_check(&out->checks[1], "reduction", bits_of(got) == bits_of(want), buf);
    // 3. The bounded-NN kernel of sherd_nearest_below over the whole batch.
    double *dc = (double *)c->scratch(2, sizeof(double) * 3 * NN_POINTS, &e);
    double *dq = dc ? (double *)c->scratch(3, sizeof(double) * 3 * m, &e) : nullptr;
    int64_t *di = dq ? (int64_t *)c->scratch(4, sizeof(int64_t) * m, &e) : nullptr;
    double *dd = di ? (double *)c->scratch(5, sizeof(double) * m, &e) : nullptr;
    if (!dd) return hip_fail(e, "scratch");
    HIPCK(hipMemcpyAsync(dc, nc.target.data(), sizeof(double) * 3 * NN_POINTS, hipMemcpyHostToDevice, c->stream),
          "upload");
    HIPCK(hipMemcpyAsync(dq, nc.queries.data(), sizeof(double) * 3 * m, hipMemcpyHostToDevice, c->stream), "upload");
    DevGrid g;
    struct Guard {
      DevGrid *g;
      ~Guard() { grid_free(g); }
    } guard{&g};
    HIPCK(grid_build(dc, NN_POINTS, NN_RADIUS, &g, c->stream), "grid build");
    // Warm once, then time the kernel alone (host wall time around launch and wait).
    HIPCK(launch_nearest(view_of(g), dq, (int)m, NN_RADIUS, di, dd, c->stream), "nearest kernel");
    HIPCK(hipStreamSynchronize(c->stream), "synchronize");
    auto t0 = std::chrono::steady_clock::now();
    HIPCK(launch_nearest(view_of(g), dq, (int)m, NN_RADIUS, di, dd, c->stream), "nearest kernel");
    HIPCK(hipStreamSynchronize(c->stream), "synchronize");
    gpu_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    HIPCK(hipMemcpyAsync(di_h.data(), di, sizeof(int64_t) * m, hipMemcpyDeviceToHost, c->stream), "download");
    HIPCK(hipMemcpyAsync(dd_h.data(), dd, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream), "download");
    SYNC_RECEIPT(c);
  }
  // The host search over every hardware thread, warmed the same way.
  HostGrid hg(nc.target.data(), NN_POINTS, NN_RADIUS);
  std::vector<int64_t> hi(m);
  std::vector<double> hd(m);
  int threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  auto host_batch = [&]() {
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; ++w)
      pool.emplace_back([&, w]() {
        for (size_t i = w; i < m; i += threads) hg.nearest(&nc.queries[3 * i], &hi[i], &hd[i]);
      });
    for (auto &th : pool) th.join();
  };
  host_batch();
  auto t0 = std::chrono::steady_clock::now();
  host_batch();
  double cpu_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  size_t differing = 0, misses = 0, hits = 0;
  double max_delta = 0.0, max_tie = 0.0;
  for (size_t i = 0; i < m; ++i) {
    if ((di_h[i] < 0) != (hi[i] < 0)) {
      ++misses;
      continue;
    }
    if (hi[i] < 0) continue;
    ++hits;
    double d = std::fabs(dd_h[i] - hd[i]);
    max_delta = std::fmax(max_delta, d);
    if (di_h[i] != hi[i]) {
      ++differing;
      max_tie = std::fmax(max_tie, d);
    }
  }
  double share = (double)differing / (double)m;
  std::snprintf(buf, sizeof buf,
                "%zu queries, %zu hits: %zu differing neighbours (%.2e, limit %.0e), each a tie to %.3g; %zu hit/miss "
                "disagreements; max |delta d| %.3g (limit %.0e)",
                m, hits, differing, share, NN_MAX_DIFFERING_SHARE, max_tie, misses, max_delta, NN_MAX_DELTA);
  set_check(&out->checks[2], "bounded nn",
            misses == 0 && max_delta <= NN_MAX_DELTA && max_tie <= NN_MAX_DELTA && share <= NN_MAX_DIFFERING_SHARE &&
                hits > m / 2,
            buf);
  // 4. Throughput: recorded, never a failure.
  double speedup = gpu_ms > 0.0 ? cpu_ms / gpu_ms : 0.0;
  std::snprintf(buf, sizeof buf, "gpu %.3f ms (%.2f ns/query), cpu %.2f ms over %d threads: %.1fx", gpu_ms,
                gpu_ms * 1e6 / (double)m, cpu_ms, threads, speedup);
  set_check(&out->checks[3], "throughput", true, buf);
  out->n_checks = 4;
  out->passed = 1;
  for (int k = 0; k < 4; ++k) out->passed &= out->checks[k].passed;
  out->queries = m;
  out->gpu_ms = gpu_ms;
  out->cpu_ms = cpu_ms;
  out->speedup = speedup;
  out->cpu_threads = threads;
  return SHERD_OK;
}
