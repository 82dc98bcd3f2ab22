// Calibration program for hardware byte counters (not part of the library): one launch of a plain
// 16-byte-vector copy kernel that reads and writes a known number of bytes, so that rocprofv3's
// FETCH_SIZE and WRITE_SIZE readings of the library's kernels can be scaled to bytes.
//
//     sherd-refit_amd/calib/copy_probe [MiB]      (default 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char **argv) {
  size_t mib = argc > 1 ? (size_t)std::strtoull(argv[1], nullptr, 10) : 1024;
  if (mib == 0 || mib > 65536) {
    std::fprintf(stderr, "copy_probe: MiB must be in 1..65536\n");
    return 2;
  }
  size_t bytes = mib << 20, n = bytes / sizeof(uint4);
  uint4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
    std::fprintf(stderr, "copy_probe: hipMalloc of 2 x %zu bytes failed\n", bytes);
    return 1;
  }
  // Written first by a copy engine so the kernel reads defined memory; the kernel is the one launch.
  if (hipMemcpy(src, dst, bytes, hipMemcpyDeviceToDevice) != hipSuccess) return 1;
  int blocks = (int)((n + 255) / 256 < (size_t)256 * 64 ? (n + 255) / 256 : (size_t)256 * 64);
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, src, dst, n);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    std::fprintf(stderr, "copy_probe: %s\n", hipGetErrorString(e));
    return 1;
  }
  std::printf("copy probe: %zu bytes read and %zu bytes written by one launch of k_copy_probe\n", n * 16, n * 16);
  (void)hipFree(src);
  (void)hipFree(dst);
  return 0;
}
