// hbm_copy: a plain device-to-device copy of a known size, used to calibrate the HBM traffic
// counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports per dispatch. Each launch reads and
// writes exactly `bytes` bytes once, 16 bytes per lane per iteration, grid-stride.
//
// Usage: hbm_copy [MiB=1024] [launches=5]; prints bytes per launch and the mean launch time.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) hbm_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       size_t count) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

#define CHECK(x)                                                                   \
  do {                                                                             \
    const hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "hbm_copy: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  const int launches = argc > 2 ? std::atoi(argv[2]) : 5;
  if (mib == 0 || mib > 16384 || launches <= 0 || launches > 1000) {
    std::fprintf(stderr, "hbm_copy: MiB must be in [1, 16384] and launches in [1, 1000]\n");
    return 2;
  }
  const size_t bytes = mib << 20;
  const size_t count = bytes / sizeof(uint4);
  uint4 *src = nullptr, *dst = nullptr;
  CHECK(hipMalloc(&src, bytes));
  CHECK(hipMalloc(&dst, bytes));
  CHECK(hipMemset(src, 1, bytes));
  CHECK(hipMemset(dst, 0, bytes));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int blocks = prop.multiProcessorCount * 8;
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  hbm_copy_kernel<<<blocks, 256>>>(src, dst, count);  // warm
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(a));
  for (int i = 0; i < launches; ++i) hbm_copy_kernel<<<blocks, 256>>>(src, dst, count);
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, a, b));
  const double per = ms / launches;
  std::printf("{\"bytes_read\": %zu, \"bytes_written\": %zu, \"launches\": %d, \"ms_per_launch\": %.4f, "
              "\"gbs\": %.1f}\n",
              bytes, bytes, launches + 1, per, 2.0 * static_cast<double>(bytes) / (per * 1e6));
  CHECK(hipFree(src));
  CHECK(hipFree(dst));
  return 0;
}
