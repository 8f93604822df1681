// stream_probe.hip -- achievable-bandwidth probe for this GPU (a plain copy and
// a read-only sweep, float4 accesses, hipEvent timing).  Used to put the
// engine's roofline numbers next to what a trivial kernel reaches, and as the
// known-size copy that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE.
//
//   stream_probe [MiB ...]      (default: 64 128 256 512 2048)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                    \
  do {                                                                              \
    hipError_t e = (x);                                                             \
    if (e != hipSuccess) {                                                          \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
      std::exit(1);                                                                 \
    }                                                                               \
  } while (0)

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b,
                                              size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ a, float* __restrict__ out,
                                              size_t n) {
  float s = 0.f;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    float4 v = a[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1234.5f) out[0] = s;  // keeps the loads alive; never true for zeroed input
}

int main(int argc, char** argv) {
  std::vector<size_t> mib;
  for (int i = 1; i < argc; ++i) mib.push_back(std::strtoull(argv[i], nullptr, 10));
  if (mib.empty()) mib = {64, 128, 256, 512, 2048};
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int grid = prop.multiProcessorCount * 8;
  std::printf("device %s, %d CUs\n", prop.gcnArchName, prop.multiProcessorCount);
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  for (size_t m : mib) {
    const size_t bytes = m << 20, n = bytes / sizeof(float4);
    float4 *a, *b;
    float* out;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMalloc(&out, 16));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    const int reps = 20;
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms_copy;
    CHECK(hipEventElapsedTime(&ms_copy, e0, e1));
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, 0, a, out, n);
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, 0, a, out, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms_read;
    CHECK(hipEventElapsedTime(&ms_read, e0, e1));
    const double copy_gbs = 2.0 * bytes * reps / (ms_copy * 1e-3) / 1e9;
    const double read_gbs = 1.0 * bytes * reps / (ms_read * 1e-3) / 1e9;
    std::printf("%6zu MiB  copy %8.1f GB/s (%.4f ms/launch)  read %8.1f GB/s (%.4f ms/launch)\n", m,
                copy_gbs, ms_copy / reps, read_gbs, ms_read / reps);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(out));
  }
  return 0;
}
