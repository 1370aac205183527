// Diagnostics: a plain streaming copy kernel that calibrates HBM counters
// and measures the achievable copy bandwidth of the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/skx.h"

namespace {

// One 16-byte element per thread over an exact grid: every wave's loads are
// issued at once, which keeps more bytes in flight than a grid-stride loop
// (6.2 against 4.6 TB/s for 1 GiB on MI355X).
__global__ void k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" int skx_diag_stream_copy(int32_t device, size_t bytes, int32_t reps, double* gbs, double* ms_per_copy) {
  if (bytes < 16 || reps < 1) return SKX_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return SKX_EHIP;
  const size_t n = bytes / sizeof(float4);
  const int threads = 256;
  if ((n + threads - 1) / threads > 0x7fffffffu) return SKX_EINVAL;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess) return SKX_ENOMEM;
  if (hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
    (void)hipFree(a);
    return SKX_ENOMEM;
  }
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    return SKX_EHIP;
  }
  (void)hipMemsetAsync(a, 0, n * sizeof(float4), s);
  const unsigned blocks = (unsigned)((n + threads - 1) / threads);
  hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(threads), 0, s, a, b, n);  // warm-up
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, s);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(threads), 0, s, a, b, n);
  (void)hipEventRecord(e1, s);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) return SKX_EHIP;
  const double per = ms / reps;
  if (ms_per_copy) *ms_per_copy = per;
  if (gbs) *gbs = 2.0 * (double)(n * sizeof(float4)) / (per * 1e-3) / 1e9;
  return SKX_OK;
}

extern "C" int skx_device_info(int32_t device, char* arch, int32_t n, int32_t* cu_count) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SKX_EHIP;
  if (arch && n > 0) {
    std::string a(prop.gcnArchName);
    const size_t colon = a.find(':');  // drop feature suffixes (":sramecc+:xnack-")
    if (colon != std::string::npos) a.resize(colon);
    const size_t m = std::min(a.size(), (size_t)n - 1);
    a.copy(arch, m);
    arch[m] = '\0';
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return SKX_OK;
}
