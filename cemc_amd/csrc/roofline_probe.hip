// Roofline probes for gfx950: a streaming HBM copy of known size and an LDS
// read kernel.  They measure the two ceilings the sweep kernel is priced
// against and give the byte counters (FETCH_SIZE / WRITE_SIZE) a kernel of
// known traffic to be scaled by.  Built as its own library
// (libcemcprobe.so); nothing in the engine depends on it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#define PROBE_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string g_err;

// grid-stride 16-byte copy: reads `n` float4 and writes `n` float4 once each
__global__ __launch_bounds__(256) void hbm_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                       uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * 256u;
  for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// every lane reads 16 B from LDS per iteration, consecutive lanes on
// consecutive 16-byte words (no bank conflicts).  The index passes through an
// empty asm statement each iteration: without it the compiler sees that the
// addresses repeat every 8 iterations and hoists the reads out of the loop.
__global__ __launch_bounds__(256) void lds_read_kernel(uint32_t* __restrict__ out, int iters) {
  __shared__ uint4 buf[1024];  // 16 KiB: up to 8 workgroups per CU
  const int t = threadIdx.x;
  for (int k = t; k < 1024; k += 256) buf[k] = make_uint4(k, 3u * k + 1u, 5u * k + 2u, 7u * k + 3u);
  __syncthreads();
  uint32_t acc = 0;
  uint32_t idx = uint32_t(t);
  for (int it = 0; it < iters; it += 2) {  // two independent reads in flight
    asm volatile("" : "+v"(idx));
    const uint4 v = buf[idx & 1023u];
    const uint4 w = buf[(idx + 512u) & 1023u];
    acc ^= (v.x ^ v.y) ^ (v.z ^ v.w) ^ (w.x ^ w.y) ^ (w.z ^ w.w);
    idx += 256u;
  }
  // written only when the xor vanishes, which the compiler cannot rule out
  if (acc == 0x9e3779b9u) out[blockIdx.x * 256 + t] = acc;
}

bool fail(hipError_t e, const char* what) {
  if (e == hipSuccess) return false;
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return true;
}

}  // namespace

PROBE_API const char* probe_last_error(void) { return g_err.c_str(); }

// Copies `bytes` (rounded down to 16) device to device `reps` times; returns
// the mean kernel time of one launch in ms.  Traffic per launch: `bytes` read
// plus `bytes` written.
PROBE_API int probe_hbm_copy(uint64_t bytes, int reps, double* ms_per_launch) {
  if (!ms_per_launch || reps < 1 || bytes < 16) {
    g_err = "invalid argument";
    return 1;
  }
  const uint64_t n = bytes / 16;
  float4 *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 2;
  do {
    if (fail(hipMalloc(&src, n * 16), "hipMalloc")) break;
    if (fail(hipMalloc(&dst, n * 16), "hipMalloc")) break;
    if (fail(hipMemset(src, 1, n * 16), "hipMemset")) break;
    if (fail(hipEventCreate(&e0), "hipEventCreate") || fail(hipEventCreate(&e1), "hipEventCreate")) break;
    const unsigned grid = 256u * 16u;  // 16 workgroups per CU
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(grid), dim3(256), 0, nullptr, src, dst, n);  // warm
    if (fail(hipGetLastError(), "launch")) break;
    if (fail(hipEventRecord(e0, nullptr), "hipEventRecord")) break;
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(hbm_copy_kernel, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
    if (fail(hipEventRecord(e1, nullptr), "hipEventRecord")) break;
    if (fail(hipEventSynchronize(e1), "hipEventSynchronize")) break;
    float ms = 0.f;
    if (fail(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) break;
    *ms_per_launch = double(ms) / reps;
    rc = 0;
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

// LDS read rate: `blocks` workgroups of 256 lanes, `iters` (even) 16-byte
// reads per lane.  Returns the kernel time in ms and the bytes read from LDS.
PROBE_API int probe_lds_read(int blocks, int iters, double* ms, double* bytes) {
  if (!ms || !bytes || blocks < 1 || iters < 2 || (iters & 1)) {
    g_err = "invalid argument";
    return 1;
  }
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 2;
  do {
    if (fail(hipMalloc(&out, sizeof(uint32_t) * 256u * uint64_t(blocks)), "hipMalloc")) break;
    if (fail(hipEventCreate(&e0), "hipEventCreate") || fail(hipEventCreate(&e1), "hipEventCreate")) break;
    hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(256), 0, nullptr, out, 64);  // warm
    if (fail(hipGetLastError(), "launch")) break;
    if (fail(hipEventRecord(e0, nullptr), "hipEventRecord")) break;
    hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(256), 0, nullptr, out, iters);
    if (fail(hipEventRecord(e1, nullptr), "hipEventRecord")) break;
    if (fail(hipEventSynchronize(e1), "hipEventSynchronize")) break;
    float t = 0.f;
    if (fail(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime")) break;
    *ms = t;
    *bytes = 16.0 * 256.0 * double(blocks) * double(iters);
    rc = 0;
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  return rc;
}
