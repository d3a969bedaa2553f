// calibrate.hip -- kernels of known HBM traffic, for reading rocprofv3's
// FETCH_SIZE and WRITE_SIZE counters on this device.  Measurement support
// for the benchmark (built into its own library, libgm_calib.so); nothing
// in the mining library calls it.
//
// Each kernel touches every byte of its buffers exactly once, with the
// access widths the level kernels use:
//   stream16  16-byte loads and 16-byte stores, coalesced: reads N, writes N
//   gather4   4-byte loads of every word of an N-byte buffer in a scattered
//             order (consecutive lanes land on distinct cache lines); one
//             4-byte store per lane of the sums
//   atomic8   one 64-bit atomic add on every word of an N-byte buffer, in
//             the same scattered order (the counters' access pattern)
// The scattered order is word (i * kStride) mod n for i = 0 .. n-1, a
// permutation of the n words because kStride is odd and n a power of two.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace {

constexpr uint64_t kStride = 0x9E3779B97F4A7C15ull | 1;
constexpr int kThreads = 256;
constexpr int kBlocks = 4096;

__global__ void __launch_bounds__(kThreads) k_stream16(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                       uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
    dst[i] = src[i];
}

__global__ void __launch_bounds__(kThreads) k_gather4(const uint32_t *__restrict__ src, uint32_t *__restrict__ out,
                                                      uint64_t n) {
  const uint64_t t = blockIdx.x * (uint64_t)kThreads + threadIdx.x;
  uint32_t acc = 0;
  for (uint64_t i = t; i < n; i += (uint64_t)gridDim.x * kThreads) acc += src[(i * kStride) & (n - 1)];
  out[t] = acc;
}

__global__ void __launch_bounds__(kThreads) k_atomic8(unsigned long long *dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
    atomicAdd(&dst[(i * kStride) & (n - 1)], 1ull);
}

}  // namespace

// Run one calibration kernel over `bytes` (a power of two, at least 1 MiB)
// of fresh device memory; *ms gets its duration from HIP events.  Returns 0,
// or a negative value on a bad argument or a HIP error.
extern "C" int gmc_run(const char *kind, uint64_t bytes, double *ms) {
  if (!kind || bytes < (1u << 20) || (bytes & (bytes - 1))) return -1;
  const uint64_t out_bytes = (uint64_t)kBlocks * kThreads * sizeof(uint32_t);
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -2;
  if (hipMalloc(&a, bytes) != hipSuccess) return -2;
  if (hipMalloc(&b, bytes > out_bytes ? bytes : out_bytes) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  if (hipMemset(a, 1, bytes) == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
      hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    rc = 0;
    (void)hipEventRecord(e0, nullptr);
    if (!std::strcmp(kind, "stream16"))
      k_stream16<<<kBlocks, kThreads>>>(static_cast<const uint4 *>(a), static_cast<uint4 *>(b), bytes / 16);
    else if (!std::strcmp(kind, "gather4"))
      k_gather4<<<kBlocks, kThreads>>>(static_cast<const uint32_t *>(a), static_cast<uint32_t *>(b), bytes / 4);
    else if (!std::strcmp(kind, "atomic8"))
      k_atomic8<<<kBlocks, kThreads>>>(static_cast<unsigned long long *>(a), bytes / 8);
    else
      rc = -1;
    (void)hipEventRecord(e1, nullptr);
    if (rc == 0 && (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = -2;
    float f = 0.f;
    if (rc == 0 && hipEventElapsedTime(&f, e0, e1) == hipSuccess && ms) *ms = f;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
