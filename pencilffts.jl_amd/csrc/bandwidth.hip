// Streaming copy used as the measured HBM ceiling beside the pass kernels
// (bench.py) and as the known-size reference that scales rocprofv3's HBM byte
// counters.  Not part of the transform library: built apart, as
// libpfft_bandwidth.so.
//
// A persistent grid (a few workgroups per CU) walks the buffer in chunks of
// UNROLL x 256 threads x 16 bytes: every lane issues UNROLL independent
// 16-byte loads before its UNROLL stores, and the stores carry the
// nontemporal hint (the copy's output is not read again).  The tail below a
// whole 16-byte vector is copied bytewise by the first workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 8;

typedef float v4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreads) void k_stream_copy(v4* __restrict__ dst,
                                                         const v4* __restrict__ src, int64_t nvec,
                                                         char* __restrict__ dst_tail,
                                                         const char* __restrict__ src_tail,
                                                         int tail) {
  const int64_t chunk = (int64_t)kUnroll * kThreads;
  const int64_t nchunk = nvec / chunk;
  for (int64_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = c * chunk + threadIdx.x;
    v4 r[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) r[u] = src[base + (int64_t)u * kThreads];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) __builtin_nontemporal_store(r[u], &dst[base + (int64_t)u * kThreads]);
  }
  // vectors after the last whole chunk
  for (int64_t i = nchunk * chunk + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < nvec;
       i += (int64_t)gridDim.x * kThreads)
    dst[i] = src[i];
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < tail; i += kThreads) dst_tail[i] = src_tail[i];
}

}  // namespace

extern "C" {

// Copies nbytes from src to dst (device pointers, 16-byte aligned) on
// `stream` with `blocks_per_cu` resident workgroups per CU (0: 4).  Returns
// a hipError_t value (0 = success).
int pfft_bandwidth_copy(void* dst, const void* src, size_t nbytes, int blocks_per_cu,
                        void* stream) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return (int)e;
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return (int)hipErrorInvalidValue;
  const int64_t nvec = (int64_t)(nbytes / 16);
  const int tail = (int)(nbytes - (size_t)nvec * 16);
  const int per_cu = blocks_per_cu > 0 ? blocks_per_cu : 4;
  const int64_t need = (nvec + kThreads - 1) / kThreads;
  int64_t grid = (int64_t)cus * per_cu;
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)grid), dim3(kThreads), 0, (hipStream_t)stream,
                     (v4*)dst, (const v4*)src, nvec, (char*)dst + (size_t)nvec * 16,
                     (const char*)src + (size_t)nvec * 16, tail);
  return (int)hipGetLastError();
}

}  // extern "C"
