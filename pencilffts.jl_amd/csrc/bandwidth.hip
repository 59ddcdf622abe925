// Streaming copy used as the measured HBM ceiling beside the pass kernels
// (bench.py) and as the known-size reference that scales rocprofv3's HBM byte
// counters.  Not part of the transform library: built apart, as
// libpfft_bandwidth.so.
//
// A grid of workgroups walks the buffer in chunks of UNROLL x 256 threads x
// 16 bytes: every lane issues UNROLL independent 16-byte loads before its
// UNROLL stores, and the stores carry the nontemporal hint (the copy's output
// is not read again); NTL puts the hint on the loads too.  The grid is either
// persistent (a given number of resident workgroups per CU, each walking
// every grid-th chunk) or one workgroup per chunk.  The bench sweeps these
// shapes and keeps the fastest, so the ceiling is the best this copy does on
// the device, not one arbitrary configuration.  The tail below a whole
// 16-byte vector is copied bytewise by the first workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 256;

typedef float v4 __attribute__((ext_vector_type(4)));

template <int UNROLL, bool NTL>
__global__ __launch_bounds__(kThreads) void k_stream_copy(v4* __restrict__ dst,
                                                         const v4* __restrict__ src, int64_t nvec,
                                                         char* __restrict__ dst_tail,
                                                         const char* __restrict__ src_tail,
                                                         int tail) {
  const int64_t chunk = (int64_t)UNROLL * kThreads;
  const int64_t nchunk = nvec / chunk;
  for (int64_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = c * chunk + threadIdx.x;
    v4 r[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      if constexpr (NTL) r[u] = __builtin_nontemporal_load(&src[base + (int64_t)u * kThreads]);
      else r[u] = src[base + (int64_t)u * kThreads];
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u)
      __builtin_nontemporal_store(r[u], &dst[base + (int64_t)u * kThreads]);
  }
  // vectors after the last whole chunk
  for (int64_t i = nchunk * chunk + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < nvec;
       i += (int64_t)gridDim.x * kThreads)
    dst[i] = src[i];
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < tail; i += kThreads) dst_tail[i] = src_tail[i];
}

template <int UNROLL, bool NTL>
hipError_t launch(int64_t grid, hipStream_t s, v4* dst, const v4* src, int64_t nvec,
                  char* dt, const char* st, int tail) {
  hipLaunchKernelGGL((k_stream_copy<UNROLL, NTL>), dim3((unsigned)grid), dim3(kThreads), 0, s,
                     dst, src, nvec, dt, st, tail);
  return hipGetLastError();
}

}  // namespace

extern "C" {

// Copies nbytes from src to dst (device pointers, 16-byte aligned) on
// `stream`.  blocks_per_cu > 0: that many resident workgroups per CU, each
// walking the chunks; 0: one workgroup per chunk.  unroll: 4, 8 or 16
// vectors in flight per lane.  nt_loads: nontemporal hint on the loads.
// Returns a hipError_t value (0 = success).
int pfft_bandwidth_copy(void* dst, const void* src, size_t nbytes, int blocks_per_cu,
                        int unroll, int nt_loads, void* stream) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return (int)e;
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return (int)hipErrorInvalidValue;
  if (unroll != 4 && unroll != 8 && unroll != 16) return (int)hipErrorInvalidValue;
  const int64_t nvec = (int64_t)(nbytes / 16);
  const int tail = (int)(nbytes - (size_t)nvec * 16);
  const int64_t chunks = (nvec + (int64_t)unroll * kThreads - 1) / ((int64_t)unroll * kThreads);
  int64_t grid = blocks_per_cu > 0 ? (int64_t)cus * blocks_per_cu : chunks;
  if (grid > chunks) grid = chunks;
  if (grid > (int64_t)1 << 30) grid = (int64_t)1 << 30;
  if (grid < 1) grid = 1;
  v4* d = (v4*)dst;
  const v4* s = (const v4*)src;
  char* dt = (char*)dst + (size_t)nvec * 16;
  const char* st = (const char*)src + (size_t)nvec * 16;
  hipStream_t hs = (hipStream_t)stream;
  if (unroll == 4)
    e = nt_loads ? launch<4, true>(grid, hs, d, s, nvec, dt, st, tail)
                 : launch<4, false>(grid, hs, d, s, nvec, dt, st, tail);
  else if (unroll == 8)
    e = nt_loads ? launch<8, true>(grid, hs, d, s, nvec, dt, st, tail)
                 : launch<8, false>(grid, hs, d, s, nvec, dt, st, tail);
  else
    e = nt_loads ? launch<16, true>(grid, hs, d, s, nvec, dt, st, tail)
                 : launch<16, false>(grid, hs, d, s, nvec, dt, st, tail);
  return (int)e;
}

}  // extern "C"
