// stream_probe.hip — the achievable HBM streaming rate on this device, the
// ceiling the loops' achieved bandwidth is read against beside the 8 TB/s
// peak.  Copy kernels of several shapes over buffers far larger than the L2
// and the 256 MiB Infinity Cache, plus read-only and write-only forms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "op2_hip_rt.h"

namespace {

// v0: one 16-byte vector per thread, a grid covering the buffer
__global__ void __launch_bounds__(256) k_copy1(const double2 *__restrict s,
                                               double2 *__restrict d, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// v1/v2: U vectors per thread, consecutive lanes on consecutive vectors for
// each of the U loads (coalesced), all loads issued before the stores;
// NT: non-temporal loads and stores
template <int U, bool NT>
__global__ void __launch_bounds__(256) k_copyU(const double2 *__restrict s,
                                               double2 *__restrict d, size_t n) {
  const size_t base = blockIdx.x * (size_t)blockDim.x * U + threadIdx.x;
  double2 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n) {
      if (NT) {
        v[u].x = __builtin_nontemporal_load(&s[i].x);
        v[u].y = __builtin_nontemporal_load(&s[i].y);
      } else {
        v[u] = s[i];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n) {
      if (NT) {
        __builtin_nontemporal_store(v[u].x, &d[i].x);
        __builtin_nontemporal_store(v[u].y, &d[i].y);
      } else {
        d[i] = v[u];
      }
    }
  }
}

// v4: a persistent grid striding over the buffer, 4 vectors per iteration
__global__ void __launch_bounds__(256) k_copy_stride(const double2 *__restrict s,
                                                     double2 *__restrict d, size_t n) {
  const size_t step = (size_t)gridDim.x * blockDim.x * 4;
  for (size_t base = blockIdx.x * (size_t)blockDim.x * 4 + threadIdx.x; base < n;
       base += step) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (base + (size_t)u * blockDim.x < n) v[u] = s[base + (size_t)u * blockDim.x];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (base + (size_t)u * blockDim.x < n) d[base + (size_t)u * blockDim.x] = v[u];
  }
}

// v5: read only (a sum per thread written once per thread, 1/U of the reads)
template <int U>
__global__ void __launch_bounds__(256) k_read(const double2 *__restrict s,
                                              double *__restrict sink, size_t n) {
  const size_t base = blockIdx.x * (size_t)blockDim.x * U + threadIdx.x;
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n) {
      const double2 v = s[i];
      acc += v.x + v.y;
    }
  }
  if (acc == 1234.5) sink[0] = acc;  // keeps the loads; never true for the data
}

// v6: write only
template <int U>
__global__ void __launch_bounds__(256) k_fill(double2 *__restrict d, size_t n) {
  const size_t base = blockIdx.x * (size_t)blockDim.x * U + threadIdx.x;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n) d[i] = make_double2(1.0, 2.0);
  }
}

#define PCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2h_stream_probe: %s: %s\n", #x, hipGetErrorString(e_));      \
      return -2;                                                             \
    }                                                                        \
  } while (0)

}  // namespace

extern "C" int op2h_stream_probe(size_t bytes, int variant, int reps, double *gbs,
                                 double *us_per_call) {
  if (bytes < (1u << 20) || variant < 0 || variant >= OP2H_STREAM_VARIANTS || reps < 1 || !gbs)
    return -1;
  const size_t n = bytes / 16;
  double2 *s = nullptr, *d = nullptr;
  double *sink = nullptr;
  PCHECK(hipMalloc(&s, n * 16));
  PCHECK(hipMalloc(&d, n * 16));
  PCHECK(hipMalloc(&sink, 64));
  hipStream_t st = nullptr;
  PCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  k_fill<4><<<(unsigned)((n + 1023) / 1024), 256, 0, st>>>(s, n);  // data
  int ncu = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess) ncu = prop.multiProcessorCount;
  auto launch = [&]() {
    switch (variant) {
      case 0: k_copy1<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(s, d, n); break;
      case 1: k_copyU<4, false><<<(unsigned)((n + 1023) / 1024), 256, 0, st>>>(s, d, n); break;
      case 2: k_copyU<4, true><<<(unsigned)((n + 1023) / 1024), 256, 0, st>>>(s, d, n); break;
      case 3: k_copyU<8, false><<<(unsigned)((n + 2047) / 2048), 256, 0, st>>>(s, d, n); break;
      case 4: k_copy_stride<<<ncu * 8, 256, 0, st>>>(s, d, n); break;
      case 5: k_read<4><<<(unsigned)((n + 1023) / 1024), 256, 0, st>>>(s, sink, n); break;
      case 6: k_fill<4><<<(unsigned)((n + 1023) / 1024), 256, 0, st>>>(d, n); break;
    }
  };
  hipEvent_t a, b;
  PCHECK(hipEventCreate(&a));
  PCHECK(hipEventCreate(&b));
  launch();  // warm-up
  PCHECK(hipGetLastError());
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    PCHECK(hipEventRecord(a, st));
    launch();
    PCHECK(hipEventRecord(b, st));
    PCHECK(hipEventSynchronize(b));
    float ms = 0.f;
    PCHECK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  const double moved = (variant == 5 || variant == 6) ? (double)n * 16 : 2.0 * n * 16;
  *gbs = moved / (best * 1e-3) / 1e9;
  if (us_per_call) *us_per_call = best * 1e3;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipStreamDestroy(st);
  (void)hipFree(s);
  (void)hipFree(d);
  (void)hipFree(sink);
  return 0;
}
