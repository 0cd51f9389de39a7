// The HBM ceiling of this device, measured: streaming kernels with no
// arithmetic (read only, write only, copy) over buffers far larger than the
// last-level cache, each lane keeping U independent 16-byte loads in flight
// and storing with non-temporal stores.  The best of several launch shapes
// is the practical bandwidth every HBM-bound kernel of the library is
// measured against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "hst_internal.h"

namespace hst {
namespace {

constexpr int PB = 256;  // threads per block
typedef float v4f __attribute__((ext_vector_type(4)));  // a 16-byte load (the nontemporal builtins take native vectors)

// grid-stride over n float4 values, U loads in flight per lane
template <int U>
__global__ void __launch_bounds__(PB) k_read(const v4f *__restrict__ a, size_t n, float *__restrict__ sink) {
  const size_t step = (size_t)gridDim.x * PB * U;
  float acc = 0.0f;
  for (size_t i = (size_t)blockIdx.x * PB * U + threadIdx.x; i + (size_t)(U - 1) * PB < n; i += step) {
    v4f v[U];
#pragma unroll
    for (int k = 0; k < U; k++) v[k] = __builtin_nontemporal_load(a + i + (size_t)k * PB);
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k].x + v[k].y + v[k].z + v[k].w;
  }
  // one value per block keeps the loads live
  __shared__ float red[PB / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sink[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

template <int U>
__global__ void __launch_bounds__(PB) k_write(v4f *__restrict__ b, size_t n, float val) {
  const size_t step = (size_t)gridDim.x * PB * U;
  const v4f v = {val, val, val, val};
  for (size_t i = (size_t)blockIdx.x * PB * U + threadIdx.x; i + (size_t)(U - 1) * PB < n; i += step) {
#pragma unroll
    for (int k = 0; k < U; k++) __builtin_nontemporal_store(v, b + i + (size_t)k * PB);
  }
}

template <int U>
__global__ void __launch_bounds__(PB) k_copy(const v4f *__restrict__ a, v4f *__restrict__ b, size_t n) {
  const size_t step = (size_t)gridDim.x * PB * U;
  for (size_t i = (size_t)blockIdx.x * PB * U + threadIdx.x; i + (size_t)(U - 1) * PB < n; i += step) {
    v4f v[U];
#pragma unroll
    for (int k = 0; k < U; k++) v[k] = __builtin_nontemporal_load(a + i + (size_t)k * PB);
#pragma unroll
    for (int k = 0; k < U; k++) __builtin_nontemporal_store(v[k], b + i + (size_t)k * PB);
  }
}

}  // namespace
}  // namespace hst

using namespace hst;

extern "C" int hst_gpu_bandwidth_probe(double gib, int reps, double out[6]) {
  if (!out) {
    set_error("bandwidth_probe: out is NULL");
    return 1;
  }
  if (!(gib >= 0.25 && gib <= 64.0) || reps < 1 || reps > 100) {
    set_error("bandwidth_probe: gib must be in [0.25, 64] and reps in [1, 100]");
    return 1;
  }
  constexpr int U = 4;
  // n a multiple of PB * U * 1024: every grid below covers it exactly
  const size_t unit = (size_t)PB * U * 1024;
  const size_t n = std::max<size_t>(1, (size_t)(gib * (double)(1ull << 30) / 16.0) / unit) * unit;
  v4f *a = nullptr, *b = nullptr;
  float *sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, n * sizeof(v4f));
  if (e == hipSuccess) e = hipMalloc(&b, n * sizeof(v4f));
  const int grids[4] = {1024, 2048, 4096, (int)(n / ((size_t)PB * U))};
  int maxgrid = 0;
  for (int g : grids) maxgrid = std::max(maxgrid, g);
  if (e == hipSuccess) e = hipMalloc(&sink, (size_t)maxgrid * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * sizeof(v4f), st);
  double best[3] = {0.0, 0.0, 0.0};
  int best_grid[3] = {0, 0, 0};
  for (int which = 0; which < 3 && e == hipSuccess; which++) {
    for (int g : grids) {
      if (e != hipSuccess) break;
      for (int r = 0; r <= reps && e == hipSuccess; r++) {  // r = 0: warm-up
        e = hipEventRecord(e0, st);
        if (which == 0) hipLaunchKernelGGL(k_read<U>, dim3(g), dim3(PB), 0, st, a, n, sink);
        else if (which == 1) hipLaunchKernelGGL(k_write<U>, dim3(g), dim3(PB), 0, st, b, n, 1.0f);
        else hipLaunchKernelGGL(k_copy<U>, dim3(g), dim3(PB), 0, st, a, b, n);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && r > 0 && ms > 0.0f) {
          const double bytes = (which == 2 ? 2.0 : 1.0) * (double)n * 16.0;
          const double gbs = bytes / (ms * 1e-3) / 1e9;
          if (gbs > best[which]) {
            best[which] = gbs;
            best_grid[which] = g;
          }
        }
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (sink) (void)hipFree(sink);
  if (b) (void)hipFree(b);
  if (a) (void)hipFree(a);
  if (e != hipSuccess) {
    set_error(std::string("bandwidth_probe: ") + hipGetErrorString(e));
    return (int)e;
  }
  for (int k = 0; k < 3; k++) {
    out[k] = best[k];
    out[3 + k] = best_grid[k];
  }
  return 0;
}
