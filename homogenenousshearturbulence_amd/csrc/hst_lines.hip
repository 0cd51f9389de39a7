// Batched FP64 transforms of whole lines, for the per-field entry points of
// the Fortran host's hst_fft (FFT, IFT, RFT, HFT of src/hst_fft.f90:181-237)
// and the product pass of its hst_transforms (build_products,
// src/hst_transforms.f90:97-121).  These work on the reference's own buffer
// shapes, outside the fused time step:
//
//   complex lines of N points, contiguous (VVdz: N = nzd, in place)
//   real x lines of 2M points stored as M + 1 complex (VVdx, VVdp: M = nxd,
//     the padded in-place layout: a line of M + 1 complex modes and a line of
//     2(M + 1) reals are the same bytes)
//
// One block holds G lines in LDS and runs the library's Stockham stages on
// them (fft_line).  The real transforms of length 2M run as complex
// transforms of length M, with the pre- and post-processing of the x pass.
#include <hip/hip_runtime.h>

#include "hst_cx.h"
#include "hst_fft_dev.h"
#include "hst_internal.h"

namespace hst {

typedef Cx<double> CD;

// lines per block: G T threads (at most 256) and G lines of LDS within 48 KB
template <int N, int PER_LINE>
struct LCfg {
  static constexpr int T = fft_threads(N);
  static constexpr int G0 = 256 / T > 0 ? 256 / T : 1;
  static constexpr int G1 = (48 * 1024) / (PER_LINE * (int)sizeof(CD));
  static constexpr int G = G0 < G1 ? (G0 > 0 ? G0 : 1) : (G1 > 0 ? G1 : 1);
  static constexpr int NT = G * T;
};

// SIGN -1: forward (exp(-i...)), +1: backward; unnormalised, in place.
template <int N, int SIGN>
__global__ void __launch_bounds__(256) k_cfft_lines(double2 *__restrict__ data, int64_t nlines,
                                                    const CD *__restrict__ tw) {
  typedef LCfg<N, N> C;
  __shared__ CD lds[C::G * N];
  const int g = threadIdx.x / C::T, t = threadIdx.x % C::T;
  const int64_t l0 = (int64_t)blockIdx.x * C::G;
  for (int e = threadIdx.x; e < C::G * N; e += blockDim.x) {
    const int64_t l = l0 + e / N;
    lds[e] = l < nlines ? ld(data + l * N + e % N) : CD{0.0, 0.0};
  }
  __syncthreads();
  fft_line<double, N, C::T, SIGN>(lds + g * N, t, tw);
  for (int e = threadIdx.x; e < C::G * N; e += blockDim.x) {
    const int64_t l = l0 + e / N;
    if (l < nlines) st(data + l * N + e % N, lds[e]);
  }
}

// DIR +1 (RFT, c2r): the modes X(0..M) of a line (imaginary parts of X(0) and
// X(M) ignored) -> its 2M real points; DIR -1 (HFT, r2c): 2M real points ->
// X(0..M).  twh[k] = exp(-i pi k / M), k = 0..M.  The two reals after the 2M
// points of a c2r line are set to zero.
template <int M, int DIR>
__global__ void __launch_bounds__(256) k_rfft_lines(double2 *__restrict__ data, int64_t nlines,
                                                    const CD *__restrict__ tw, const CD *__restrict__ twh) {
  typedef LCfg<M, 2 * M + 1> C;
  constexpr int W = M + 1;  // complex entries per stored line
  __shared__ CD X[C::G * W];
  __shared__ CD Z[C::G * M];
  const int g = threadIdx.x / C::T, t = threadIdx.x % C::T;
  const int64_t l0 = (int64_t)blockIdx.x * C::G;
  for (int e = threadIdx.x; e < C::G * W; e += blockDim.x) {
    const int64_t l = l0 + e / W;
    X[e] = l < nlines ? ld(data + l * W + e % W) : CD{0.0, 0.0};
  }
  __syncthreads();
  CD *x = X + g * W, *z = Z + g * M;
  if constexpr (DIR > 0) {
    for (int k = t; k < M; k += C::T) {
      CD a = x[k], b = conj(x[M - k]);
      if (k == 0) {
        a.y = 0.0;
        b.y = 0.0;
      }
      z[k] = (a + b) + mul_i(conj(twh[k]) * (a - b));
    }
    __syncthreads();
    fft_line<double, M, C::T, +1>(z, t, tw);
    // z[m] = x(2m) + i x(2m+1): the line's reals in order, then two zeros
    for (int e = threadIdx.x; e < C::G * W; e += blockDim.x) {
      const int64_t l = l0 + e / W;
      const int k = e % W;
      if (l < nlines) st(data + l * W + k, k < M ? Z[(e / W) * M + k] : CD{0.0, 0.0});
    }
  } else {
    for (int k = t; k < M; k += C::T) z[k] = x[k];  // the reals of a line as M complex
    __syncthreads();
    fft_line<double, M, C::T, -1>(z, t, tw);
    for (int k = t; k <= M; k += C::T) {
      const CD Zk = z[k % M], Zc = conj(z[(M - k) % M]);
      const CD sm = (Zk + Zc) * 0.5, df = (Zk - Zc) * 0.5;
      x[k] = sm - mul_i(twh[k] * df);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < C::G * W; e += blockDim.x) {
      const int64_t l = l0 + e / W;
      if (l < nlines) st(data + l * W + e % W, X[e]);
    }
  }
}

#define HST_LINE_SIZES(X) X(6) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536)

int launch_cfft_lines(int n, int sign, double2 *data, int64_t nlines, const void *tw, hipStream_t s) {
  switch (n) {
#define CASE(N)                                                                                                    \
  case N: {                                                                                                        \
    typedef LCfg<N, N> C;                                                                                          \
    const dim3 grid((unsigned)((nlines + C::G - 1) / C::G));                                                       \
    if (sign > 0) hipLaunchKernelGGL((k_cfft_lines<N, +1>), grid, dim3(C::NT), 0, s, data, nlines, (const CD *)tw); \
    else hipLaunchKernelGGL((k_cfft_lines<N, -1>), grid, dim3(C::NT), 0, s, data, nlines, (const CD *)tw);          \
    return hipGetLastError();                                                                                      \
  }
    HST_LINE_SIZES(CASE)
#undef CASE
  }
  return hipErrorInvalidValue;
}

int launch_rfft_lines(int m, int dir, double2 *data, int64_t nlines, const void *tw, const void *twh,
                      hipStream_t s) {
  switch (m) {
#define CASE(N)                                                                                                  \
  case N: {                                                                                                      \
    typedef LCfg<N, 2 * N + 1> C;                                                                                \
    const dim3 grid((unsigned)((nlines + C::G - 1) / C::G));                                                     \
    if (dir > 0)                                                                                                 \
      hipLaunchKernelGGL((k_rfft_lines<N, +1>), grid, dim3(C::NT), 0, s, data, nlines, (const CD *)tw,           \
                         (const CD *)twh);                                                                       \
    else                                                                                                         \
      hipLaunchKernelGGL((k_rfft_lines<N, -1>), grid, dim3(C::NT), 0, s, data, nlines, (const CD *)tw,           \
                         (const CD *)twh);                                                                       \
    return hipGetLastError();                                                                                    \
  }
    HST_LINE_SIZES(CASE)
#undef CASE
  }
  return hipErrorInvalidValue;
}

// build_products(g) on the real x lines of three fields (stride `plane`
// doubles between the fields, lines of `lw` doubles of which the first
// `used` are points): group 1 uu, vv, ww; 2 uv, vw, uw; each times factor.
// The other doubles of a line are left as they are.
__global__ void __launch_bounds__(256) k_products(const double *__restrict__ vel, double *__restrict__ prod,
                                                  int64_t plane, int lw, int used, int g, double factor) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane || (int)(i % lw) >= used) return;
  const double u = vel[i], v = vel[plane + i], w = vel[2 * plane + i];
  if (g == 1) {
    prod[i] = u * u * factor;
    prod[plane + i] = v * v * factor;
    prod[2 * plane + i] = w * w * factor;
  } else {
    prod[i] = u * v * factor;
    prod[plane + i] = v * w * factor;
    prod[2 * plane + i] = u * w * factor;
  }
}

int launch_products(const double *vel, double *prod, int64_t plane, int lw, int used, int g, double factor,
                    hipStream_t s) {
  const dim3 grid((unsigned)((plane + 255) / 256));
  hipLaunchKernelGGL(k_products, grid, dim3(256), 0, s, vel, prod, plane, lw, used, g, factor);
  return hipGetLastError();
}

}  // namespace hst
