// Device building blocks of the in-LDS FFTs shared by the transform kernels:
// the radix plan, the small DFTs, one Stockham stage and a whole line, and
// the twiddle tables.
#pragma once

#include <hip/hip_runtime.h>

#include "hst_cx.h"
#include "hst_internal.h"

namespace hst {

// ------------------------------------------------------------------ transform buffers

// Layout of one (component, row) plane of the transform buffers Hv (the
// velocities after the z inverse transform) and Hp (the products after the x
// forward transform): nzd z points x NXp x modes, tiled by HT modes (4 in
// FP32, 2 in FP64), each tile z-major ([ix / HT][zp][ix % HT]).  A block of
// a z transform, which takes HT modes, then reads (k_zfft) or writes (k_zift)
// one contiguous range; a block of the x pass, which takes four neighbouring
// z points, writes each tile's four points as one 128-byte run (its lines
// staged in LDS) and reads its lines in HT-mode pieces that the block's other
// lines share.
template <typename R>
__host__ __device__ constexpr int htile() { return sizeof(R) == 4 ? 4 : 2; }
template <typename R>
__device__ __forceinline__ size_t hoff(int NZD, int NXp, int zp, int ix) {
  constexpr int HT = htile<R>();
  (void)NXp;
  return (size_t)(ix / HT) * ((size_t)NZD * HT) + (size_t)zp * HT + ix % HT;
}

// ------------------------------------------------------------------ radix plan

__host__ __device__ constexpr int pow2_part(int N) { return N % 3 == 0 ? N / 3 : N; }
__host__ __device__ constexpr int n_radix8(int N) {
  int m = pow2_part(N), c = 0;
  while (m % 8 == 0) { m /= 8; c++; }
  return c;
}
__host__ __device__ constexpr int radix_rem(int N) {
  int m = pow2_part(N);
  while (m % 8 == 0) m /= 8;
  return m;  // 1, 2 or 4
}
__host__ __device__ constexpr int n_stages(int N) { return (N % 3 == 0 ? 1 : 0) + n_radix8(N) + (radix_rem(N) > 1 ? 1 : 0); }
__host__ __device__ constexpr int radix_at(int N, int s) {
  return (N % 3 == 0) ? (s == 0 ? 3 : (s - 1 < n_radix8(N) ? 8 : radix_rem(N))) : (s < n_radix8(N) ? 8 : radix_rem(N));
}
// threads per line
__host__ __device__ constexpr int fft_threads(int N) { return N >= 16 ? (N / 8 > 192 ? 192 : N / 8) : 1; }

// ------------------------------------------------------------------ small DFTs

template <typename R, int SIGN>
__device__ __forceinline__ void dft2(Cx<R> *v) {
  Cx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft3(Cx<R> *v) {
  const R h = (R)0.86602540378443864676;  // sin(2 pi/3)
  Cx<R> t1 = v[1] + v[2];
  Cx<R> t2 = v[0] - t1 * (R)0.5;
  Cx<R> d = (v[1] - v[2]) * (h * (R)SIGN);
  v[0] = v[0] + t1;
  v[1] = t2 + mul_i(d);
  v[2] = t2 - mul_i(d);
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft4(Cx<R> *v) {
  Cx<R> a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = (v[1] - v[3]);
  d = SIGN > 0 ? mul_i(d) : mul_mi(d);
  v[0] = a + c;
  v[2] = a - c;
  v[1] = b + d;
  v[3] = b - d;
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft8(Cx<R> *v) {
  const R r2 = (R)0.70710678118654752440;
  Cx<R> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  dft4<R, SIGN>(e);
  dft4<R, SIGN>(o);
  // o[k] *= exp(SIGN 2 pi i k / 8)
  Cx<R> o1 = o[1], o3 = o[3];
  if (SIGN > 0) {
    o[1] = Cx<R>{(o1.x - o1.y) * r2, (o1.x + o1.y) * r2};
    o[2] = mul_i(o[2]);
    o[3] = Cx<R>{(-o3.x - o3.y) * r2, (o3.x - o3.y) * r2};
  } else {
    o[1] = Cx<R>{(o1.x + o1.y) * r2, (o1.y - o1.x) * r2};
    o[2] = mul_mi(o[2]);
    o[3] = Cx<R>{(-o3.x + o3.y) * r2, (-o3.x - o3.y) * r2};
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[k] = e[k] + o[k];
    v[k + 4] = e[k] - o[k];
  }
}

template <typename R, int SIGN, int RAD>
__device__ __forceinline__ void dft(Cx<R> *v) {
  if constexpr (RAD == 2) dft2<R, SIGN>(v);
  else if constexpr (RAD == 3) dft3<R, SIGN>(v);
  else if constexpr (RAD == 4) dft4<R, SIGN>(v);
  else dft8<R, SIGN>(v);
}

// ------------------------------------------------------------------ in-LDS FFT

// One Stockham stage (decimation in frequency): sequences of the current
// length len = N/s interleaved with stride s; butterfly (q, n1) reads
// x[q + s (n1 + L n2)], n2 < r, and writes its DFT times w_len^(n1 k2) to
// y[q + s k2 + s r n1].  In place: every thread holds its butterflies in
// registers across the barrier.
// The last stage (S RAD = N, so n1 = 0) has no twiddles; with EMIT it hands
// each result to emit(k, value) (k the output index) instead of storing it.
template <typename R, int N, int T, int SIGN, int RAD, int S, bool EMIT = false, class F = int>
__device__ __forceinline__ void fft_stage(Cx<R> *line, int t, const Cx<R> *__restrict__ tw, F emit = 0) {
  constexpr int len = N / S, L = len / RAD, B = N / RAD, K = (B + T - 1) / T;
  constexpr bool LAST = (S * RAD == N);
  Cx<R> v[K][RAD];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int b = t + T * k;
    if (B % T == 0 || b < B) {
      const int q = b % S, n1 = b / S;
#pragma unroll
      for (int n2 = 0; n2 < RAD; n2++) v[k][n2] = line[q + S * (n1 + L * n2)];
    }
  }
  if constexpr (!EMIT) __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int b = t + T * k;
    if (B % T == 0 || b < B) {
      const int q = b % S, n1 = b / S;
      dft<R, SIGN, RAD>(v[k]);
      if constexpr (EMIT) {
#pragma unroll
        for (int k2 = 0; k2 < RAD; k2++) emit(q + S * k2, v[k][k2]);
      } else {
        line[q + S * RAD * n1] = v[k][0];
#pragma unroll
        for (int k2 = 1; k2 < RAD; k2++) {
          if constexpr (LAST) {
            line[q + S * k2] = v[k][k2];
          } else {
            Cx<R> w = tw[n1 * k2 * S];
            if (SIGN > 0) w.y = -w.y;
            line[q + S * k2 + S * RAD * n1] = v[k][k2] * w;
          }
        }
      }
    }
  }
  if constexpr (!EMIT) __syncthreads();
}

template <typename R, int N, int T, int SIGN, int ST = 0, int S = 1>
__device__ __forceinline__ void fft_line(Cx<R> *line, int t, const Cx<R> *__restrict__ tw) {
  if constexpr (ST < n_stages(N)) {
    constexpr int RAD = radix_at(N, ST);
    fft_stage<R, N, T, SIGN, RAD, S>(line, t, tw);
    fft_line<R, N, T, SIGN, ST + 1, S * RAD>(line, t, tw);
  }
}

// fft_line whose last stage hands its results to emit(k, value) instead of
// writing them back to the line (which it leaves as the stage before left it)
template <typename R, int N, int T, int SIGN, class F, int ST = 0, int S = 1>
__device__ __forceinline__ void fft_line_emit(Cx<R> *line, int t, const Cx<R> *__restrict__ tw, F emit) {
  constexpr int RAD = radix_at(N, ST);
  if constexpr (ST + 1 < n_stages(N)) {
    fft_stage<R, N, T, SIGN, RAD, S>(line, t, tw);
    fft_line_emit<R, N, T, SIGN, F, ST + 1, S * RAD>(line, t, tw, emit);
  } else {
    fft_stage<R, N, T, SIGN, RAD, S, true>(line, t, tw, emit);
  }
}

// ------------------------------------------------------------------ tables

template <typename R>
struct Tables {
  Cx<R> *twz;   // exp(-2 pi i k / nzd), k < nzd
  Cx<R> *twx;   // exp(-2 pi i k / M), k < M = nxd
  Cx<R> *twh;   // exp(-i pi k / M), k < M
};

struct TableSet {
  int mixed;
  Tables<double> d;
  Tables<float> f;
};

}  // namespace hst
