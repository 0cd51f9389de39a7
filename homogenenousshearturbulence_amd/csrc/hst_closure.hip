// The learned subgrid closure (no counterpart in the reference):
//
//   k_filter     spectral box in x, z and a y top-hat: fine spectra -> coarse
//   k_czinv      z inverse transform of coarse fields, d/dx and d/dz fused
//                into the load (FP64)
//   k_cxc2r      x inverse real transform to physical values (FP64)
//   k_samples    (9 gradients, 6 stresses) per coarse grid point
//   k_mlp_train  forward + backward of the 9-64-64-6 tanh network for 128
//                samples per wavefront on the matrix cores (bfloat16 operands,
//                FP32 accumulation), per-block partial gradients
//   k_adam       ordered reduction of the partial gradients and the Adam step
//   k_mlp_infer  the network on samples or on the physical gradient fields
//   k_cxr2c      x forward real transform of the modelled stresses (FP64),
//                then the solver's own z forward transform and ghost rows
//   k_add_stress P += tau-hat on the live state's products
//
// Matrix-core layout (v_mfma_f32_16x16x32_bf16, one wavefront): lane l holds
// A[l%16][8(l/16)+v] and B[8(l/16)+v][l%16], v < 8, and D[4(l/16)+v][l%16],
// v < 4.  Samples run along the 16 columns; a layer's D fragments of two
// consecutive 16-row tiles are the B operand of the next layer once the K
// slots are renamed (kfeat): slot v of lane group q is row 4q+v of the first
// tile for v < 4 and of the second for v >= 4, and the weights are staged in
// LDS in the same order.  Weight gradients contract over samples, so the
// activations and errors are also staged as [feature][sample] in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "hst_cx.h"
#include "hst_fft_dev.h"
#include "hst_internal.h"
#include "../../include/hst_closure.h"

namespace hst {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

constexpr int NP = HST_MLP_NPARAM;
constexpr int NPP = NP + 2;  // row of the partial-gradient buffer: the parameters, the loss
constexpr int OW1 = 0, OB1 = 64 * 9, OW2 = OB1 + 64, OB2 = OW2 + 64 * 64, OW3 = OB2 + 64, OB3 = OW3 + 6 * 64;
static_assert(OB3 + 6 == NP, "parameter layout");
constexpr int SPB = 128;        // samples per training block (one wavefront)
constexpr int SP = 40;          // LDS row stride (bf16) of the [feature][32 samples] staging
constexpr int MAX_PART = 2048;  // partial-gradient rows: at most SPB * MAX_PART samples per launch

// ================================================================== transforms

struct FilterArgs {
  const void *src[9];
  double2 *dst[9];
};

// One lane per coarse column; coarse row k = sum_t w[8k+t] * fine row j0[k]+t.
template <typename RS>
__global__ void __launch_bounds__(256)
    k_filter(FilterArgs a, int ncol_f, int NZm_f, int nz_f, Dims Dc, const int *__restrict__ j0,
             const double *__restrict__ w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Dc.ncol) return;
  const int ixc = c / Dc.NZm, izc = c % Dc.NZm - Dc.nz;
  const size_t fcol = (size_t)ixc * NZm_f + izc + nz_f;
  const Cx<RS> *src = (const Cx<RS> *)a.src[blockIdx.y];
  double2 *dst = a.dst[blockIdx.y];
  for (int k = 0; k < Dc.ny; k++) {
    double sx = 0.0, sy = 0.0;
    const int r0 = j0[k] + 2;
    for (int t = 0; t < 8; t++) {
      const double wt = w[k * 8 + t];
      if (wt == 0.0) continue;
      const Cx<RS> v = src[(size_t)(r0 + t) * ncol_f + fcol];
      sx += wt * (double)v.x;
      sy += wt * (double)v.y;
    }
    dst[(size_t)(k + 2) * Dc.ncol + c] = make_double2(sx, sy);
  }
}

struct ZSrc {
  const double2 *p[18];
  int op[18];  // 0 the field, 1 d/dx, 2 d/dz
};

constexpr int ZG = 8;  // z lines (x modes) per block
constexpr int XG = 8;  // x lines (z points) per block

// z inverse transform of the interior rows of nf coarse fields into
// Hc[f][row][zp][ix] (ix < NXp), the padding and the derivative factor fused
// into the load.
template <int NZD>
__global__ void __launch_bounds__(ZG * fft_threads(NZD))
    k_czinv(ZSrc s, Cx<double> *__restrict__ Hc, Dims D, const Cx<double> *__restrict__ tw) {
  constexpr int T = fft_threads(NZD), LS = NZD + 1;
  __shared__ Cx<double> lds[ZG * LS];
  __shared__ Cx<double> twl[NZD];
  for (int k = threadIdx.x; k < NZD; k += blockDim.x) twl[k] = tw[k];
  const int f = blockIdx.z, rowc = blockIdx.y, ix0 = blockIdx.x * ZG, nz = D.nz;
  const double2 *src = s.p[f] + (size_t)(rowc + 2) * D.ncol;
  const int op = s.op[f];
  for (int e = threadIdx.x; e < ZG * NZD; e += blockDim.x) {
    const int j = e / NZD, k = e % NZD, ix = ix0 + j;
    const int iz = k <= nz ? k : (k >= NZD - nz ? k - NZD : INT_MAX);
    Cx<double> v{0.0, 0.0};
    if (ix < D.NXm && iz != INT_MAX) {
      const double2 a = src[(size_t)ix * D.NZm + iz + nz];
      v = Cx<double>{a.x, a.y};
      if (op) {
        const double kk = op == 1 ? D.alfa0 * ix : D.beta0 * iz;
        v = Cx<double>{-kk * a.y, kk * a.x};
      }
    }
    lds[j * LS + k] = v;
  }
  __syncthreads();
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  fft_line<double, NZD, T, +1>(lds + g * LS, t, twl);
  Cx<double> *dst = Hc + ((size_t)f * D.ny + rowc) * NZD * D.NXp;
  for (int e = threadIdx.x; e < ZG * NZD; e += blockDim.x) {
    const int j = e % ZG, zp = e / ZG;
    if (ix0 + j < D.NXm) dst[(size_t)zp * D.NXp + ix0 + j] = lds[j * LS + zp];
  }
}

// x inverse real transform of one (field, row, z point) line per T threads:
// phys[f][row][zp][x], x < 2M.  The half-length complex form of the
// reference's c2r, the imaginary part of the kx = 0 bin ignored.
template <int M>
__global__ void __launch_bounds__(XG * fft_threads(M))
    k_cxc2r(const Cx<double> *__restrict__ Hc, double *__restrict__ phys, Dims D,
            const Cx<double> *__restrict__ twx, const Cx<double> *__restrict__ twh) {
  constexpr int T = fft_threads(M), LS = M + 1;
  __shared__ Cx<double> lin[XG * LS], wk[XG * LS];
  const int f = blockIdx.z, rowc = blockIdx.y, g = threadIdx.x / T, t = threadIdx.x % T;
  const int zp = blockIdx.x * XG + g, nx = D.nx;
  const Cx<double> *src = Hc + (((size_t)f * D.ny + rowc) * D.nzd + zp) * D.NXp;
  Cx<double> *wl = wk + g * LS, *vl = lin + g * LS;
  for (int k = t; k <= nx; k += T) wl[k] = src[k];
  __syncthreads();
  for (int k = t; k < M; k += T) {
    Cx<double> a = (k <= nx) ? wl[k] : Cx<double>{0.0, 0.0};
    if (k == 0) a.y = 0.0;
    const int j = M - k;
    const Cx<double> b = (k > 0 && j <= nx) ? conj(wl[j]) : Cx<double>{0.0, 0.0};
    const Cx<double> e = conj(twh[k]);
    vl[k] = (a + b) + mul_i(e * (a - b));
  }
  __syncthreads();
  fft_line<double, M, T, +1>(vl, t, twx);
  double *out = phys + (((size_t)f * D.ny + rowc) * D.nzd + zp) * (2 * M);
  for (int m = t; m < M; m += T) {
    out[2 * m] = vl[m].x;
    out[2 * m + 1] = vl[m].y;
  }
}

// x forward real transform (times factor) of six physical fields into the
// product buffer in the tiled layout the z forward transform reads (hoff).
template <int M>
__global__ void __launch_bounds__(XG * fft_threads(M))
    k_cxr2c(const double *__restrict__ phys, Cx<double> *__restrict__ Hp, Dims D,
            const Cx<double> *__restrict__ twx, const Cx<double> *__restrict__ twh) {
  constexpr int T = fft_threads(M), LS = M + 1;
  __shared__ Cx<double> lin[XG * LS];
  const int p = blockIdx.z, rowc = blockIdx.y, g = threadIdx.x / T, t = threadIdx.x % T;
  const int zp = blockIdx.x * XG + g, nx = D.nx;
  const double *in = phys + (((size_t)p * D.ny + rowc) * D.nzd + zp) * (2 * M);
  Cx<double> *vl = lin + g * LS;
  for (int m = t; m < M; m += T) vl[m] = Cx<double>{in[2 * m] * D.factor, in[2 * m + 1] * D.factor};
  __syncthreads();
  fft_line<double, M, T, -1>(vl, t, twx);
  Cx<double> *dst = Hp + ((size_t)p * D.ny + rowc) * D.nzd * D.NXp;
  for (int k = t; k <= nx; k += T) {
    const Cx<double> Z = vl[k], Zc = conj(vl[(M - k) % M]);
    const Cx<double> e = twh[k];
    const Cx<double> s = (Z + Zc) * 0.5, d = (Z - Zc) * 0.5;
    dst[hoff<double>(D.nzd, D.NXp, zp, k)] = s - mul_i(e * d);
  }
}

// (gradients, stresses) of every point: phys fields 0..8 the gradients
// (i-major), 9..11 u v w, 12..17 the filtered products.
__global__ void __launch_bounds__(256)
    k_samples(const double *__restrict__ phys, size_t fs, int64_t npts, float *__restrict__ X,
              float *__restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
#pragma unroll
  for (int k = 0; k < 9; k++) X[i * 9 + k] = (float)phys[k * fs + i];
  const double u = phys[9 * fs + i], v = phys[10 * fs + i], w = phys[11 * fs + i];
  const double uu[6] = {u * u, v * v, w * w, u * v, v * w, u * w};
#pragma unroll
  for (int p = 0; p < 6; p++) Y[i * 6 + p] = (float)(phys[(12 + p) * fs + i] - uu[p]);
}

// P[p] += T[p] over all rows of the six products
template <typename R>
__global__ void __launch_bounds__(256) k_add_stress(Cx<R> *__restrict__ P, const double2 *__restrict__ T, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 t = T[i];
  Cx<R> p = P[i];
  p.x = (R)((double)p.x + t.x);
  p.y = (R)((double)p.y + t.y);
  P[i] = p;
}

// ================================================================== the network

struct Norm {
  float mx[9], rx[9], my[6], ry[6], sy[6];
};

__device__ __forceinline__ f4 mfma(const bf8 &a, const bf8 &b, const f4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// feature of K slot v of lane group q in K block kb, when the B operand is
// formed from the D fragments of output tiles 2kb and 2kb+1
__device__ __forceinline__ int kfeat(int kb, int q, int v) { return 32 * kb + (v < 4 ? 4 * q + v : 12 + 4 * q + v); }

struct WLds {
  bf8 A1[4][64];     // W1[16m+i][8q+v], K padded 9 -> 32
  bf8 A2[4][2][64];  // W2[16m+i][kfeat]
  bf8 A3[2][64];     // W3[i][kfeat], rows 6..15 zero
  bf8 B3[4][64];     // W3[4q+v][16m+i] (v < 4): W3^T against the output errors
  bf8 B2[4][2][64];  // W2[kfeat][16m+i]: W2^T
  float b1[64], b2[64], b3[16];
};

__device__ void stage_weights(const float *__restrict__ par, WLds &W, int l) {
  const int i = l & 15, q = l >> 4;
  for (int m = 0; m < 4; m++) {
    bf8 a;
    for (int v = 0; v < 8; v++) {
      const int k = 8 * q + v;
      a[v] = (__bf16)(k < 9 ? par[OW1 + (16 * m + i) * 9 + k] : 0.0f);
    }
    W.A1[m][l] = a;
    bf8 b;
    for (int v = 0; v < 8; v++) {
      const int o = 4 * q + v;
      b[v] = (__bf16)((v < 4 && o < 6) ? par[OW3 + o * 64 + 16 * m + i] : 0.0f);
    }
    W.B3[m][l] = b;
    for (int kb = 0; kb < 2; kb++) {
      bf8 x, y;
      for (int v = 0; v < 8; v++) {
        const int f = kfeat(kb, q, v);
        x[v] = (__bf16)par[OW2 + (16 * m + i) * 64 + f];
        y[v] = (__bf16)par[OW2 + f * 64 + 16 * m + i];
      }
      W.A2[m][kb][l] = x;
      W.B2[m][kb][l] = y;
    }
  }
  for (int kb = 0; kb < 2; kb++) {
    bf8 a;
    for (int v = 0; v < 8; v++) a[v] = (__bf16)(i < 6 ? par[OW3 + i * 64 + kfeat(kb, q, v)] : 0.0f);
    W.A3[kb][l] = a;
  }
  W.b1[l] = par[OB1 + l];
  W.b2[l] = par[OB2 + l];
  if (l < 16) W.b3[l] = l < 6 ? par[OB3 + l] : 0.0f;
}

// B operands of the next layer from the D fragments of four 16-row tiles
__device__ __forceinline__ void to_operands(const f4 h[4], bf8 B[2]) {
#pragma unroll
  for (int kb = 0; kb < 2; kb++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      B[kb][v] = (__bf16)h[2 * kb][v];
      B[kb][4 + v] = (__bf16)h[2 * kb + 1][v];
    }
}

// Forward pass of one 16-sample tile: x the input operand; h1, h2 the hidden
// activations (FP32 D fragments), y the outputs (rows 4q+v, valid below 6).
__device__ __forceinline__ void forward(const WLds &W, int l, const bf8 &x, f4 h1[4], f4 h2[4], f4 &y) {
  const int q = l >> 4;
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const f4 z = mfma(W.A1[m][l], x, zero);
#pragma unroll
    for (int v = 0; v < 4; v++) h1[m][v] = tanhf(z[v] + W.b1[16 * m + 4 * q + v]);
  }
  bf8 B[2];
  to_operands(h1, B);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    f4 z = mfma(W.A2[m][0][l], B[0], zero);
    z = mfma(W.A2[m][1][l], B[1], z);
#pragma unroll
    for (int v = 0; v < 4; v++) h2[m][v] = tanhf(z[v] + W.b2[16 * m + 4 * q + v]);
  }
  to_operands(h2, B);
  y = mfma(W.A3[0][l], B[0], zero);
  y = mfma(W.A3[1][l], B[1], y);
#pragma unroll
  for (int v = 0; v < 4; v++) y[v] += W.b3[4 * q + v];
}

// the input operand of the sample of lane l: x[0..7] in group 0, x[8] in group 1
template <class LoadX>
__device__ __forceinline__ bf8 input_operand(int q, bool valid, const Norm &nm, LoadX ld) {
  bf8 x;
#pragma unroll
  for (int v = 0; v < 8; v++) {
    const int k = 8 * q + v;
    float xv = 0.0f;
    if (valid && k < 9) xv = (ld(k) - nm.mx[k]) * nm.rx[k];
    x[v] = (__bf16)xv;
  }
  return x;
}

struct TrainArgs {
  const float *X, *Y;
  int64_t n;              // samples of this launch
  int64_t base;           // first position
  int64_t mul, add, mod;  // sample index = (mul * (base + b) + add) % mod
  float scale;            // 2 / (samples of the loss * 6)
};

// GRAD: forward, backward and the per-block partial gradients (part[blk][NPP],
// the last two entries the sum of squared errors); otherwise the sum of
// squared errors alone into lossp[blk].
template <bool GRAD>
__global__ void __launch_bounds__(64)
    k_mlp_train(const float *__restrict__ par, Norm nm, TrainArgs ta, float *__restrict__ part,
                double *__restrict__ lossp) {
  __shared__ WLds W;
  __shared__ __attribute__((aligned(16))) __bf16 Xs[16][SP], H1s[64][SP], H2s[64][SP], DYs[16][SP], DZ2s[64][SP],
      DZ1s[64][SP];
  const int l = threadIdx.x, j = l & 15, q = l >> 4;
  stage_weights(par, W, l);
  for (int r = 9; r < 16; r++)
    if (l < SP) Xs[r][l] = (__bf16)0.0f;
  __syncthreads();
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  f4 gW2[4][4], gW1[4], gW3[4];
  float gb1[4][4], gb2[4][4], gb3[4];
  for (int a = 0; a < 4; a++) {
    gW1[a] = zero;
    gW3[a] = zero;
    for (int b = 0; b < 4; b++) { gW2[a][b] = zero; gb1[a][b] = 0.f; gb2[a][b] = 0.f; }
    gb3[a] = 0.f;
  }
  float lsum = 0.0f;
  for (int kb = 0; kb < SPB / 32; kb++) {
#pragma unroll 1
    for (int jt = 0; jt < 2; jt++) {
      const int64_t b = (int64_t)blockIdx.x * SPB + kb * 32 + jt * 16 + j;
      const bool valid = b < ta.n;
      int64_t idx = 0;
      if (valid) idx = (ta.mul * (ta.base + b) + ta.add) % ta.mod;
      const bf8 x = input_operand(q, valid, nm, [&](int k) { return ta.X[idx * 9 + k]; });
      f4 h1[4], h2[4], y;
      forward(W, l, x, h1, h2, y);
      f4 dy = zero;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int o = 4 * q + v;
        if (valid && o < 6) {
          const float yh = (ta.Y[idx * 6 + o] - nm.my[o]) * nm.ry[o];
          const float e = y[v] - yh;
          lsum += e * e;
          dy[v] = ta.scale * e;
        }
      }
      if constexpr (!GRAD) continue;
      bf8 Bd;
#pragma unroll
      for (int v = 0; v < 4; v++) { Bd[v] = (__bf16)dy[v]; Bd[4 + v] = (__bf16)0.0f; }
      f4 dz2[4], dz1[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const f4 a = mfma(W.B3[m][l], Bd, zero);
#pragma unroll
        for (int v = 0; v < 4; v++) dz2[m][v] = a[v] * (1.0f - h2[m][v] * h2[m][v]);
      }
      bf8 Bz[2];
      to_operands(dz2, Bz);
#pragma unroll
      for (int m = 0; m < 4; m++) {
        f4 a = mfma(W.B2[m][0][l], Bz[0], zero);
        a = mfma(W.B2[m][1][l], Bz[1], a);
#pragma unroll
        for (int v = 0; v < 4; v++) dz1[m][v] = a[v] * (1.0f - h1[m][v] * h1[m][v]);
      }
      // stage [feature][sample] for the weight gradients; bias gradients in FP32
      const int col = jt * 16 + j;
#pragma unroll
      for (int v = 0; v < 8; v++)
        if (8 * q + v < 9) Xs[8 * q + v][col] = x[v];
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = 16 * m + 4 * q + v;
          H1s[r][col] = (__bf16)h1[m][v];
          H2s[r][col] = (__bf16)h2[m][v];
          DZ2s[r][col] = (__bf16)dz2[m][v];
          DZ1s[r][col] = (__bf16)dz1[m][v];
          gb1[m][v] += dz1[m][v];
          gb2[m][v] += dz2[m][v];
        }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        DYs[4 * q + v][col] = Bd[v];
        gb3[v] += dy[v];
      }
    }
    if constexpr (GRAD) {
      __syncthreads();
      // K = the 32 samples of this block step: A rows from one array, B columns from another
      auto rd = [&](const __bf16 *row) { return *(const bf8 *)(row + 8 * q); };
#pragma unroll
      for (int mn = 0; mn < 4; mn++) {
        const bf8 dz2r = rd(DZ2s[16 * mn + j]);
#pragma unroll
        for (int mk = 0; mk < 4; mk++) gW2[mn][mk] = mfma(dz2r, rd(H1s[16 * mk + j]), gW2[mn][mk]);
        gW1[mn] = mfma(rd(DZ1s[16 * mn + j]), rd(Xs[j]), gW1[mn]);
        gW3[mn] = mfma(rd(DYs[j]), rd(H2s[16 * mn + j]), gW3[mn]);
      }
      __syncthreads();
    }
  }
  // sum of squared errors over the wavefront
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if constexpr (!GRAD) {
    if (l == 0) lossp[blockIdx.x] = (double)lsum;
    return;
  }
  float *pr = part + (size_t)blockIdx.x * NPP;
#pragma unroll
  for (int mn = 0; mn < 4; mn++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * mn + 4 * q + v;
#pragma unroll
      for (int mk = 0; mk < 4; mk++) pr[OW2 + r * 64 + 16 * mk + j] = gW2[mn][mk][v];
      if (j < 9) pr[OW1 + r * 9 + j] = gW1[mn][v];
      if (4 * q + v < 6) pr[OW3 + (4 * q + v) * 64 + 16 * mn + j] = gW3[mn][v];
    }
  // bias gradients: sum over the 16 sample columns
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int v = 0; v < 4; v++)
      for (int o = 1; o < 16; o <<= 1) {
        gb1[m][v] += __shfl_xor(gb1[m][v], o);
        gb2[m][v] += __shfl_xor(gb2[m][v], o);
      }
#pragma unroll
  for (int v = 0; v < 4; v++)
    for (int o = 1; o < 16; o <<= 1) gb3[v] += __shfl_xor(gb3[v], o);
  if (j == 0) {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        pr[OB1 + 16 * m + 4 * q + v] = gb1[m][v];
        pr[OB2 + 16 * m + 4 * q + v] = gb2[m][v];
      }
#pragma unroll
    for (int v = 0; v < 4; v++)
      if (4 * q + v < 6) pr[OB3 + 4 * q + v] = gb3[v];
  }
  if (l == 0) pr[NP] = lsum;
}

// Ordered sum of nblk partial rows per parameter, then (update) the Adam step
// or (!update) the gradient into grad; slot NP: the summed squared error.
// A block holds 64 parameters (one per lane of a wavefront, so every row read
// is 256 contiguous bytes) and AR row groups (one per wavefront); the group
// sums are added in a fixed order, so the result does not depend on timing.
constexpr int AR = 8;
__global__ void __launch_bounds__(64 * AR)
    k_adam(float *__restrict__ par, float *__restrict__ m1, float *__restrict__ m2, const float *__restrict__ part,
           int nblk, int update, float lr, float c1, float c2, float *__restrict__ grad, double *__restrict__ sse) {
  __shared__ double red[AR][64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + l;
  double g = 0.0;
  if (p <= NP) {
#pragma unroll 8
    for (int b = w; b < nblk; b += AR) g += (double)part[(size_t)b * NPP + p];
  }
  red[w][l] = g;
  __syncthreads();
  if (w != 0 || p > NP) return;
  g = 0.0;
  for (int k = 0; k < AR; k++) g += red[k][l];
  if (p == NP) {
    *sse = g;
    return;
  }
  if (!update) {
    grad[p] = (float)g;
    return;
  }
  const float gf = (float)g;
  const float a = 0.9f * m1[p] + 0.1f * gf, s = 0.999f * m2[p] + 0.001f * gf * gf;
  m1[p] = a;
  m2[p] = s;
  par[p] -= lr * (a * c1) / (sqrtf(s * c2) + 1e-8f);
}

struct InferIO {
  const float *X;      // mode 0: x[n][9]
  float *Y;            // mode 0: y[n][6]
  const double *gin;   // mode 1: 9 gradient fields of stride fs
  double *tout;        // mode 1: 6 stress fields of stride fs
  size_t fs;
  int64_t n;
  int mode;
};

constexpr int INFER_TILES = 32;  // 16-sample tiles per block: the weight staging is paid once per 512 samples

__global__ void __launch_bounds__(64) k_mlp_infer(const float *__restrict__ par, Norm nm, InferIO io) {
  __shared__ WLds W;
  const int l = threadIdx.x, j = l & 15, q = l >> 4;
  stage_weights(par, W, l);
  __syncthreads();
  for (int jt = 0; jt < INFER_TILES; jt++) {
    const int64_t s = ((int64_t)blockIdx.x * INFER_TILES + jt) * 16 + j;
    const bool valid = s < io.n;
    bf8 x;
    if (io.mode == 0) x = input_operand(q, valid, nm, [&](int k) { return io.X[s * 9 + k]; });
    else x = input_operand(q, valid, nm, [&](int k) { return (float)io.gin[(size_t)k * io.fs + s]; });
    f4 h1[4], h2[4], y;
    forward(W, l, x, h1, h2, y);
    if (!valid) continue;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int o = 4 * q + v;
      if (o >= 6) continue;
      const float t = y[v] * nm.sy[o] + nm.my[o];
      if (io.mode == 0) io.Y[s * 6 + o] = t;
      else io.tout[(size_t)o * io.fs + s] = (double)t;
    }
  }
}

// sums (pass 0) or sums of squared deviations (pass 1) of the 15 columns, per block
__global__ void __launch_bounds__(256)
    k_moments(const float *__restrict__ X, const float *__restrict__ Y, int64_t n, const double *__restrict__ mean,
              int pass, double *__restrict__ part) {
  __shared__ double red[15][4];
  double acc[15];
  for (int k = 0; k < 15; k++) acc[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    for (int k = 0; k < 15; k++) {
      double v = k < 9 ? (double)X[i * 9 + k] : (double)Y[i * 6 + k - 9];
      if (pass) {
        v -= mean[k];
        v *= v;
      }
      acc[k] += v;
    }
  }
  for (int k = 0; k < 15; k++) {
    double v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 15) {
    const int k = threadIdx.x;
    part[(size_t)blockIdx.x * 15 + k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
  }
}

// ================================================================== host side

template <int NZD>
static void run_czinv(const Dims &D, const ZSrc &s, int nf, Cx<double> *Hc, const Cx<double> *tw, hipStream_t st) {
  dim3 grid((D.NXm + ZG - 1) / ZG, D.ny, nf);
  hipLaunchKernelGGL(k_czinv<NZD>, grid, dim3(ZG * fft_threads(NZD)), 0, st, s, Hc, D, tw);
}
template <int M>
static void run_cxc2r(const Dims &D, int nf, const Cx<double> *Hc, double *phys, const Cx<double> *twx,
                      const Cx<double> *twh, hipStream_t st) {
  dim3 grid(D.nzd / XG, D.ny, nf);
  hipLaunchKernelGGL(k_cxc2r<M>, grid, dim3(XG * fft_threads(M)), 0, st, Hc, phys, D, twx, twh);
}
template <int M>
static void run_cxr2c(const Dims &D, const double *phys, Cx<double> *Hp, const Cx<double> *twx,
                      const Cx<double> *twh, hipStream_t st) {
  dim3 grid(D.nzd / XG, D.ny, 6);
  hipLaunchKernelGGL(k_cxr2c<M>, grid, dim3(XG * fft_threads(M)), 0, st, phys, Hp, D, twx, twh);
}

#define CLOSURE_Z_SIZES(X) X(48) X(64) X(96) X(128) X(192)
#define CLOSURE_X_SIZES(X) X(32) X(48) X(64) X(96) X(128)

static bool closure_size_ok(const Dims &D) {
  bool z = false, x = false;
#define ZC(n) z = z || D.nzd == n;
#define XC(n) x = x || D.nxd == n;
  CLOSURE_Z_SIZES(ZC)
  CLOSURE_X_SIZES(XC)
#undef ZC
#undef XC
  return z && x && D.nzd % XG == 0;
}

}  // namespace hst

using namespace hst;

struct hst_closure {
  hst_deck deck;
  hst_grid_info g;
  Dims D;
  std::vector<double> y, dyl, der;
  D0Factor f0;
  D0Const fc;
  double *d_der = nullptr, *d_dyl = nullptr, *d_f0 = nullptr;
  Cx<double> *twz = nullptr, *twx = nullptr, *twh = nullptr;
  void *tables = nullptr;  // FP64 tables of the solver's z forward transform
  double2 *fields = nullptr;  // 15 coarse fields: U[3], Pbar[6], Tau[6]
  double2 *Gy = nullptr;      // 3 y derivatives
  double2 *XR = nullptr;      // D0 solve workspace
  Cx<double> *Hc = nullptr;   // z-inverse output, 18 fields
  double *phys = nullptr;     // 18 physical fields of the interior rows
  double2 *stage = nullptr;   // layout staging, one field
  int *d_j0 = nullptr;
  double *d_w = nullptr;
  int fine_ny = 0;
  int64_t cap = 0, count = 0;
  float *X = nullptr, *Yd = nullptr;
  float *par = nullptr, *m1 = nullptr, *m2 = nullptr, *part = nullptr, *gradd = nullptr;
  double *sse = nullptr, *lossp = nullptr, *mom = nullptr, *meand = nullptr;
  int64_t adam_t = 0;
  double norm[HST_CLOSURE_NORM];
  bool have_norm = false, have_params = false, attached = false;
  hipStream_t own = nullptr;
  size_t nf() const { return (size_t)D.NYg * D.ncol; }
  double2 *field(int k) { return fields + (size_t)k * nf(); }
  size_t fs() const { return (size_t)D.ny * D.nzd * 2 * D.nxd; }
};

namespace {

int cfail(const std::string &m, int rc = 1) {
  set_error("closure: " + m);
  return rc ? rc : 1;
}

#define CHK(call)                                                                               \
  do {                                                                                          \
    hipError_t _e = (call);                                                                     \
    if (_e != hipSuccess) return cfail(std::string(#call) + ": " + hipGetErrorString(_e), (int)_e); \
  } while (0)

hst_closure *g_attached = nullptr;

// the stream: the live state's when there is one (its data are read), else the context's
hipStream_t stream_of(hst_closure *c) {
  StateView v;
  if (state_view(v)) return v.stream;
  return c->own;
}

Norm make_norm(const double *n) {
  Norm nm;
  for (int k = 0; k < 9; k++) {
    nm.mx[k] = (float)n[k];
    nm.rx[k] = 1.0f / (float)n[9 + k];
  }
  for (int k = 0; k < 6; k++) {
    nm.my[k] = (float)n[18 + k];
    nm.ry[k] = 1.0f / (float)n[24 + k];
    nm.sy[k] = (float)n[24 + k];
  }
  return nm;
}

int filter_weights(int nyf, int nyc, std::vector<int> &j0, std::vector<double> &w) {
  if (nyc < 4 || nyf < nyc) return 1;
  j0.assign(nyc, 0);
  w.assign((size_t)nyc * 8, 0.0);
  const double r = 0.5 * (double)nyf / (double)nyc;
  auto cdf = [](double x) {  // integral of the unit hat function up to x
    if (x <= -1.0) return 0.0;
    if (x <= 0.0) return 0.5 * (x + 1.0) * (x + 1.0);
    if (x <= 1.0) return 1.0 - 0.5 * (1.0 - x) * (1.0 - x);
    return 1.0;
  };
  for (int k = 0; k < nyc; k++) {
    const double cen = (double)k * nyf / nyc, a = cen - r, b = cen + r;
    const int lo = (int)std::floor(a), hi = (int)std::ceil(b);
    if (hi - lo + 1 > 8) return 2;
    j0[k] = lo;
    for (int t = 0; t <= hi - lo; t++) w[(size_t)k * 8 + t] = (cdf(b - (lo + t)) - cdf(a - (lo + t))) / (b - a);
  }
  return 0;
}

// gradients (and, with sample, u and the filtered products) of the context's
// fields to physical values: phys fields 0..8 (and 9..17)
int to_physical(hst_closure *c, const double2 *const U[3], bool sample, double sx, double sz, hipStream_t st) {
  const Dims &D = c->D;
  for (int i = 0; i < 3; i++) {
    int rc = launch_d0solve(D, 1, 0, U[i], c->Gy + (size_t)i * c->nf(), nullptr, c->XR, c->d_f0, c->fc, c->d_der,
                            c->d_dyl, sx, sz, nullptr, nullptr, st);
    if (rc) return cfail("y derivative", rc);
  }
  ZSrc zs;
  int nf = 0;
  for (int i = 0; i < 3; i++) {
    zs.p[nf] = U[i]; zs.op[nf++] = 1;
    zs.p[nf] = c->Gy + (size_t)i * c->nf(); zs.op[nf++] = 0;
    zs.p[nf] = U[i]; zs.op[nf++] = 2;
  }
  if (sample) {
    for (int i = 0; i < 3; i++) { zs.p[nf] = U[i]; zs.op[nf++] = 0; }
    for (int p = 0; p < 6; p++) { zs.p[nf] = c->field(HST_CLOSURE_PBAR + p); zs.op[nf++] = 0; }
  }
  switch (D.nzd) {
#define ZC(n) case n: run_czinv<n>(D, zs, nf, c->Hc, c->twz, st); break;
    CLOSURE_Z_SIZES(ZC)
#undef ZC
    default: return cfail("z size");
  }
  switch (D.nxd) {
#define XC(n) case n: run_cxc2r<n>(D, nf, c->Hc, c->phys, c->twx, c->twh, st); break;
    CLOSURE_X_SIZES(XC)
#undef XC
    default: return cfail("x size");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cfail(std::string("transform launch: ") + hipGetErrorString(e), (int)e);
  return 0;
}

// modelled stress spectra of U into the Tau fields (ghosts at sx, sz)
int stress(hst_closure *c, const double2 *const U[3], double sx, double sz, hipStream_t st) {
  if (!c->have_norm || !c->have_params) return cfail("the network has no parameters or no standardisation");
  const Dims &D = c->D;
  if (int rc = to_physical(c, U, false, sx, sz, st)) return rc;
  InferIO io;
  io.X = nullptr; io.Y = nullptr;
  io.gin = c->phys;
  io.tout = c->phys + 9 * c->fs();
  io.fs = c->fs();
  io.n = (int64_t)c->fs();
  io.mode = 1;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((io.n + 16 * INFER_TILES - 1) / (16 * INFER_TILES))), dim3(64), 0,
                     st, c->par, make_norm(c->norm), io);
  Cx<double> *Hp = c->Hc;  // the z-inverse output is free again
  switch (D.nxd) {
#define XC(n) case n: run_cxr2c<n>(D, c->phys + 9 * c->fs(), Hp, c->twx, c->twh, st); break;
    CLOSURE_X_SIZES(XC)
#undef XC
    default: return cfail("x size");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cfail(std::string("stress launch: ") + hipGetErrorString(e), (int)e);
  int rc = launch_zfft(D, 0, 6, Hp, c->field(HST_CLOSURE_TAU), c->nf(), 2, D.ny, c->tables, st);
  if (rc) return cfail("z forward transform", rc);
  double2 *T[6];
  for (int p = 0; p < 6; p++) T[p] = c->field(HST_CLOSURE_TAU + p);
  rc = launch_ghosts(D, 6, T, sx, sz, st);
  if (rc) return cfail("ghost rows", rc);
  return 0;
}

// the live state has the grid and box of the context (the sizes its kernels assume)
bool same_grid(const StateView &v, const hst_closure *c) {
  const Dims &F = v.D, &D = c->D;
  return F.nx == D.nx && F.ny == D.ny && F.nz == D.nz && v.deck.alfa0 == c->deck.alfa0 &&
         v.deck.beta0 == c->deck.beta0 && v.deck.ly == c->deck.ly && v.deck.S == c->deck.S &&
         v.deck.ystretch == c->deck.ystretch;
}

// the timestep hook: stresses of the live state's velocity added to its products
int hook(void *ctx, double time) {
  hst_closure *c = (hst_closure *)ctx;
  StateView v;
  if (!state_view(v)) return cfail("no live state");
  // the state may have been freed and set up again on another grid since the attach
  if (!same_grid(v, c)) return cfail("the live state's grid is no longer the attached context's");
  double sx, sz;
  deck_shifts(c->deck, c->g.lx, c->g.lz, time, sx, sz);
  const double2 *U[3] = {v.F[HST_V1], v.F[HST_V2], v.F[HST_V3]};
  return profiled("closure", 0.0, [&]() -> int {
    if (int rc = stress(c, U, sx, sz, v.stream)) return rc;
    const size_t n = 6 * c->nf();
    if (v.mixed)
      hipLaunchKernelGGL(k_add_stress<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, v.stream,
                         (Cx<float> *)v.P, (const double2 *)c->field(HST_CLOSURE_TAU), n);
    else
      hipLaunchKernelGGL(k_add_stress<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, v.stream,
                         (Cx<double> *)v.P, (const double2 *)c->field(HST_CLOSURE_TAU), n);
    return (int)hipGetLastError();
  });
}

void release(hst_closure *c) {
  if (c->own) (void)hipStreamSynchronize(c->own);
  void *ptrs[] = {c->d_der, c->d_dyl, c->d_f0, c->twz, c->twx, c->twh, c->fields, c->Gy, c->XR, c->Hc, c->phys,
                  c->stage, c->d_j0, c->d_w, c->X, c->Yd, c->par, c->m1, c->m2, c->part, c->gradd, c->sse,
                  c->lossp, c->mom, c->meand};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->tables) free_fft_tables(c->tables);
  if (c->own) (void)hipStreamDestroy(c->own);
}

int fill_tw(Cx<double> **dst, int n, double den) {
  std::vector<Cx<double>> h(n);
  for (int k = 0; k < n; k++) {
    const double a = -2.0 * M_PI * (double)k / den;
    h[k] = Cx<double>{std::cos(a), std::sin(a)};
  }
  if (hipMalloc(dst, sizeof(Cx<double>) * n) != hipSuccess) return 1;
  return hipMemcpy(*dst, h.data(), sizeof(Cx<double>) * n, hipMemcpyHostToDevice) != hipSuccess;
}

}  // namespace

extern "C" {

int hst_closure_filter_weights(int ny_fine, int ny_coarse, int32_t *first_row, double *weights) {
  std::vector<int> j0;
  std::vector<double> w;
  if (filter_weights(ny_fine, ny_coarse, j0, w)) return cfail("filter: need 4 <= ny_coarse <= ny_fine, ratio below 6");
  for (int k = 0; k < ny_coarse; k++) first_row[k] = j0[k];
  std::memcpy(weights, w.data(), sizeof(double) * w.size());
  return 0;
}

int hst_closure_create(const hst_deck *coarse, int64_t capacity, hst_closure **out) {
  if (!coarse || !out) return cfail("NULL argument");
  *out = nullptr;
  std::string err = check_deck(*coarse);
  if (!err.empty()) return cfail(err);
  if (capacity < 0) return cfail("negative capacity");
  hst_closure *c = new hst_closure();
  c->deck = *coarse;
  grid_setup(c->deck, c->g, c->y, c->dyl, c->der);
  c->D = make_dims(c->deck, c->g);
  const Dims &D = c->D;
  if (!closure_size_ok(D)) {
    delete c;
    return cfail("coarse grid transform sizes nxd = " + std::to_string(D.nxd) + ", nzd = " + std::to_string(D.nzd) +
                 " are not built for the closure (nxd 32..128, nzd 48..192)");
  }
  d0_factor(c->der, D.ny, c->f0);
  c->fc = d0_const(c->f0);
  auto bad = [&](const char *what, hipError_t e) {
    release(c);
    delete c;
    return cfail(std::string(what) + ": " + hipGetErrorString(e), (int)e);
  };
#define MAL(p, bytes)                                    \
  do {                                                   \
    hipError_t _e = hipMalloc(&(p), (bytes));            \
    if (_e != hipSuccess) return bad("allocation", _e); \
  } while (0)
  hipError_t e = hipStreamCreate(&c->own);
  if (e != hipSuccess) return bad("stream", e);
  std::vector<double> derd((size_t)D.ny * 20);
  for (int iy = 0; iy < D.ny; iy++)
    for (int n = 0; n < 4; n++)
      for (int j = 0; j < 5; j++) derd[(size_t)iy * 20 + n * 5 + j] = c->der[(size_t)iy + (size_t)D.ny * (n + 4 * j)];
  MAL(c->d_der, derd.size() * sizeof(double));
  MAL(c->d_dyl, c->dyl.size() * sizeof(double));
  MAL(c->d_f0, c->f0.rows.size() * sizeof(double));
  (void)hipMemcpy(c->d_der, derd.data(), derd.size() * sizeof(double), hipMemcpyHostToDevice);
  (void)hipMemcpy(c->d_dyl, c->dyl.data(), c->dyl.size() * sizeof(double), hipMemcpyHostToDevice);
  (void)hipMemcpy(c->d_f0, c->f0.rows.data(), c->f0.rows.size() * sizeof(double), hipMemcpyHostToDevice);
  if (fill_tw(&c->twz, D.nzd, D.nzd) || fill_tw(&c->twx, D.nxd, D.nxd) || fill_tw(&c->twh, D.nxd, 2.0 * D.nxd))
    return bad("twiddles", hipErrorOutOfMemory);
  if (make_fft_tables(D, 0, &c->tables)) return bad("tables", hipErrorOutOfMemory);
  MAL(c->fields, 15 * c->nf() * sizeof(double2));
  MAL(c->Gy, 3 * c->nf() * sizeof(double2));
  MAL(c->stage, c->nf() * sizeof(double2));
  MAL(c->XR, (size_t)(D.ny - 2) * D.ncol * sizeof(double2));
  MAL(c->Hc, 18 * (size_t)D.ny * D.nzd * D.NXp * sizeof(Cx<double>));
  MAL(c->phys, 18 * c->fs() * sizeof(double));
  (void)hipMemset(c->fields, 0, 15 * c->nf() * sizeof(double2));
  (void)hipMemset(c->Gy, 0, 3 * c->nf() * sizeof(double2));
  c->cap = capacity;
  if (capacity > 0) {
    MAL(c->X, (size_t)capacity * 9 * sizeof(float));
    MAL(c->Yd, (size_t)capacity * 6 * sizeof(float));
  }
  MAL(c->par, NP * sizeof(float));
  MAL(c->m1, NP * sizeof(float));
  MAL(c->m2, NP * sizeof(float));
  MAL(c->gradd, NP * sizeof(float));
  MAL(c->part, (size_t)MAX_PART * NPP * sizeof(float));
  MAL(c->sse, sizeof(double) * 1);
  MAL(c->lossp, sizeof(double) * 4096);
  MAL(c->mom, sizeof(double) * 15 * 1024);
  MAL(c->meand, sizeof(double) * 15);
  (void)hipMemset(c->m1, 0, NP * sizeof(float));
  (void)hipMemset(c->m2, 0, NP * sizeof(float));
#undef MAL
  *out = c;
  return 0;
}

int hst_closure_free(hst_closure *c) {
  if (!c) return 0;
  if (g_attached == c) {
    set_stress_hook(nullptr, nullptr);
    g_attached = nullptr;
  }
  release(c);
  delete c;
  return 0;
}

int hst_closure_filter(hst_closure *c, double time) {
  if (!c) return cfail("NULL context");
  StateView v;
  if (!state_view(v)) return cfail("filter: no live state (hst_gpu_init)");
  const Dims &F = v.D, &D = c->D;
  if (F.nx < D.nx || F.nz < D.nz || F.ny < D.ny || v.deck.alfa0 != c->deck.alfa0 || v.deck.beta0 != c->deck.beta0 ||
      v.deck.ly != c->deck.ly || v.deck.ystretch != 0.0 || c->deck.ystretch != 0.0)
    return cfail("filter: the state's grid must be finer, uniform in y, with the same alfa0, beta0, ly");
  if (c->fine_ny != F.ny) {
    std::vector<int> j0;
    std::vector<double> w;
    if (filter_weights(F.ny, D.ny, j0, w)) return cfail("filter: y ratio not supported (at most 6)");
    for (int k = 0; k < D.ny; k++)
      for (int t = 0; t < 8; t++)
        if (w[(size_t)k * 8 + t] != 0.0 && (j0[k] + t < -2 || j0[k] + t > F.ny + 1))
          return cfail("filter: a weight falls outside the fine rows");
    if (c->d_j0) (void)hipFree(c->d_j0);
    if (c->d_w) (void)hipFree(c->d_w);
    c->d_j0 = nullptr;
    c->d_w = nullptr;
    CHK(hipMalloc(&c->d_j0, sizeof(int) * D.ny));
    CHK(hipMalloc(&c->d_w, sizeof(double) * 8 * D.ny));
    CHK(hipMemcpy(c->d_j0, j0.data(), sizeof(int) * D.ny, hipMemcpyHostToDevice));
    CHK(hipMemcpy(c->d_w, w.data(), sizeof(double) * 8 * D.ny, hipMemcpyHostToDevice));
    c->fine_ny = F.ny;
  }
  hipStream_t st = v.stream;
  const size_t cs = v.mixed ? sizeof(float2) : sizeof(double2);
  FilterArgs fv, fp;
  for (int i = 0; i < 3; i++) {
    fv.src[i] = v.F[HST_V1 + i];
    fv.dst[i] = c->field(HST_CLOSURE_U + i);
  }
  for (int p = 0; p < 6; p++) {
    fp.src[p] = (const char *)v.P + (size_t)p * F.NYg * F.ncol * cs;
    fp.dst[p] = c->field(HST_CLOSURE_PBAR + p);
  }
  const unsigned gx = (unsigned)((D.ncol + 255) / 256);
  hipLaunchKernelGGL(k_filter<double>, dim3(gx, 3), dim3(256), 0, st, fv, F.ncol, F.NZm, F.nz, D, c->d_j0, c->d_w);
  if (v.mixed)
    hipLaunchKernelGGL(k_filter<float>, dim3(gx, 6), dim3(256), 0, st, fp, F.ncol, F.NZm, F.nz, D, c->d_j0, c->d_w);
  else
    hipLaunchKernelGGL(k_filter<double>, dim3(gx, 6), dim3(256), 0, st, fp, F.ncol, F.NZm, F.nz, D, c->d_j0, c->d_w);
  CHK(hipGetLastError());
  double sx, sz;
  deck_shifts(c->deck, c->g.lx, c->g.lz, time, sx, sz);
  double2 *f[8];
  for (int i = 0; i < 8; i++) f[i] = c->field(i);
  int rc = launch_ghosts(D, 8, f, sx, sz, st);
  if (!rc) {
    double2 *last[1] = {c->field(8)};
    rc = launch_ghosts(D, 1, last, sx, sz, st);
  }
  if (rc) return cfail("ghost rows", rc);
  CHK(hipStreamSynchronize(st));
  return 0;
}

int hst_closure_get_field(hst_closure *c, int which, double *host) {
  if (!c || !host) return cfail("NULL argument");
  if (which < 0 || which > 14) return cfail("field index out of range (0..14)");
  hipStream_t st = stream_of(c);
  int rc = launch_to_ref_layout(c->D, c->field(which), c->stage, st);
  if (rc) return cfail("layout", rc);
  CHK(hipMemcpyAsync(host, c->stage, c->nf() * sizeof(double2), hipMemcpyDeviceToHost, st));
  CHK(hipStreamSynchronize(st));
  return 0;
}

int hst_closure_set_field(hst_closure *c, int which, const double *host) {
  if (!c || !host) return cfail("NULL argument");
  if (which < 0 || which > 14) return cfail("field index out of range (0..14)");
  hipStream_t st = stream_of(c);
  CHK(hipMemcpyAsync(c->stage, host, c->nf() * sizeof(double2), hipMemcpyHostToDevice, st));
  int rc = launch_to_device_layout(c->D, c->stage, c->field(which), st);
  if (rc) return cfail("layout", rc);
  CHK(hipStreamSynchronize(st));
  return 0;
}

int hst_closure_sample(hst_closure *c, double time, int64_t *count) {
  if (!c) return cfail("NULL context");
  const int64_t npts = (int64_t)c->fs();
  if (c->count + npts > c->cap)
    return cfail("data set full: " + std::to_string(c->count) + " + " + std::to_string(npts) + " > capacity " +
                 std::to_string(c->cap));
  hipStream_t st = stream_of(c);
  double sx, sz;
  deck_shifts(c->deck, c->g.lx, c->g.lz, time, sx, sz);
  const double2 *U[3] = {c->field(0), c->field(1), c->field(2)};
  if (int rc = to_physical(c, U, true, sx, sz, st)) return rc;
  hipLaunchKernelGGL(k_samples, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, c->phys, c->fs(), npts,
                     c->X + (size_t)c->count * 9, c->Yd + (size_t)c->count * 6);
  CHK(hipGetLastError());
  CHK(hipStreamSynchronize(st));
  c->count += npts;
  if (count) *count = c->count;
  return 0;
}

int hst_closure_set_dataset(hst_closure *c, int64_t n, const float *x, const float *y) {
  if (!c || !x || !y) return cfail("NULL argument");
  if (n < 0 || n > c->cap) return cfail("data set size beyond capacity");
  CHK(hipMemcpy(c->X, x, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice));
  CHK(hipMemcpy(c->Yd, y, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice));
  c->count = n;
  return 0;
}

int hst_closure_get_dataset(hst_closure *c, int64_t first, int64_t n, float *x, float *y) {
  if (!c) return cfail("NULL context");
  if (first < 0 || n < 0 || first + n > c->count) return cfail("range beyond the data set");
  if (x) CHK(hipMemcpy(x, c->X + (size_t)first * 9, (size_t)n * 9 * sizeof(float), hipMemcpyDeviceToHost));
  if (y) CHK(hipMemcpy(y, c->Yd + (size_t)first * 6, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int hst_closure_count(hst_closure *c, int64_t *count) {
  if (!c || !count) return cfail("NULL argument");
  *count = c->count;
  return 0;
}

int hst_closure_normalise(hst_closure *c, double *norm) {
  if (!c) return cfail("NULL context");
  if (c->count < 2) return cfail("normalise: fewer than two samples");
  hipStream_t st = c->own;
  const int nblk = 1024;
  double s[15], q[15];
  std::vector<double> h((size_t)nblk * 15);
  for (int pass = 0; pass < 2; pass++) {
    hipLaunchKernelGGL(k_moments, dim3(nblk), dim3(256), 0, st, c->X, c->Yd, c->count, c->meand, pass, c->mom);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(h.data(), c->mom, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    CHK(hipStreamSynchronize(st));
    double *acc = pass ? q : s;
    for (int k = 0; k < 15; k++) {
      acc[k] = 0.0;
      for (int b = 0; b < nblk; b++) acc[k] += h[(size_t)b * 15 + k];
    }
    if (pass == 0) {
      double mean[15];
      for (int k = 0; k < 15; k++) mean[k] = s[k] / (double)c->count;
      CHK(hipMemcpy(c->meand, mean, sizeof(mean), hipMemcpyHostToDevice));
    }
  }
  for (int k = 0; k < 15; k++) {
    const double mean = s[k] / (double)c->count, sd = std::sqrt(q[k] / (double)c->count);
    const int mi = k < 9 ? k : 18 + (k - 9), si = k < 9 ? 9 + k : 24 + (k - 9);
    c->norm[mi] = mean;
    c->norm[si] = sd > 0.0 ? sd : 1.0;
  }
  c->have_norm = true;
  if (norm) std::memcpy(norm, c->norm, sizeof(c->norm));
  return 0;
}

int hst_closure_set_norm(hst_closure *c, const double *norm) {
  if (!c || !norm) return cfail("NULL argument");
  for (int k = 0; k < 9; k++)
    if (!(norm[9 + k] > 0.0)) return cfail("standard deviations must be positive");
  for (int k = 0; k < 6; k++)
    if (!(norm[24 + k] > 0.0)) return cfail("standard deviations must be positive");
  std::memcpy(c->norm, norm, sizeof(c->norm));
  c->have_norm = true;
  return 0;
}

int hst_closure_get_norm(hst_closure *c, double *norm) {
  if (!c || !norm) return cfail("NULL argument");
  if (!c->have_norm) return cfail("no standardisation yet");
  std::memcpy(norm, c->norm, sizeof(c->norm));
  return 0;
}

int hst_closure_init_params(hst_closure *c, uint64_t seed) {
  if (!c) return cfail("NULL context");
  std::mt19937_64 gen(seed);
  std::vector<float> p(NP, 0.0f);
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < fan_out * fan_in; i++) p[off + i] = (float)u(gen);
  };
  fill(OW1, 64, 9);
  fill(OW2, 64, 64);
  fill(OW3, 6, 64);
  return hst_closure_set_params(c, p.data());
}

int hst_closure_set_params(hst_closure *c, const float *params) {
  if (!c || !params) return cfail("NULL argument");
  CHK(hipMemcpy(c->par, params, NP * sizeof(float), hipMemcpyHostToDevice));
  CHK(hipMemset(c->m1, 0, NP * sizeof(float)));
  CHK(hipMemset(c->m2, 0, NP * sizeof(float)));
  c->adam_t = 0;
  c->have_params = true;
  return 0;
}

int hst_closure_get_params(hst_closure *c, float *params) {
  if (!c || !params) return cfail("NULL argument");
  CHK(hipMemcpy(params, c->par, NP * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int hst_closure_grad(hst_closure *c, int64_t first, int64_t n, float *grad, double *loss) {
  if (!c) return cfail("NULL context");
  if (!c->have_norm || !c->have_params) return cfail("grad: parameters and standardisation needed");
  if (first < 0 || n < 1 || first + n > c->count) return cfail("grad: range beyond the data set");
  if (n > (int64_t)SPB * MAX_PART) return cfail("grad: at most " + std::to_string(SPB * MAX_PART) + " samples");
  hipStream_t st = c->own;
  TrainArgs ta{c->X, c->Yd, n, first, 1, 0, INT64_MAX, (float)(2.0 / ((double)n * 6.0))};
  const int nblk = (int)((n + SPB - 1) / SPB);
  hipLaunchKernelGGL(k_mlp_train<true>, dim3(nblk), dim3(64), 0, st, c->par, make_norm(c->norm), ta, c->part,
                     c->lossp);
  hipLaunchKernelGGL(k_adam, dim3((NP + 1 + 63) / 64), dim3(64 * AR), 0, st, c->par, c->m1, c->m2, c->part, nblk, 0,
                     0.0f, 0.0f, 0.0f, c->gradd, c->sse);
  CHK(hipGetLastError());
  double sse = 0.0;
  if (grad) CHK(hipMemcpyAsync(grad, c->gradd, NP * sizeof(float), hipMemcpyDeviceToHost, st));
  CHK(hipMemcpyAsync(&sse, c->sse, sizeof(double), hipMemcpyDeviceToHost, st));
  CHK(hipStreamSynchronize(st));
  if (loss) *loss = sse / ((double)n * 6.0);
  return 0;
}

// loss over [first, first + n) without gradients
static int eval_loss(hst_closure *c, int64_t first, int64_t n, double *loss) {
  hipStream_t st = c->own;
  double total = 0.0;
  const int64_t chunk = (int64_t)SPB * 4096;
  std::vector<double> h(4096);
  for (int64_t b0 = 0; b0 < n; b0 += chunk) {
    const int64_t m = std::min(chunk, n - b0);
    TrainArgs ta{c->X, c->Yd, m, first + b0, 1, 0, INT64_MAX, 0.0f};
    const int nblk = (int)((m + SPB - 1) / SPB);
    hipLaunchKernelGGL(k_mlp_train<false>, dim3(nblk), dim3(64), 0, st, c->par, make_norm(c->norm), ta, c->part,
                       c->lossp);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(h.data(), c->lossp, nblk * sizeof(double), hipMemcpyDeviceToHost, st));
    CHK(hipStreamSynchronize(st));
    for (int b = 0; b < nblk; b++) total += h[b];
  }
  *loss = total / ((double)n * 6.0);
  return 0;
}

int hst_closure_train(hst_closure *c, int epochs, int batch, double lr, uint64_t seed, int64_t n_train,
                      double *loss, double *val_loss) {
  if (!c) return cfail("NULL context");
  if (!c->have_norm || !c->have_params) return cfail("train: parameters and standardisation needed");
  if (n_train <= 0 || n_train > c->count) n_train = c->count;
  if (batch < SPB || batch % SPB || batch > n_train || batch > SPB * MAX_PART)
    return cfail("train: batch must be a multiple of 128, at most the training samples and " +
                 std::to_string(SPB * MAX_PART));
  if (epochs < 1) return cfail("train: epochs < 1");
  hipStream_t st = c->own;
  const Norm nm = make_norm(c->norm);
  const int64_t nb = n_train / batch;
  const int nblk = batch / SPB;
  std::mt19937_64 gen(seed);
  double *dsse = nullptr;
  CHK(hipMalloc(&dsse, sizeof(double) * nb));
  std::vector<double> h(nb);
  for (int e = 0; e < epochs; e++) {
    // this epoch's order: sample = (mul * position + add) mod n_train, mul coprime to n_train
    std::uniform_int_distribution<int64_t> um(1, std::max<int64_t>(1, n_train - 1)), ua(0, n_train - 1);
    int64_t mul = um(gen);
    while (std::gcd(mul, n_train) != 1) mul = um(gen);
    const int64_t add = ua(gen);
    for (int64_t b = 0; b < nb; b++) {
      c->adam_t++;
      const double t = (double)c->adam_t;
      const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, t))), c2 = (float)(1.0 / (1.0 - std::pow(0.999, t)));
      TrainArgs ta{c->X, c->Yd, batch, b * batch, mul, add, n_train, (float)(2.0 / ((double)batch * 6.0))};
      hipLaunchKernelGGL(k_mlp_train<true>, dim3(nblk), dim3(64), 0, st, c->par, nm, ta, c->part, c->lossp);
      hipLaunchKernelGGL(k_adam, dim3((NP + 1 + 63) / 64), dim3(64 * AR), 0, st, c->par, c->m1, c->m2, c->part, nblk,
                         1, (float)lr, c1, c2, c->gradd, dsse + b);
    }
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpyAsync(h.data(), dsse, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) {
      (void)hipFree(dsse);
      return cfail(std::string("train: ") + hipGetErrorString(err), (int)err);
    }
    double s = 0.0;
    for (int64_t b = 0; b < nb; b++) s += h[b] / ((double)batch * 6.0);
    if (loss) loss[e] = s / (double)nb;
    if (val_loss) {
      val_loss[e] = 0.0;
      if (n_train < c->count) {
        int rc = eval_loss(c, n_train, c->count - n_train, &val_loss[e]);
        if (rc) {
          (void)hipFree(dsse);
          return rc;
        }
      }
    }
  }
  (void)hipFree(dsse);
  return 0;
}

int hst_closure_predict(hst_closure *c, int64_t n, const float *x, float *y) {
  if (!c || !x || !y) return cfail("NULL argument");
  if (!c->have_norm || !c->have_params) return cfail("predict: parameters and standardisation needed");
  if (n < 1) return 0;
  hipStream_t st = c->own;
  float *dx = nullptr, *dy = nullptr;
  CHK(hipMalloc(&dx, (size_t)n * 9 * sizeof(float)));
  if (hipMalloc(&dy, (size_t)n * 6 * sizeof(float)) != hipSuccess) {
    (void)hipFree(dx);
    return cfail("predict: allocation");
  }
  InferIO io{dx, dy, nullptr, nullptr, 0, n, 0};
  hipError_t e = hipMemcpyAsync(dx, x, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 16 * INFER_TILES - 1) / (16 * INFER_TILES))), dim3(64), 0,
                       st, c->par, make_norm(c->norm), io);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(y, dy, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (e != hipSuccess) return cfail(std::string("predict: ") + hipGetErrorString(e), (int)e);
  return 0;
}

int hst_closure_stress(hst_closure *c, double time) {
  if (!c) return cfail("NULL context");
  hipStream_t st = stream_of(c);
  double sx, sz;
  deck_shifts(c->deck, c->g.lx, c->g.lz, time, sx, sz);
  const double2 *U[3] = {c->field(0), c->field(1), c->field(2)};
  if (int rc = stress(c, U, sx, sz, st)) return rc;
  CHK(hipStreamSynchronize(st));
  return 0;
}

int hst_closure_attach(hst_closure *c, int on) {
  if (!c) return cfail("NULL context");
  if (!on) {
    if (g_attached == c) {
      set_stress_hook(nullptr, nullptr);
      g_attached = nullptr;
    }
    return 0;
  }
  StateView v;
  if (!state_view(v)) return cfail("attach: no live state");
  if (!same_grid(v, c)) return cfail("attach: the live state's grid is not the context's");
  if (!c->have_norm || !c->have_params) return cfail("attach: the network has no parameters or no standardisation");
  set_stress_hook(hook, c);
  g_attached = c;
  return 0;
}

}  // extern "C"
