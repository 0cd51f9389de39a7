// Kernels of the spectral side of the RK3 substep on gfx950: the right-hand
// sides, the shear shift, the cyclic pentadiagonal line solves, the recovery
// of u and w, the ghost rows and the statistics.
//
// Layout: a field is [row][col] with row = iy + 2 (iy = -2..ny+1) and
// col = ix*(2nz+1) + iz + nz, so one lane owns one Fourier column and a
// wavefront reads 64 consecutive columns of a row in one coalesced access.
// The compact-stencil weights der(iy, n, j) are the same for every lane of a
// wavefront (all lanes are at the same row), so they come through the scalar
// cache.  Every kernel with a recurrence in y runs one column per lane.
#include <hip/hip_runtime.h>

#include "hst_cx.h"
#include "hst_internal.h"

namespace hst {

typedef Cx<double> C;

enum { KD2V = 1, KETA = 2, KPOISSON = 3, KD0 = 4, KDY = 5 };

__device__ __forceinline__ void col_ix_iz(const Dims &D, int col, int &ix, int &iz) {
  ix = col / D.NZm;
  iz = col - ix * D.NZm - D.nz;
}

__device__ __forceinline__ C cexpi(double th) {
  double s, c;
  sincos(th, &s, &c);
  return C{c, s};
}

// der(iy, n, j) on the device: [iy][n][j + 2]
__device__ __forceinline__ double DER(const double *der, int iy, int n, int j) { return der[iy * 20 + n * 5 + j + 2]; }

// ------------------------------------------------------------------ right-hand sides

template <typename PR>
struct PTraits;
template <>
struct PTraits<double> {
  typedef double2 T2;
};
template <>
struct PTraits<float> {
  typedef float2 T2;
};

template <typename PR>
__device__ __forceinline__ Cx<PR> ldp(const typename PTraits<PR>::T2 *p) {
  typename PTraits<PR>::T2 v = *p;
  return Cx<PR>{v.x, v.y};
}

// buildrhs_prepare + buildrhs(g = 1, 2) (+ the default node-phase shift) of one
// column over a chunk of rows, with five-row windows of u, v, w and of the six
// product spectra in registers.  PR: the arithmetic of the product stencils.
// MODE 3: prepare and both product groups (the fused substep); 0: prepare only;
// 1, 2: add the products of group 1 or 2 to rhs, oldrhs as buildrhs(ODE, g);
// 4: prepare and the five combined products of the x pass's MODE 3
// (vv - uu, vv - ww, uv, vw, uw), the same sums regrouped.
constexpr int RHS_ROWS = 64;
template <typename PR, int MODE, bool SHIFT>
__global__ void __launch_bounds__(256)
    k_buildrhs(Dims D, RhsArgs a, const double2 *__restrict__ V1, const double2 *__restrict__ V2,
               const double2 *__restrict__ V3, const typename PTraits<PR>::T2 *__restrict__ P0, size_t pstride,
               double2 *__restrict__ R1, double2 *__restrict__ R2, typename PTraits<PR>::T2 *__restrict__ O1,
               typename PTraits<PR>::T2 *__restrict__ O2, const double *__restrict__ der,
               const double *__restrict__ y, int rch) {
  // FIVE: the five combinations; CMP (MODE 5): V1 holds eta and V3 dv/dy (the
  // time step's own form between substeps), so the eta equation's f is V1
  constexpr bool FIVE = (MODE == 4 || MODE == 5), CMP = (MODE == 5);
  constexpr bool NONLIN = MODE != 0, PREP = (MODE == 0 || MODE == 3 || FIVE);
  constexpr int Q0 = (MODE == 2 ? 3 : 0), Q1 = (MODE == 1 ? 3 : (FIVE ? 5 : 6));
  const int iy0 = blockIdx.y * rch, iy1 = min(iy0 + rch, D.ny);
  // FP32 products: the stencil rows of the block's chunk from LDS (read as
  // scalar loads, each row's waited for in turn; 2.5 % faster).  FP64 keeps
  // the scalar loads: read from LDS the uniform values take VGPRs (218 -> 254)
  // and it ran 2 % slower.
  constexpr bool LROWS = sizeof(PR) == 4;
  __shared__ double dl[LROWS ? RHS_ROWS * 20 : 1];
  if constexpr (LROWS) {
    constexpr int PER = (RHS_ROWS * 20 + 255) / 256;
    const int n = (iy1 - iy0) * 20;
    double t[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) t[u] = der[iy0 * 20 + min(u * 256 + (int)threadIdx.x, n - 1)];
#pragma unroll
    for (int u = 0; u < PER; u++)
      if (u * 256 + (int)threadIdx.x < n) dl[u * 256 + threadIdx.x] = t[u];
    __syncthreads();
  }
  auto DERL = [&](int iy, int n, int j) -> double {
    if constexpr (LROWS) return dl[(iy - iy0) * 20 + n * 5 + j + 2];
    return DER(der, iy, n, j);
  };
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const double al = D.alfa0 * ix, be = D.beta0 * iz;
  const double kk = al * al + be * be;
  const bool m00 = (ix == 0 && iz == 0);
  const double gam = a.s2now * (ix * D.alfa0) - D.S * (iz * D.beta0);
  const double gd = a.gdamp ? a.gsoft * a.gdamp[col] : 0.0;
  const bool masked = a.hard && a.gdamp && a.gdamp[col] > 0.0;
  const size_t nc = D.ncol;
  // windows: w[k] holds row iy - 2 + k
  C u[5], v[5], w[5];
  Cx<PR> p[6][5];
#pragma unroll
  for (int k = 1; k < 5; k++) {
    const size_t r = (size_t)(iy0 - 2 + k - 1 + 2) * nc + col;
    if constexpr (PREP) { u[k] = ld(V1 + r); v[k] = ld(V2 + r); if constexpr (!CMP) w[k] = ld(V3 + r); }
    if constexpr (NONLIN) {
#pragma unroll
      for (int q = Q0; q < Q1; q++) p[q][k] = ldp<PR>(P0 + q * pstride + r);
    }
  }
  const PR alp = (PR)al, bep = (PR)be, kkp = (PR)kk;
  for (int iy = iy0; iy < iy1; iy++) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if constexpr (PREP) { u[k] = u[k + 1]; v[k] = v[k + 1]; if constexpr (!CMP) w[k] = w[k + 1]; }
      if constexpr (NONLIN) {
#pragma unroll
        for (int q = Q0; q < Q1; q++) p[q][k] = p[q][k + 1];
      }
    }
    {
      const size_t r = (size_t)(iy + 2 + 2) * nc + col;
      if constexpr (PREP) { u[4] = ld(V1 + r); v[4] = ld(V2 + r); if constexpr (!CMP) w[4] = ld(V3 + r); }
      if constexpr (NONLIN) {
#pragma unroll
        for (int q = Q0; q < Q1; q++) p[q][4] = ldp<PR>(P0 + q * pstride + r);
      }
    }
    const size_t c0 = (size_t)(iy + 2) * nc + col;
    C r1, r2, o1, o2;
    if constexpr (!PREP) {
      r1 = ld(R1 + c0); r2 = ld(R2 + c0); o1 = ld(O1 + c0); o2 = ld(O2 + c0);
    } else {
    // v equation (src/hst_equations.f90:108-117)
    C unkn{0, 0}, impl{0, 0};
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const double d0 = DERL(iy, 0, j - 2), d2 = DERL(iy, 2, j - 2), d4 = DERL(iy, 3, j - 2);
      const double helm = d2 - kk * d0;
      const double biharm = d4 - 2.0 * kk * d2 + kk * kk * d0;
      unkn = unkn + helm * v[j];
      impl = impl + (D.ni * biharm) * v[j];
    }
    C o2old = ld(O2 + c0), o1old = ld(O1 + c0);
    // ODE(1) unkn / deltat + impl - ODE(3) oldrhs - ODE(2) g unkn, g the
    // column's explicit damping rate
    r2 = a.ode0 * unkn;
    r2 = C{r2.x / a.deltat, r2.y / a.deltat} + impl - a.ode2 * o2old;
    o2 = C{0.0, 0.0};
    if (gd != 0.0) {
      r2 = r2 - (a.ode1 * gd) * unkn;
      o2 = C{-gd * unkn.x, -gd * unkn.y};
    }
    // eta equation (:118-137), the (0,0) column packs (u, w)
    C unk1{0, 0}, imp1{0, 0}, expl{0, 0};
    double force = 0.0;
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const double d0 = DERL(iy, 0, j - 2), d2 = DERL(iy, 2, j - 2);
      const double helm = d2 - kk * d0;
      if (m00) {
        const C f = CMP ? u[j] : C{u[j].x, w[j].x};
        unk1 = unk1 + d0 * f;
        imp1 = imp1 + (D.ni * d2) * f;
        if (a.fy) force = force + d0 * a.fy[iy + j];  // the layer's body force on the mean w, D0-weighted
      } else {
        const C f = CMP ? u[j] : imul(be, u[j]) - imul(al, w[j]);
        unk1 = unk1 + d0 * f;
        imp1 = imp1 + (D.ni * helm) * f;
      }
      expl = expl + imul(gam * d0, v[j]);
    }
    if (m00 && a.fy) expl = expl + C{0.0, force};
    if (gd != 0.0) expl = expl - gd * unk1;
    r1 = (a.ode0 * unk1);
    r1 = C{r1.x / a.deltat, r1.y / a.deltat} + imp1 - a.ode2 * o1old + a.ode1 * expl;
    o1 = expl;
    }
    if constexpr (FIVE) {
      // buildrhs (:169-207) summed over the six products and regrouped by
      // the combinations P0 = vv - uu, P1 = vv - ww, P2 = uv, P3 = vw, P4 = uw:
      //   eta: alfa beta D0 (P1 - P0) + (beta^2 - alfa^2) D0 P4 - i beta D1 P2 + i alfa D1 P3
      //        ((0,0): (-Re D1 P2, -Re D1 P3))
      //   d2v: alfa^2 D1 P0 + beta^2 D1 P1 - 2 alfa beta D1 P4 + i alfa (D2 + k^2 D0) P2
      //        + i beta (D2 + k^2 D0) P3
      Cx<PR> d0[5], d1[5], d2[5];
#pragma unroll
      for (int q = 0; q < 5; q++) { d0[q] = Cx<PR>{0, 0}; d1[q] = Cx<PR>{0, 0}; d2[q] = Cx<PR>{0, 0}; }
#pragma unroll
      for (int j = 0; j < 5; j++) {
        const PR c0 = (PR)DERL(iy, 0, j - 2), c1 = (PR)DERL(iy, 1, j - 2), c2 = (PR)DERL(iy, 2, j - 2);
#pragma unroll
        for (int q = 0; q < 5; q++) {  // (D2 of P0, P1, P4 is not used and drops out)
          d0[q] = d0[q] + c0 * p[q][j];
          d1[q] = d1[q] + c1 * p[q][j];
          d2[q] = d2[q] + c2 * p[q][j];
        }
      }
      Cx<PR> e, ex;
      if (m00) {
        const PR rw = a.no_mean_vw ? (PR)0 : -d1[3].x;
        e = Cx<PR>{-d1[2].x, rw};
      } else {
        e = (alp * bep) * (d0[1] - d0[0]) + (bep * bep - alp * alp) * d0[4] - imul(bep, d1[2]) + imul(alp, d1[3]);
      }
      ex = (alp * alp) * d1[0] + (bep * bep) * d1[1] - ((PR)2 * alp * bep) * d1[4] +
           imul(alp, d2[2] + kkp * d0[2]) + imul(bep, d2[3] + kkp * d0[3]);
      const C exd{(double)ex.x, (double)ex.y}, ed{(double)e.x, (double)e.y};
      r2 = r2 + a.ode1 * exd;
      o2 = o2 + exd;
      r1 = r1 + a.ode1 * ed;
      o1 = o1 + ed;
    } else if constexpr (NONLIN) {
      // buildrhs for the six products in the order uu vv ww uv vw uw (:169-207)
      PR e0[5], e1[5], e2[5];
#pragma unroll
      for (int j = 0; j < 5; j++) {
        e0[j] = (PR)DERL(iy, 0, j - 2);
        e1[j] = (PR)DERL(iy, 1, j - 2);
        e2[j] = (PR)DERL(iy, 2, j - 2);
      }
#pragma unroll
      for (int q = Q0; q < Q1; q++) {
        Cx<PR> d0{0, 0}, d1{0, 0}, d2{0, 0};
#pragma unroll
        for (int j = 0; j < 5; j++) {
          d0 = d0 + e0[j] * p[q][j];
          d1 = d1 + e1[j] * p[q][j];
          d2 = d2 + e2[j] * p[q][j];
        }
        Cx<PR> ru{0, 0}, rw{0, 0}, ex;
        if (q == 0) { ru = -imul(alp, d0); ex = d1 * (-(alp * alp)); }
        else if (q == 1) { ex = kkp * d1; }
        else if (q == 2) { rw = -imul(bep, d0); ex = d1 * (-(bep * bep)); }
        else if (q == 3) { ru = -d1; ex = imul(alp, d2) + imul(alp * kkp, d0); }
        else if (q == 4) {
          rw = (a.no_mean_vw && m00) ? Cx<PR>{0, 0} : -d1;
          ex = imul(bep, d2) + imul(bep * kkp, d0);
        }
        else { ru = -imul(bep, d0); rw = -imul(alp, d0); ex = d1 * (-((PR)2 * alp * bep)); }
        const C exd{(double)ex.x, (double)ex.y};
        r2 = r2 + a.ode1 * exd;
        o2 = o2 + exd;
        C e;
        if (m00) e = C{(double)ru.x, (double)rw.x};
        else {
          const Cx<PR> ep = imul(bep, ru) - imul(alp, rw);
          e = C{(double)ep.x, (double)ep.y};
        }
        r1 = r1 + a.ode1 * e;
        o1 = o1 + e;
      }
    }
    if constexpr (SHIFT) {
      // shear_shift, default advection (src/hst_equations.f90:245-257)
      const C f = cexpi(-((D.alfa0 * ix) * a.dgx + (D.beta0 * iz) * a.dgz) * y[iy + 2]);
      r1 = r1 * f; r2 = r2 * f; o1 = o1 * f; o2 = o2 * f;
    }
    if (masked) {  // damping_mask (src/hst_damping.f90:119-133), before the shift
      r1 = C{0, 0}; r2 = C{0, 0}; o1 = C{0, 0}; o2 = C{0, 0};
    }
    stn(R1 + c0, r1); stn(R2 + c0, r2); stn(O1 + c0, o1); stn(O2 + c0, o2);
  }
}

int launch_buildrhs(const Dims &D, int mixed, const RhsArgs &a, const double2 *V1, const double2 *V2,
                    const double2 *V3, const void *const P[6], double2 *R1, double2 *R2, void *O1, void *O2,
                    const double *der, const double *y, hipStream_t s) {
  const int rch = RHS_ROWS;  // rows per lane: the 5-row windows re-read 4 rows per chunk
  dim3 grid((D.ncol + 255) / 256, (D.ny + rch - 1) / rch);
  const size_t pstride = (size_t)D.NYg * D.ncol;
  // 0 prepare, 1/2 one group, 3 all, 4 all from the five combinations, 5 as 4
  // with eta and dv/dy in V1 and V3
  const int mode = a.nonlinear;
  const bool sh = a.shift != 0;
#define L(PR, MD, SH)                                                                                            \
  hipLaunchKernelGGL((k_buildrhs<PR, MD, SH>), grid, dim3(256), 0, s, D, a, V1, V2, V3,                          \
                     (const typename PTraits<PR>::T2 *)P[0], pstride, R1, R2, (typename PTraits<PR>::T2 *)O1,    \
                     (typename PTraits<PR>::T2 *)O2, der, y, rch)
#define M(PR)                                                      \
  switch (mode) {                                                  \
    case 0: if (sh) L(PR, 0, true); else L(PR, 0, false); break;   \
    case 1: L(PR, 1, false); break;                                \
    case 2: L(PR, 2, false); break;                                \
    case 4: if (sh) L(PR, 4, true); else L(PR, 4, false); break;   \
    case 5: if (sh) L(PR, 5, true); else L(PR, 5, false); break;   \
    default: if (sh) L(PR, 3, true); else L(PR, 3, false); break;  \
  }
  if (mixed) { M(float) } else { M(double) }
#undef M
#undef L
  return hipGetLastError();
}

// u, w from eta (in V1) and dv/dy (in V3) in place, on every row (ghost rows
// included): the time step leaves the velocity in that form between its
// calls, and this restores u, v, w.
__global__ void __launch_bounds__(256) k_uw_from_eta(Dims D, double2 *__restrict__ V1, double2 *__restrict__ V3) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const double al = D.alfa0 * ix, be = D.beta0 * iz, kk = al * al + be * be;
  const size_t o = (size_t)blockIdx.y * D.ncol + col;
  C u, w;
  recover_uw(al, be, kk, ix == 0 && iz == 0, ld(V3 + o), ld(V1 + o), u, w);
  st(V1 + o, u);
  st(V3 + o, w);
}

int launch_uw_from_eta(const Dims &D, double2 *V1, double2 *V3, hipStream_t s) {
  dim3 grid((D.ncol + 255) / 256, D.NYg);
  hipLaunchKernelGGL(k_uw_from_eta, grid, dim3(256), 0, s, D, V1, V3);
  return hipGetLastError();
}

// shear_shift with the default advection as a pass of its own
// (src/hst_equations.f90:245-257): rhs, oldrhs times the node phase.
template <typename OT>
__global__ void __launch_bounds__(256)
    k_phase4(Dims D, double2 *R1, double2 *R2, OT *O1, OT *O2, const double *__restrict__ y, double dgx,
             double dgz) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const int iy = blockIdx.y;
  const size_t o = (size_t)(iy + 2) * D.ncol + col;
  const C f = cexpi(-((D.alfa0 * ix) * dgx + (D.beta0 * iz) * dgz) * y[iy + 2]);
  st(R1 + o, ld(R1 + o) * f); st(R2 + o, ld(R2 + o) * f);
  st(O1 + o, ld(O1 + o) * f); st(O2 + o, ld(O2 + o) * f);
}

int launch_phase4(const Dims &D, int mixed, double2 *R1, double2 *R2, void *O1, void *O2, const double *y,
                  double dgx, double dgz, hipStream_t s) {
  dim3 grid((D.ncol + 255) / 256, D.ny);
  if (mixed)
    hipLaunchKernelGGL(k_phase4<float2>, grid, dim3(256), 0, s, D, R1, R2, (float2 *)O1, (float2 *)O2, y, dgx, dgz);
  else
    hipLaunchKernelGGL(k_phase4<double2>, grid, dim3(256), 0, s, D, R1, R2, (double2 *)O1, (double2 *)O2, y, dgx,
                       dgz);
  return hipGetLastError();
}

// ------------------------------------------------------------------ the border system

// The reduced (2x2) system of a line with one slab (src/hst_linsolve.f90:383-453)
// from the record: rr = the real record entries by 1-based index, the complex
// ones (x0 = rec 1-2, xx1 = rec 7-8, xm2 = rec 13-14, xm1 = rec 19-20) and the
// right-hand sides of the two border rows.  Returns the border (b1, b2).
__device__ __forceinline__ void border_solve(const double *rr, double c1m4, double c1m3, double d11, double d12,
                                             double c10, double c2m3, double d21, double d22, double c20, double c21,
                                             C ph, C x0, C xx1, C xm2, C xm1, C r25, C r27, C &b1, C &b2) {
  const C cph = conj(ph), php = ph;
  C A11{0, 0}, A12{0, 0}, A21{0, 0}, A22{0, 0}, f;
  b1 = r25 - c1m4 * xm2 - c1m3 * xm1 - (c10 * php) * x0;
  A11 = A11 - cph * (c1m4 * rr[15] + c1m3 * rr[21]);
  A12 = A12 - cph * (c1m4 * rr[16] + c1m3 * rr[22]);
  A11 = A11 + C{d11, 0.0} - C{c1m4 * rr[17] + c1m3 * rr[23], 0.0};
  A12 = A12 + C{d12, 0.0} - C{c1m4 * rr[18] + c1m3 * rr[24], 0.0};
  f = (c10 * php) * conj(ph);
  A11 = A11 - f * rr[3];
  A12 = A12 - f * rr[4];
  A11 = A11 - (c10 * php) * rr[5];
  A12 = A12 - (c10 * php) * rr[6];
  b2 = r27 - c2m3 * xm1 - php * (c20 * x0 + c21 * xx1);
  A21 = A21 - (cph * c2m3) * rr[21];
  A22 = A22 - (cph * c2m3) * rr[22];
  A21 = A21 + C{d21, 0.0} - C{c2m3 * rr[23], 0.0};
  A22 = A22 + C{d22, 0.0} - C{c2m3 * rr[24], 0.0};
  f = php * conj(ph);
  A21 = A21 - f * (c20 * rr[3] + c21 * rr[9]);
  A22 = A22 - f * (c20 * rr[4] + c21 * rr[10]);
  A21 = A21 - php * (c20 * rr[5] + c21 * rr[11]);
  A22 = A22 - php * (c20 * rr[6] + c21 * rr[12]);
  const C piv = cdiv(C{1.0, 0.0}, A11);
  f = A21 * piv;
  A22 = A22 - f * A12;
  b2 = b2 - f * b1;
  b2 = cdiv(b2, A22);
  b1 = cdiv(b1 - A12 * b2, A11);
}

// ------------------------------------------------------------------ per-column solves

template <int K>
__device__ __forceinline__ double coefk(double lam, double ni, double kk, const double *der, int iy, int j) {
  const double d0 = DER(der, iy, 0, j);
  if constexpr (K == KD2V) {
    const double d2 = DER(der, iy, 2, j), d4 = DER(der, iy, 3, j);
    return lam * (d2 - kk * d0) - ni * (d4 - 2.0 * kk * d2 + kk * kk * d0);
  } else if constexpr (K == KETA) {
    return lam * d0 - ni * (DER(der, iy, 2, j) - kk * d0);
  } else if constexpr (K == KPOISSON) {
    return DER(der, iy, 2, j) - kk * d0;
  } else {
    return d0;
  }
}

// Row pipelines of the per-column recurrences: body(i, v) for i = 0..n-1
// (ascending) or n-1..0 (descending) with v = load(i) issued PF rows ahead,
// so that a wavefront keeps PF rows of loads in flight while its recurrence
// runs.  n, i are uniform over the wavefront.
template <int PF, class T, class LOAD, class BODY>
__device__ __forceinline__ void rows_ascending(int n, LOAD load, BODY body) {
  T buf[PF];
#pragma unroll
  for (int k = 0; k < PF; k++)
    if (k < n) buf[k] = load(k);
  int i = 0;
  for (; i + PF <= n; i += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const T v = buf[k];
      if (i + k + PF < n) buf[k] = load(i + k + PF);
      body(i + k, v);
    }
  }
#pragma unroll
  for (int k = 0; k < PF; k++)
    if (i + k < n) body(i + k, buf[k]);
}

template <int PF, class T, class LOAD, class BODY>
__device__ __forceinline__ void rows_descending(int n, LOAD load, BODY body) {
  rows_ascending<PF, T>(
      n, [&](int k) { return load(n - 1 - k); }, [&](int k, const T &v) { body(n - 1 - k, v); });
}

struct PentaArgs {
  const double2 *src[2];
  double2 *dst[2];
  PentaWS ws[2];
  const double4 *tab[2];  // factor rows per (ix, |iz|) group, or null: the workspace holds them
};

// a row of the elimination workspace, as the back substitution reads it
struct WSRow {
  C xr;
  double b1, b2, u1, u2;
};

// rows of loads in flight per sweep (measured at nx = nz = 256, ny = 513)
constexpr int PF_FWD = 8, PF_BWD = 2, PF_PENTA = 4;

// The elimination of one cyclic pentadiagonal line with the rows generated
// from der (line_solve, src/hst_linsolve.f90:282-380, one slab), split into
// the part that depends only on the matrix (kind, lambda, k^2) and the part
// that carries the right-hand side.  step() advances the matrix part by row i
// and returns the multipliers l2, l1, the inverse pivot rp and the row's
// upper entries t1, t2 and border columns p, q; the record sums accumulate.
struct PentaFactor {
  double u11 = 0, u12 = 0, p1 = 0, q1 = 0, u21 = 0, u22 = 0, p2 = 0, q2 = 0;
  double w1 = 0, w2 = 0, v1 = 0, v2 = 0, sp0 = 0, sq0 = 0, sp1 = 0, sq1 = 0;
  double ww = 0, vv = 0;  // the weights of the row just stepped
  template <int K>
  __device__ __forceinline__ void step(int i, double lam, double ni, double kk, const double *__restrict__ der,
                                       double &l2, double &l1, double &rp, double &t1, double &t2, double &p,
                                       double &q) {
    double a[5];
#pragma unroll
    for (int j = 0; j < 5; j++) a[j] = coefk<K>(lam, ni, kk, der, i, j - 2);
    p = 0.0;
    q = 0.0;
    if (i == 0) { p = a[0]; q = a[1]; }
    if (i == 1) q = a[0];
    l2 = 0.0;
    l1 = 0.0;
    if (i >= 2) {
      const double l = a[0];
      a[1] = a[1] - l * u21;
      a[2] = a[2] - l * u22;
      p = p - l * p2; q = q - l * q2;
      l2 = l;
    }
    if (i >= 1) {
      const double l = a[1];
      a[2] = a[2] - l * u11;
      a[3] = a[3] - l * u12;
      p = p - l * p1; q = q - l * q1;
      l1 = l;
    }
    rp = 1.0 / a[2];
    t1 = a[3] * rp;
    t2 = a[4] * rp;
    p = p * rp; q = q * rp;
    ww = -(u11 * w1 + u22 * w2);
    if (i == 0) ww = 1.0;
    vv = -(u11 * v1 + u22 * v2);
    if (i == 1) vv = 1.0;
    sp0 = sp0 + ww * p; sq0 = sq0 + ww * q;
    sp1 = sp1 + vv * p; sq1 = sq1 + vv * q;
    w2 = w1; w1 = ww; v2 = v1; v1 = vv;
    u21 = u11; u22 = u12; p2 = p1; q2 = q1;
    u11 = t1; u12 = t2; p1 = p; q1 = q;
  }
  // the state that step() carries from row to row into the matrix rows
  // (t1, t2, p, q): the last two rows' upper entries and border columns
  __device__ __forceinline__ void save(double4 *a, double4 *b) const {
    *a = make_double4(u11, u12, p1, q1);
    *b = make_double4(u21, u22, p2, q2);
  }
  __device__ __forceinline__ void restore(const double4 &a, const double4 &b) {
    u11 = a.x; u12 = a.y; p1 = a.z; q1 = a.w;
    u21 = b.x; u22 = b.y; p2 = b.z; q2 = b.w;
  }
};

__device__ __forceinline__ int penta_group(const Dims &D, int ix, int iz) { return ix * (D.nz + 1) + (iz < 0 ? -iz : iz); }

// Lane order of the line kernels that read the factor tables: slot p of the
// (ix, |iz|) groups (p = ix (nz+1) + |iz|, the table's own index) runs the
// column (ix, |iz|) in lane p%32 of the first half of wavefront p/32 and the
// column (ix, -|iz|) in the same lane of the second half, so that one load of
// a table row (and of a z phase) serves both; the second-half lane of iz = 0
// idles.  Each half-wavefront reads 32 neighbouring columns.  False for an
// idle lane.
__device__ __forceinline__ bool pair_lane(const Dims &D, int L, int &ix, int &iz, int &col, int &g) {
  const int h = (L >> 5) & 1;
  g = (L >> 6) * 32 + (L & 31);
  if (g >= D.NXm * (D.nz + 1)) return false;
  ix = g / (D.nz + 1);
  const int k = g - ix * (D.nz + 1);
  if (h && k == 0) return false;
  iz = h ? -k : k;
  col = ix * D.NZm + iz + D.nz;
  return true;
}

static int pair_lanes(const Dims &D) { return (D.NXm * (D.nz + 1) + 31) / 32 * 64; }

// The matrix part of the back substitution for every (ix, |iz|) group, one
// kind and lambda: the rows (t1, t2, p, q) come from step() of the matrix
// factorisation, whose state entering every PC-th row is kept here,
// tab[2 c + h][g] for chunk c (h = 0: u11, u12, p1, q1; 1: u21, u22, p2, q2);
// a line recomputes a chunk's rows from it when its back substitution reaches
// the chunk.  Columns (ix, iz) and (ix, -iz) share k^2 and so every bit of
// these rows.
constexpr int PC = 8;
template <int K>
__global__ void __launch_bounds__(256) k_penta_factor(Dims D, double lam, const double *__restrict__ der,
                                                      double4 *__restrict__ tab) {
  const int ng = (D.nx + 1) * (D.nz + 1);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int ix = g / (D.nz + 1), iz = g - ix * (D.nz + 1);
  const double al = D.alfa0 * ix, be = D.beta0 * iz;
  const double kk = al * al + be * be;
  PentaFactor f;
  for (int i = 0; i < D.ny - 2; i++) {
    if (i % PC == 0) f.save(tab + (size_t)(2 * (i / PC)) * ng + g, tab + (size_t)(2 * (i / PC) + 1) * ng + g);
    double l2, l1, rp, t1, t2, p, q;
    f.step<K>(i, lam, D.ni, kk, der, l2, l1, rp, t1, t2, p, q);
  }
}

// The back substitution of one line from its top row down (penta_backward,
// src/hst_linsolve.f90:359-483, one slab) with the matrix rows recomputed
// chunk by chunk from the group's checkpoints: x(i) = xr(i) - p zb1 - q zb2 -
// t1 x(i+1) - t2 x(i+2), xr(i) = loadx(i), each x(i) to emit(i, x).  x1, x2
// enter as x(mi), x(mi+1) (the border rows).
template <int K, class LOADX, class EMIT>
__device__ __forceinline__ void penta_back(const Dims &D, int g, double lam, double kk,
                                           const double *__restrict__ der, const double4 *__restrict__ tab,
                                           C zb1, C zb2, C x1, C x2, LOADX loadx, EMIT emit) {
  const int mi = D.ny - 2, ng = (D.nx + 1) * (D.nz + 1);
  // the next chunk's checkpoint is loaded while this chunk is substituted
  // (3-4 % faster in the fused shift and solve than loading it at its chunk)
  double4 ta = tab[(size_t)(2 * ((mi - 1) / PC)) * ng + g], tb = tab[(size_t)(2 * ((mi - 1) / PC) + 1) * ng + g];
  for (int c = (mi - 1) / PC; c >= 0; c--) {
    const int c0 = c * PC, n = min(PC, mi - c0);
    C xr[PC];
#pragma unroll
    for (int j = 0; j < PC; j++)
      if (j < n) xr[j] = loadx(c0 + j);
    PentaFactor f;
    f.restore(ta, tb);
    double t1[PC], t2[PC], pp[PC], qq[PC];
#pragma unroll
    for (int j = 0; j < PC; j++) {
      if (j < n) {
        double l2, l1, rp;
        f.step<K>(c0 + j, lam, D.ni, kk, der, l2, l1, rp, t1[j], t2[j], pp[j], qq[j]);
      }
    }
    if (c > 0) {
      ta = tab[(size_t)(2 * c - 2) * ng + g];
      tb = tab[(size_t)(2 * c - 1) * ng + g];
    }
#pragma unroll
    for (int j = PC - 1; j >= 0; j--) {
      if (j < n) {
        const C x = xr[j] - pp[j] * zb1 - qq[j] * zb2 - t1[j] * x1 - t2[j] * x2;
        emit(c0 + j, x);
        x2 = x1;
        x1 = x;
      }
    }
  }
}

// One cyclic pentadiagonal line per lane (line_solve, src/hst_linsolve.f90:147-483,
// one slab): forward elimination of rows 0..ny-3 with the rows generated from
// der, record in registers, 2x2 border system, back substitution.  With a
// factor table the workspace holds only the eliminated right-hand side.
template <int K, bool TAB>
__device__ __forceinline__ void penta_line(const Dims &D, int col, int ix, int iz, int g, double lam,
                                           const double *__restrict__ der, const double2 *__restrict__ src,
                                           double2 *__restrict__ dst, const PentaWS &ws,
                                           const double4 *__restrict__ tab, double sx, double sz) {
  const size_t nc = D.ncol;
  const int m = D.ny, mi = m - 2;
  if (ix == 0 && iz == 0 && K == KD2V) {
    for (int i = 0; i < m; i++) st(dst + (size_t)(i + 2) * nc + col, C{0.0, 0.0});
    return;
  }
  if (ix == 0 && iz == 0 && K == KPOISSON) return;
  const double al = D.alfa0 * ix, be = D.beta0 * iz;
  const double kk = al * al + be * be, ni = D.ni;
  PentaFactor f;
  C x1{0, 0}, x2{0, 0}, sx0{0, 0}, sx1{0, 0};
  rows_ascending<PF_PENTA, C>(
      mi, [&](int i) { return ld(src + (size_t)(i + 2) * nc + col); },
      [&](int i, const C &bin) {
        double l2, l1, rp, t1, t2, p, q;
        f.step<K>(i, lam, ni, kk, der, l2, l1, rp, t1, t2, p, q);
        C bx = bin;
        if (i >= 2) bx = bx - l2 * x2;
        if (i >= 1) bx = bx - l1 * x1;
        bx = bx * rp;
        const size_t o = (size_t)i * nc + col;
        if constexpr (!TAB) { ws.U1[o] = t1; ws.U2[o] = t2; ws.B1[o] = p; ws.B2[o] = q; }
        st(ws.XR + o, bx);
        sx0 = sx0 + f.ww * bx;
        sx1 = sx1 + f.vv * bx;
        x2 = x1;
        x1 = bx;
      });
  double rr[25];
  rr[3] = f.sp0; rr[4] = f.sq0; rr[5] = f.w2 * f.u22 + f.w1 * f.u11; rr[6] = f.w1 * f.u12;
  rr[9] = f.sp1; rr[10] = f.sq1; rr[11] = f.v2 * f.u22 + f.v1 * f.u11; rr[12] = f.v1 * f.u12;
  rr[15] = f.p2 - f.u21 * f.p1; rr[16] = f.q2 - f.u21 * f.q1; rr[17] = f.u22 - f.u21 * f.u11; rr[18] = -f.u21 * f.u12;
  rr[21] = f.p1; rr[22] = f.q1; rr[23] = f.u11; rr[24] = f.u12;
  const C xm2 = x2 - f.u21 * x1, xm1 = x1;
  const C r25 = ld(src + (size_t)(m - 2 + 2) * nc + col), r27 = ld(src + (size_t)(m - 1 + 2) * nc + col);
  const int ga = m - 2;
  const C ph = cexpi(-(D.alfa0 * ix * sx + D.beta0 * iz * sz));
  C b1, b2;
  border_solve(rr, coefk<K>(lam, ni, kk, der, ga, -2), coefk<K>(lam, ni, kk, der, ga, -1),
               coefk<K>(lam, ni, kk, der, ga, 0), coefk<K>(lam, ni, kk, der, ga, 1),
               coefk<K>(lam, ni, kk, der, ga, 2), coefk<K>(lam, ni, kk, der, ga + 1, -2),
               coefk<K>(lam, ni, kk, der, ga + 1, -1), coefk<K>(lam, ni, kk, der, ga + 1, 0),
               coefk<K>(lam, ni, kk, der, ga + 1, 1), coefk<K>(lam, ni, kk, der, ga + 1, 2), ph, sx0, sx1, xm2,
               xm1, r25, r27, b1, b2);
  st(dst + (size_t)(m - 2 + 2) * nc + col, b1);
  st(dst + (size_t)(m - 1 + 2) * nc + col, b2);
  const C cph = conj(ph);
  const C yb1 = cph * b1, yb2 = cph * b2;
  if constexpr (TAB) {
    penta_back<K>(
        D, g, lam, kk, der, tab, yb1, yb2, b1, b2, [&](int i) { return ld(ws.XR + (size_t)i * nc + col); },
        [&](int i, const C &x) { st(dst + (size_t)(i + 2) * nc + col, x); });
  } else {
    C xk1 = b1, xk2 = b2;
    rows_descending<PF_BWD, WSRow>(
        mi,
        [&](int i) {
          const size_t o = (size_t)i * nc + col;
          return WSRow{ld(ws.XR + o), ws.B1[o], ws.B2[o], ws.U1[o], ws.U2[o]};
        },
        [&](int i, const WSRow &w) {
          const C xk = w.xr - w.b1 * yb1 - w.b2 * yb2 - w.u1 * xk1 - w.u2 * xk2;
          st(dst + (size_t)(i + 2) * nc + col, xk);
          xk2 = xk1;
          xk1 = xk;
        });
  }
}

// blockIdx.y selects the system: K0 for the first, K1 for the second
template <int K0, int K1, bool TAB>
__global__ void __launch_bounds__(256)
    k_penta(Dims D, PentaArgs pa, double lam, const double *__restrict__ der, double sx, double sz) {
  // with the factor tables the +iz and -iz columns share a wavefront (one
  // load of a table row serves both) and neighbouring blocks stay on one XCD
  const int l = TAB ? xcd_linear() : blockIdx.x + gridDim.x * blockIdx.y;
  const int sys = l / gridDim.x;
  const int L = (l % gridDim.x) * blockDim.x + threadIdx.x;
  int ix, iz, col, g;
  if constexpr (TAB) {
    if (!pair_lane(D, L, ix, iz, col, g)) return;
  } else {
    if (L >= D.ncol) return;
    col = L;
    col_ix_iz(D, col, ix, iz);
    g = penta_group(D, ix, iz);
  }
  if (sys == 0) penta_line<K0, TAB>(D, col, ix, iz, g, lam, der, pa.src[0], pa.dst[0], pa.ws[0], pa.tab[0], sx, sz);
  else penta_line<K1, TAB>(D, col, ix, iz, g, lam, der, pa.src[1], pa.dst[1], pa.ws[1], pa.tab[1], sx, sz);
}

size_t penta_table_bytes(const Dims &D) {
  return (size_t)2 * ((D.ny - 2 + PC - 1) / PC) * (D.nx + 1) * (D.nz + 1) * sizeof(double4);
}

int launch_penta_factor(const Dims &D, int kind, double lambda, const double *der, void *tab, hipStream_t s) {
  const int ng = (D.nx + 1) * (D.nz + 1);
  dim3 grid((ng + 255) / 256);
  if (kind == KD2V) hipLaunchKernelGGL(k_penta_factor<KD2V>, grid, dim3(256), 0, s, D, lambda, der, (double4 *)tab);
  else if (kind == KETA) hipLaunchKernelGGL(k_penta_factor<KETA>, grid, dim3(256), 0, s, D, lambda, der, (double4 *)tab);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

int launch_penta(const Dims &D, int nsys, const int *kinds, double lambda, const double2 *const *src,
                 double2 *const *dst, const PentaWS *ws, const void *const *tabs, const double *der, double sx,
                 double sz, hipStream_t s) {
  PentaArgs pa;
  bool tab = tabs != nullptr;
  for (int k = 0; k < nsys; k++) {
    pa.src[k] = src[k];
    pa.dst[k] = dst[k];
    pa.ws[k] = ws[k];
    pa.tab[k] = tabs ? (const double4 *)tabs[k] : nullptr;
    if (!pa.tab[k]) tab = false;
  }
  dim3 grid(((tab ? pair_lanes(D) : D.ncol) + 255) / 256, nsys);
#define L(A, B, T) hipLaunchKernelGGL((k_penta<A, B, T>), grid, dim3(256), 0, s, D, pa, lambda, der, sx, sz)
  if (nsys == 2 && kinds[0] == KD2V && kinds[1] == KETA) {
    if (tab) L(KD2V, KETA, true); else L(KD2V, KETA, false);
  } else if (nsys == 1 && kinds[0] == KD2V) L(KD2V, KD2V, false);
  else if (nsys == 1 && kinds[0] == KETA) L(KETA, KETA, false);
  else if (nsys == 1 && kinds[0] == KPOISSON) L(KPOISSON, KPOISSON, false);
  else return hipErrorInvalidValue;
#undef L
  return hipGetLastError();
}

// ------------------------------------------------------------------ shared-factor D0 solves

// rows table: [i][L2 L1 RP U1 U2 B1 B2 W V]
#define F0(i, k) f0[(i) * 9 + (k)]

// Forward sweep of D0 x = r for one lane, XR stored; r = src (RHS 0) or the
// D1 stencil of src (RHS 1, ghost rows filled).  Also returns the border
// rows' right-hand sides r(ny-2), r(ny-1).
template <int RHS, typename ST = double2, typename XT = double2>
__device__ __forceinline__ void d0_forward(const Dims &D, const double *__restrict__ f0,
                                           const double *__restrict__ der, const ST *__restrict__ src,
                                           XT *__restrict__ XR, int col, C &x1, C &x2, C &sx0, C &sx1, C &r25,
                                           C &r27) {
  const size_t nc = D.ncol;
  const int m = D.ny, mi = m - 2;
  x1 = C{0, 0}; x2 = C{0, 0}; sx0 = C{0, 0}; sx1 = C{0, 0};
  C w[5];  // RHS 1: the window src(i-2 .. i+2)
  if constexpr (RHS == 1) {
#pragma unroll
    for (int k = 0; k < 4; k++) w[k + 1] = ld(src + (size_t)k * nc + col);
  }
  auto stencil = [&](int i) {
    C r{0, 0};
#pragma unroll
    for (int j = -2; j <= 2; j++) r = r + DER(der, i, 1, j) * w[j + 2];
    return r;
  };
  auto push = [&](const C &v) {
    w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4]; w[4] = v;
  };
  rows_ascending<PF_FWD, C>(
      mi, [&](int i) { return ld(src + (size_t)(i + (RHS == 0 ? 2 : 4)) * nc + col); },
      [&](int i, const C &v) {
        C bx;
        if constexpr (RHS == 0) {
          bx = v;
        } else {
          push(v);
          bx = stencil(i);
        }
        if (i >= 2) bx = bx - F0(i, 0) * x2;
        if (i >= 1) bx = bx - F0(i, 1) * x1;
        bx = bx * F0(i, 2);
        st(XR + (size_t)i * nc + col, bx);
        sx0 = sx0 + F0(i, 7) * bx;
        sx1 = sx1 + F0(i, 8) * bx;
        x2 = x1;
        x1 = bx;
      });
  if constexpr (RHS == 0) {
    r25 = ld(src + (size_t)(m - 2 + 2) * nc + col);
    r27 = ld(src + (size_t)(m - 1 + 2) * nc + col);
  } else {
    push(ld(src + (size_t)(m + 2) * nc + col));
    r25 = stencil(m - 2);
    push(ld(src + (size_t)(m + 3) * nc + col));
    r27 = stencil(m - 1);
  }
}

__device__ __forceinline__ void d0_border(const D0Const &fc, C ph, C x1, C x2, C sx0, C sx1, C r25, C r27, C &b1,
                                          C &b2) {
  const C xm2 = x2 - fc.rec[0] * x1, xm1 = x1;
  border_solve(fc.rec, fc.c1m4, fc.c1m3, fc.d11, fc.d12, fc.c10, fc.c2m3, fc.d21, fc.d22, fc.c20, fc.c21, ph, sx0,
               sx1, xm2, xm1, r25, r27, b1, b2);
}

// D0 x = src (RHS 0) or D0 x = D1 src (RHS 1); OUT 0: x to dst; 1: recover
// u, w from x = dv/dy and eta = dst (src/hst_equations.f90:303-324); 2: block
// sums of w |x|^2 (the dissipation's y derivatives, src/hst_stats.f90:200-223).
template <int RHS, int OUT>
__global__ void __launch_bounds__(256)
    k_d0solve(Dims D, D0Const fc, const double *__restrict__ f0, const double *__restrict__ der,
              const double *__restrict__ dyl, const double2 *__restrict__ src, double2 *__restrict__ dst,
              double2 *__restrict__ dst3, double2 *__restrict__ XR, double sx, double sz, double *__restrict__ part,
              const int *__restrict__ inlayer, int skip00) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = col < D.ncol;
  const int cl = act ? col : 0;
  int ix, iz;
  col_ix_iz(D, cl, ix, iz);
  const size_t nc = D.ncol;
  const int m = D.ny;
  double acc = 0.0, ain = 0.0, aout = 0.0;
  const bool skip = skip00 && ix == 0 && iz == 0;
  if (act) {
    C x1, x2, sx0, sx1, r25, r27;
    d0_forward<RHS>(D, f0, der, src, XR, col, x1, x2, sx0, sx1, r25, r27);
    const C ph = cexpi(-(D.alfa0 * ix * sx + D.beta0 * iz * sz));
    C b1, b2;
    d0_border(fc, ph, x1, x2, sx0, sx1, r25, r27, b1, b2);
    const C cph = conj(ph);
    const C yb1 = cph * b1, yb2 = cph * b2;
    const double al = D.alfa0 * ix, be = D.beta0 * iz, kk = al * al + be * be;
    const bool m00 = (ix == 0 && iz == 0);
    auto emit = [&](int i, C x, C eta) {
      const size_t o = (size_t)(i + 2) * nc + col;
      if constexpr (OUT == 0) {
        st(dst + o, x);
      } else if constexpr (OUT == 1) {
        C uu, ww;
        recover_uw(al, be, kk, m00, x, eta, uu, ww);
        st(dst + o, uu);
        st(dst3 + o, ww);
      } else if (!skip) {
        const double wgt = (ix == 0 ? dyl[i] : 2.0 * dyl[i]);
        const double q = wgt * (x.x * x.x + x.y * x.y);
        acc += q;
        if (inlayer && inlayer[i]) ain += q;
        else aout += q;
      }
    };
    auto eta_at = [&](int i) -> C {
      if constexpr (OUT == 1) return ld(dst + (size_t)(i + 2) * nc + col);
      else return C{0, 0};
    };
    emit(m - 2, b1, eta_at(m - 2));
    emit(m - 1, b2, eta_at(m - 1));
    C xk1 = b1, xk2 = b2;
    struct Row {
      C xr, eta;
    };
    rows_descending<PF_BWD, Row>(
        m - 2, [&](int i) { return Row{ld(XR + (size_t)i * nc + col), eta_at(i)}; },
        [&](int i, const Row &r) {
          const C xk = r.xr - F0(i, 5) * yb1 - F0(i, 6) * yb2 - F0(i, 3) * xk1 - F0(i, 4) * xk2;
          emit(i, xk, r.eta);
          xk2 = xk1;
          xk1 = xk;
        });
  }
  if constexpr (OUT == 2) {
    __shared__ double red[3][4];
    double v[3] = {acc, ain, aout};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
      if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      const int k = threadIdx.x;
      part[(size_t)blockIdx.x * 3 + k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    }
  }
}

int launch_d0solve(const Dims &D, int rhs_mode, int out_mode, const double2 *src, double2 *dst, double2 *dst3,
                   double2 *XR, const double *f0rows, const D0Const &fc, const double *der, const double *dyl,
                   double sx, double sz, double *part, int *n_part, hipStream_t s, const int *inlayer, int skip00) {
  dim3 grid((D.ncol + 255) / 256);
  if (n_part) *n_part = grid.x;
#define L(R, O)                                                                                                 \
  hipLaunchKernelGGL((k_d0solve<R, O>), grid, dim3(256), 0, s, D, fc, f0rows, der, dyl, src, dst, dst3, XR, sx, sz, \
                     part, inlayer, skip00)
  if (rhs_mode == 0) {
    if (out_mode == 0) L(0, 0); else if (out_mode == 1) L(0, 1); else L(0, 2);
  } else {
    if (out_mode == 0) L(1, 0); else if (out_mode == 1) L(1, 1); else L(1, 2);
  }
#undef L
  return hipGetLastError();
}

// ------------------------------------------------------------------ dv/dy in one sweep

// D0 x = D1 v (the time step's dv/dy) in one pass over the rows, without the
// workspace: the forward elimination is d0_forward's, but the eliminated rows
// stay in registers, DYW_N = DYW_B + DYW_W of them, and every DYW_B rows the
// back substitution runs from the top of that window down through its lowest
// DYW_B rows, started from zero and without the border terms.  The inverse
// of D0's unit upper factor decays by about 0.29 a row on these grids (the
// host checks the bound for the deck, dy_window_ok), so a row emitted that
// way misses contributions below 1e-16 of the column's values.  The rows
// within DYW_W of the top stay in the window until the border is solved and
// take the exact back substitution; the first DYW_W rows, where the border
// columns have not decayed, are emitted without them and corrected at the end
// by yb1 s1(i) + yb2 s2(i), s1, s2 the back substitution of -B1 and -B2 (the
// same for every column; sig = [i][2]).  Per point v is read once and dv/dy
// written once: 32 B against the 64 B of k_d0solve's stored sweep.
constexpr int DYW_B = 8, DYW_W = 32, DYW_N = DYW_B + DYW_W;
// CL: the per-row coefficients (the D1 taps -2..2 and the F0 columns 0, 1,
// 2, 7, 8, 3, 4: DYC doubles a row) read from a copy in LDS.  Read as scalar
// loads, each row's were waited for one after another (about 50 waits a
// block of rows, the wavefronts' only other work their own recurrences):
// 0.81 against 0.64 ms at nx = nz = 256, ny = 513.  CL = 0 where the copy
// does not fit 64 KB (ny > 684).
constexpr int DYC = 12;

template <int CL>
__global__ void __launch_bounds__(256, 2)
    k_dy_window(Dims D, D0Const fc, const double *__restrict__ f0, const double *__restrict__ der,
                const double2 *__restrict__ src, double2 *__restrict__ dst, const double *__restrict__ sig, double sx,
                double sz) {
  const int m = D.ny, mi = m - 2;
  extern __shared__ double dyc[];
  if constexpr (CL) {
    const int n = mi * DYC;
    for (int b0 = 0; b0 < n; b0 += 16 * 256) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int k = min(b0 + u * 256 + (int)threadIdx.x, n - 1), i = k / DYC, j = k - i * DYC;
        const int fj = j == 5 ? 0 : (j == 6 ? 1 : (j == 7 ? 2 : (j == 8 ? 7 : (j == 9 ? 8 : (j == 10 ? 3 : 4)))));
        v[u] = j < 5 ? DER(der, i, 1, j - 2) : F0(i, fj);
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int k = b0 + u * 256 + threadIdx.x;
        if (k < n) dyc[k] = v[u];
      }
    }
    __syncthreads();
  }
  // coefficient j of row i: 0..4 the D1 taps, 5..11 F0 columns 0, 1, 2, 7, 8, 3, 4
  auto coef = [&](int i, int j) -> double {
    if constexpr (CL) return dyc[i * DYC + j];
    constexpr int fj[7] = {0, 1, 2, 7, 8, 3, 4};
    return j < 5 ? DER(der, i, 1, j - 2) : F0(i, fj[j - 5]);
  };
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const size_t nc = D.ncol;
  C w[5];  // the D1 stencil's window src(i - 2 .. i + 2)
#pragma unroll
  for (int k = 0; k < 4; k++) w[k + 1] = ld(src + (size_t)k * nc + col);
  C x1{0, 0}, x2{0, 0}, sx0{0, 0}, sx1{0, 0};
  // forward elimination of row i from src row i + 2 (array row i + 4)
  auto fwd = [&](int i, const C &v) {
    w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4]; w[4] = v;
    C bx{0, 0};
#pragma unroll
    for (int j = -2; j <= 2; j++) bx = bx + coef(i, j + 2) * w[j + 2];
    if (i >= 2) bx = bx - coef(i, 5) * x2;
    if (i >= 1) bx = bx - coef(i, 6) * x1;
    bx = bx * coef(i, 7);
    sx0 = sx0 + coef(i, 8) * bx;
    sx1 = sx1 + coef(i, 9) * bx;
    x2 = x1;
    x1 = bx;
    return bx;
  };
  C ring[DYW_N];  // ring[k]: the eliminated row r + k
  {
    C v[DYW_N];
#pragma unroll
    for (int k = 0; k < DYW_N; k++) v[k] = ld(src + (size_t)(k + 4) * nc + col);
#pragma unroll
    for (int k = 0; k < DYW_N; k++) ring[k] = fwd(k, v[k]);
  }
  // the window's lowest DYW_B rows, substituted from its top with a zero start
  auto emit_low = [&](int r) {
    C h1{0, 0}, h2{0, 0};
#pragma unroll
    for (int k = DYW_N - 1; k >= 0; k--) {
      const int i = r + k;
      const C h = ring[k] - coef(i, 10) * h1 - coef(i, 11) * h2;
      if (k < DYW_B) st(dst + (size_t)(i + 2) * nc + col, h);
      h2 = h1;
      h1 = h;
    }
  };
  int r = 0;
  for (; r + DYW_N + DYW_B <= mi; r += DYW_B) {
    C v[DYW_B];
#pragma unroll
    for (int k = 0; k < DYW_B; k++) v[k] = ld(src + (size_t)(r + DYW_N + k + 4) * nc + col);
    emit_low(r);
#pragma unroll
    for (int k = 0; k < DYW_W; k++) ring[k] = ring[k + DYW_B];
#pragma unroll
    for (int k = 0; k < DYW_B; k++) ring[DYW_W + k] = fwd(r + DYW_N + k, v[k]);
  }
  // the last, partial block: rem < DYW_B rows are left to eliminate
  const int rem = mi - (r + DYW_N);
  {
    C v[DYW_B];
#pragma unroll
    for (int k = 0; k < DYW_B; k++)
      if (k < rem) v[k] = ld(src + (size_t)(r + DYW_N + k + 4) * nc + col);
    emit_low(r);
#pragma unroll
    for (int k = 0; k < DYW_W; k++) ring[k] = ring[k + DYW_B];
#pragma unroll
    for (int k = 0; k < DYW_B; k++)
      if (k < rem) ring[DYW_W + k] = fwd(r + DYW_N + k, v[k]);
    r += DYW_B;
  }
  // the border rows (src rows m, m + 1 for the stencils at m - 2, m - 1)
  w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4]; w[4] = ld(src + (size_t)(m + 2) * nc + col);
  C r25{0, 0};
#pragma unroll
  for (int j = -2; j <= 2; j++) r25 = r25 + DER(der, m - 2, 1, j) * w[j + 2];
  w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4]; w[4] = ld(src + (size_t)(m + 3) * nc + col);
  C r27{0, 0};
#pragma unroll
  for (int j = -2; j <= 2; j++) r27 = r27 + DER(der, m - 1, 1, j) * w[j + 2];
  const C ph = cexpi(-(D.alfa0 * ix * sx + D.beta0 * iz * sz));
  C b1, b2;
  d0_border(fc, ph, x1, x2, sx0, sx1, r25, r27, b1, b2);
  const C cph = conj(ph);
  const C yb1 = cph * b1, yb2 = cph * b2;
  st(dst + (size_t)(m - 2 + 2) * nc + col, b1);
  st(dst + (size_t)(m - 1 + 2) * nc + col, b2);
  // the rows r .. mi - 1 still in the window: the exact back substitution
  const int top = mi - r;  // DYW_W + rem of them
  C xk1 = b1, xk2 = b2;
#pragma unroll
  for (int k = DYW_N - 1; k >= 0; k--) {
    if (k < top) {
      const int i = r + k;
      const C xk = ring[k] - F0(i, 5) * yb1 - F0(i, 6) * yb2 - F0(i, 3) * xk1 - F0(i, 4) * xk2;
      st(dst + (size_t)(i + 2) * nc + col, xk);
      xk2 = xk1;
      xk1 = xk;
    }
  }
  // the border columns' share of the first rows, every load issued before
  // the first store (a store between two loads of dst made each load wait
  // for the one before: DYW_W round trips one after another)
  C fx[DYW_W];
#pragma unroll
  for (int i = 0; i < DYW_W; i++) fx[i] = ld(dst + (size_t)(i + 2) * nc + col);
#pragma unroll
  for (int i = 0; i < DYW_W; i++)
    st(dst + (size_t)(i + 2) * nc + col, fx[i] + sig[2 * i] * yb1 + sig[2 * i + 1] * yb2);
}

int launch_dy_window(const Dims &D, const double2 *src, double2 *dst, const double *f0rows, const D0Const &fc,
                     const double *der, const double *sig, double sx, double sz, hipStream_t s) {
  if (D.ny - 2 < 2 * DYW_N + DYW_W) return hipErrorInvalidValue;
  dim3 grid((D.ncol + 255) / 256);
  const size_t lds = (size_t)(D.ny - 2) * DYC * sizeof(double);
  if (lds <= 65536)
    hipLaunchKernelGGL(k_dy_window<1>, grid, dim3(256), lds, s, D, fc, f0rows, der, src, dst, sig, sx, sz);
  else
    hipLaunchKernelGGL(k_dy_window<0>, grid, dim3(256), 0, s, D, fc, f0rows, der, src, dst, sig, sx, sz);
  return hipGetLastError();
}

int dy_window_rows() { return DYW_W; }

// ------------------------------------------------------------------ line solves of a column range

// line_solve over the columns col0 .. col0 + ncols - 1 only, on arrays that
// hold just those columns ([ny + 4 rows][ncols], row stride ncols): the
// wavenumbers and the wrap phase are those of the global column col0 + l,
// everything else (the order of the elimination, the border system, the
// back substitution) is that of k_penta (untabled) and k_d0solve (OUT 0), so
// a column comes out with the same bits as through the whole-field solves.
// ws: the workspace of ncols columns (U1, U2, B1, B2, XR of the untabled
// sweep; XR only for D0 and DY).
template <int K>
__global__ void __launch_bounds__(256)
    k_line_cols(Dims D, int col0, int ncols, double lam, const double *__restrict__ der,
                const double *__restrict__ f0, D0Const fc, const double2 *__restrict__ src,
                double2 *__restrict__ dst, PentaWS ws, double sx, double sz) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= ncols) return;
  int ix, iz;
  col_ix_iz(D, col0 + l, ix, iz);
  Dims Ds = D;
  Ds.ncol = ncols;  // the row stride of src, dst and ws
  Ds.ncs = ncols;
  if constexpr (K == KD0 || K == KDY) {
    const size_t nc = ncols;
    const int m = D.ny;
    C x1, x2, sx0, sx1, r25, r27;
    d0_forward<K == KDY ? 1 : 0>(Ds, f0, der, src, ws.XR, l, x1, x2, sx0, sx1, r25, r27);
    const C ph = cexpi(-(D.alfa0 * ix * sx + D.beta0 * iz * sz));
    C b1, b2;
    d0_border(fc, ph, x1, x2, sx0, sx1, r25, r27, b1, b2);
    const C cph = conj(ph);
    const C yb1 = cph * b1, yb2 = cph * b2;
    st(dst + (size_t)(m - 2 + 2) * nc + l, b1);
    st(dst + (size_t)(m - 1 + 2) * nc + l, b2);
    C xk1 = b1, xk2 = b2;
    rows_descending<PF_BWD, C>(
        m - 2, [&](int i) { return ld(ws.XR + (size_t)i * nc + l); },
        [&](int i, const C &xr) {
          const C xk = xr - F0(i, 5) * yb1 - F0(i, 6) * yb2 - F0(i, 3) * xk1 - F0(i, 4) * xk2;
          st(dst + (size_t)(i + 2) * nc + l, xk);
          xk2 = xk1;
          xk1 = xk;
        });
  } else {
    penta_line<K, false>(Ds, l, ix, iz, 0, lam, der, src, dst, ws, nullptr, sx, sz);
  }
}

size_t line_cols_ws_bytes(const Dims &D, int kind, int ncols) {
  const size_t rows = (size_t)(D.ny - 2) * (size_t)ncols;
  return rows * ((kind == KD0 || kind == KDY) ? 16 : 48);
}

int launch_line_cols(const Dims &D, int kind, double lambda, int col0, int ncols, const double2 *src, double2 *dst,
                     void *ws, const double *f0rows, const D0Const &fc, const double *der, double sx, double sz,
                     hipStream_t s) {
  const size_t rows = (size_t)(D.ny - 2) * (size_t)ncols;
  char *w = (char *)ws;
  PentaWS pw;
  if (kind == KD0 || kind == KDY) {
    pw.U1 = pw.U2 = pw.B1 = pw.B2 = nullptr;
    pw.XR = (double2 *)w;
  } else {
    pw.XR = (double2 *)w;  // 16-byte entries first: every array stays 16-byte aligned
    pw.U1 = (double *)(w + rows * 16);
    pw.U2 = (double *)(w + rows * 24);
    pw.B1 = (double *)(w + rows * 32);
    pw.B2 = (double *)(w + rows * 40);
  }
  dim3 grid((ncols + 255) / 256);
#define L(K) hipLaunchKernelGGL((k_line_cols<K>), grid, dim3(256), 0, s, D, col0, ncols, lambda, der, f0rows, fc, src, \
                                dst, pw, sx, sz)
  switch (kind) {
    case KD2V: L(KD2V); break;
    case KETA: L(KETA); break;
    case KPOISSON: L(KPOISSON); break;
    case KD0: L(KD0); break;
    case KDY: L(KDY); break;
    default: return hipErrorInvalidValue;
  }
#undef L
  return hipGetLastError();
}

// ------------------------------------------------------------------ exact shift

// shift_unweighted (src/hst_equations.f90:263-293) of one array per lane:
// x = D0^-1 q at the old shifts, x' = x exp(-i (alfa dgx + beta dgz) y),
// ghosts of x' at the new shifts, q = D0 x'.  The back substitution yields
// x from the top row down, rows 0 and 1 of x come from the record, so the
// re-weighted rows are emitted in the same sweep.
struct QList {
  void *q[4];      // double2, or float2 where f32 is set
  double2 *xr[4];  // per array a workspace of ny - 2 rows of ncol (the checkpoints use 1/4 of it)
  int f32[4];
};

// The node phases of the shift, separable in x and z:
// PX[i][ix] = exp(-i alfa0 ix dgx y_i), PZ[i][iz + nz] = exp(-i beta0 iz dgz y_i).
__global__ void __launch_bounds__(256) k_shift_phases(Dims D, const double *__restrict__ y, double dgx, double dgz,
                                                      double2 *__restrict__ PX, double2 *__restrict__ PZ) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (k < D.NXm) {
    st(PX + (size_t)i * D.NXm + k, cexpi(-(D.alfa0 * k) * dgx * y[i + 2]));
  } else if (k < D.NXm + D.NZm) {
    const int izm = k - D.NXm, iz = izm - D.nz;
    st(PZ + (size_t)i * D.NZm + izm, cexpi(-(D.beta0 * iz) * dgz * y[i + 2]));
  }
}

int launch_shift_phases(const Dims &D, const double *y, double dgx, double dgz, double2 *PX, double2 *PZ,
                        hipStream_t s) {
  dim3 grid((D.NXm + D.NZm + 255) / 256, D.ny);
  hipLaunchKernelGGL(k_shift_phases, grid, dim3(256), 0, s, D, y, dgx, dgz, PX, PZ);
  return hipGetLastError();
}

// The D0 solve's eliminated rows are not stored: the forward sweep keeps the
// recurrence's state (x1, x2) entering every EXS_CHUNK-th row in CK, and the
// back substitution, which runs from the top row down, recomputes each chunk
// of rows from its checkpoint and q before it re-weights them (a chunk's rows
// of q are all read before its re-weighting writes, and the re-weighting of
// row i writes row i + 2 of q, above every row a lower chunk reads).  Per
// array q is read twice and written once and the FP64 checkpoints (32 B per
// chunk) written and read: per row 56 B for an FP64 array against 64 B with
// the streamed workspace, and for an FP32 array (the mixed form's oldrhs) the
// same 32 B as its FP32 workspace.  Alone the kernel is slower than the
// streamed one (it waits on each chunk's loads); beside the fused shift and
// solve of a time step, chunks of 8 rows measured fastest for the whole step
// (FP64 +1.7 %, mixed unchanged; 16 and 32 rows, and loading the next chunk
// ahead, slower).
constexpr int EXS_CHUNK = 8;
template <typename QT>
__device__ __forceinline__ void shift_column(const Dims &D, const D0Const &fc, QT *__restrict__ q,
                                             double4 *__restrict__ CK, int col, const double *__restrict__ f0,
                                             const double *__restrict__ der, const double2 *__restrict__ PX,
                                             const double2 *__restrict__ PZ, double sx0s, double sz0s, double sx1s,
                                             double sz1s) {
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const size_t nc = D.ncol;
  const int m = D.ny, mi = m - 2;
  constexpr int CH = EXS_CHUNK;
  C x1{0, 0}, x2{0, 0}, sx0{0, 0}, sx1{0, 0};
  // rows travel through the prefetch as stored (float2 in the mixed form) and
  // are widened where used: a conversion at the load would wait for it there
  auto wide = [](const QT &r) { return C{(double)r.x, (double)r.y}; };
  rows_ascending<PF_FWD, QT>(
      mi, [&](int i) { return q[(size_t)(i + 2) * nc + col]; },
      [&](int i, const QT &raw) {
        if (i % CH == 0) CK[(size_t)(i / CH) * nc + col] = make_double4(x1.x, x1.y, x2.x, x2.y);
        C bx = wide(raw);
        if (i >= 2) bx = bx - F0(i, 0) * x2;
        if (i >= 1) bx = bx - F0(i, 1) * x1;
        bx = bx * F0(i, 2);
        sx0 = sx0 + F0(i, 7) * bx;
        sx1 = sx1 + F0(i, 8) * bx;
        x2 = x1;
        x1 = bx;
      });
  const C r25 = ld(q + (size_t)(m - 2 + 2) * nc + col), r27 = ld(q + (size_t)(m - 1 + 2) * nc + col);
  const C ph0 = cexpi(-(D.alfa0 * ix * sx0s + D.beta0 * iz * sz0s));
  C b1, b2;
  d0_border(fc, ph0, x1, x2, sx0, sx1, r25, r27, b1, b2);
  const C cph = conj(ph0);
  const C yb1 = cph * b1, yb2 = cph * b2;
  // rows 0 and 1 from the record: x = value - cph (p b1 + q b2) - (t b1 + u b2)
  const double *rr = fc.rec;
  const C xr0 = sx0 - cph * (rr[3] * b1 + rr[4] * b2) - (rr[5] * b1 + rr[6] * b2);
  const C xr1 = sx1 - cph * (rr[9] * b1 + rr[10] * b2) - (rr[11] * b1 + rr[12] * b2);
  auto phase = [&](int i) { return ld(PX + (size_t)i * D.NXm + ix) * ld(PZ + (size_t)i * D.NZm + iz + D.nz); };
  const C ph1 = cexpi(-(D.alfa0 * ix * sx1s + D.beta0 * iz * sz1s));
  const C xp0 = xr0 * phase(0), xp1 = xr1 * phase(1);
  const C xpm2 = b1 * phase(m - 2), xpm1 = b2 * phase(m - 1);
  // window W[k] = x'(i + 1 + k)
  C W0 = xpm2, W1 = xpm1, W2 = xp0 * ph1, W3 = xp1 * ph1;
  C xk1 = b1, xk2 = b2;
  auto body = [&](int i, const C &xr) {
        const C xk = xr - F0(i, 5) * yb1 - F0(i, 6) * yb2 - F0(i, 3) * xk1 - F0(i, 4) * xk2;
        xk2 = xk1;
        xk1 = xk;
        const C xp = xk * phase(i);
        const int r = i + 2;
        C acc{0, 0};
        acc = acc + DER(der, r, 0, -2) * xp;
        acc = acc + DER(der, r, 0, -1) * W0;
        acc = acc + DER(der, r, 0, 0) * W1;
        acc = acc + DER(der, r, 0, 1) * W2;
        acc = acc + DER(der, r, 0, 2) * W3;
        st(q + (size_t)(r + 2) * nc + col, acc);
        W3 = W2; W2 = W1; W1 = W0; W0 = xp;
      };
  for (int c = (mi - 1) / CH; c >= 0; c--) {
    const int c0 = c * CH, n = min(CH, mi - c0);
    // the chunk's rows, loaded without a branch (past the last row: the last
    // row again), as a load under a branch is waited for where the branch ends
    QT raw[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) raw[j] = q[(size_t)(min(c0 + j, mi - 1) + 2) * nc + col];
    const double4 ck = CK[(size_t)c * nc + col];
    C y1{ck.x, ck.y}, y2{ck.z, ck.w};
    C xr[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) {
      if (j < n) {
        const int i = c0 + j;
        C bx = wide(raw[j]);
        if (i >= 2) bx = bx - F0(i, 0) * y2;
        if (i >= 1) bx = bx - F0(i, 1) * y1;
        bx = bx * F0(i, 2);
        xr[j] = bx;
        y2 = y1;
        y1 = bx;
      }
    }
#pragma unroll
    for (int j = CH - 1; j >= 0; j--)
      if (j < n) body(c0 + j, xr[j]);
  }
  // W = x'(0..3); ghosts below: x'(-1) = x'(ny-1) conj(ph1), x'(-2) = x'(ny-2) conj(ph1)
  const C g1 = xpm1 * conj(ph1), g2 = xpm2 * conj(ph1);
  C acc{0, 0};
  acc = acc + DER(der, 1, 0, -2) * g1;
  acc = acc + DER(der, 1, 0, -1) * W0;
  acc = acc + DER(der, 1, 0, 0) * W1;
  acc = acc + DER(der, 1, 0, 1) * W2;
  acc = acc + DER(der, 1, 0, 2) * W3;
  st(q + (size_t)(1 + 2) * nc + col, acc);
  acc = C{0, 0};
  acc = acc + DER(der, 0, 0, -2) * g2;
  acc = acc + DER(der, 0, 0, -1) * g1;
  acc = acc + DER(der, 0, 0, 0) * W0;
  acc = acc + DER(der, 0, 0, 1) * W1;
  acc = acc + DER(der, 0, 0, 2) * W2;
  st(q + (size_t)(0 + 2) * nc + col, acc);
}

__global__ void __launch_bounds__(256)
    k_exact_shift(Dims D, D0Const fc, QList ql, const double *__restrict__ f0, const double *__restrict__ der,
                  const double2 *__restrict__ PX, const double2 *__restrict__ PZ, double sx0s, double sz0s,
                  double sx1s, double sz1s) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  const int k = blockIdx.y;
  // the checkpoints of array k in its XR buffer ((ny-2)/EXS_CHUNK + 1 rows of 32 B)
  double4 *ck = (double4 *)ql.xr[k];
  if (ql.f32[k])
    shift_column<float2>(D, fc, (float2 *)ql.q[k], ck, col, f0, der, PX, PZ, sx0s, sz0s, sx1s, sz1s);
  else
    shift_column<double2>(D, fc, (double2 *)ql.q[k], ck, col, f0, der, PX, PZ, sx0s, sz0s, sx1s, sz1s);
}

int launch_exact_shift(const Dims &D, int nq, void *const *q, const int *f32, double2 *const *XR,
                       const double *f0rows, const D0Const &fc, const double *der, const double2 *PX,
                       const double2 *PZ, double sx0, double sz0, double sx1, double sz1, hipStream_t s) {
  QList ql;
  for (int k = 0; k < nq; k++) { ql.q[k] = q[k]; ql.xr[k] = XR[k]; ql.f32[k] = f32[k]; }
  dim3 grid((D.ncol + 255) / 256, nq);
  hipLaunchKernelGGL(k_exact_shift, grid, dim3(256), 0, s, D, fc, ql, f0rows, der, PX, PZ, sx0, sz0, sx1, sz1);
  return hipGetLastError();
}

// ------------------------------------------------------------------ exact shift in one sweep

// shift_unweighted of one array per lane in one sweep, the D0 solve windowed
// as in k_dy_window: the forward elimination keeps the last SHW_N eliminated
// rows in registers, and every SHW_B rows the back substitution runs down that
// window from a zero start and yields its lowest SHW_B rows of x.  From row
// SHW_W on these are x to rounding and go on at once: x' = x times the node
// phase, and row j of q = D0 x' is written as soon as x'(j + 2) is known.  The
// first SHW_S rows of x wait in the FP64 workspace xs (the first SHW_W lack
// the border terms yb1 s1(i) + yb2 s2(i), known at the end); the top rows take
// the exact back substitution after the border solve; then the rows of q that
// read the ghosts (x'(-2), x'(-1) from the top rows, x'(m), x'(m + 1) from the
// bottom ones) are written.  q is read and written once (the stash: 2 SHW_S
// rows of 16 B), against twice read and once written with the checkpoints of
// k_exact_shift: half the bytes for the mixed form's FP32 oldrhs.
// CL: the D0 taps and F0 columns 0, 1, 2, 7, 8, 3, 4 of each row (SHC doubles)
// read from a copy in LDS (ny <= 683).
constexpr int SHW_B = 4, SHW_W = DYW_W, SHW_N = SHW_B + SHW_W, SHW_S = SHW_W + 4, SHC = 12;

template <typename QT, int CL>
__device__ __forceinline__ void shift_window_column(const Dims &D, const D0Const &fc, QT *__restrict__ q,
                                                    double2 *__restrict__ xs, int col, const double *__restrict__ f0,
                                                    const double *cfs, const double *__restrict__ der,
                                                    const double *__restrict__ sig, const double2 *__restrict__ PX,
                                                    const double2 *__restrict__ PZ, double sx0s, double sz0s,
                                                    double sx1s, double sz1s) {
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const size_t nc = D.ncol;
  const int m = D.ny, mi = m - 2;
  // D0 tap t of row j (j < m); F0 column of row i (i < mi) by its place in the copy
  auto tap = [&](int j, int t) -> double {
    if constexpr (CL) return cfs[j * SHC + t + 2];
    return DER(der, j, 0, t);
  };
  auto f0c = [&](int i, int c) -> double {
    if constexpr (CL) return cfs[i * SHC + c];
    constexpr int fj[12] = {0, 0, 0, 0, 0, 0, 1, 2, 7, 8, 3, 4};
    return F0(i, fj[c]);
  };
  auto wide = [](const QT &r) { return C{(double)r.x, (double)r.y}; };
  auto phase = [&](const C &px, const C &pz) { return px * pz; };
  C x1{0, 0}, x2{0, 0}, sx0{0, 0}, sx1{0, 0};
  auto fwd = [&](int i, const C &v) {
    C bx = v;
    if (i >= 2) bx = bx - f0c(i, 5) * x2;
    if (i >= 1) bx = bx - f0c(i, 6) * x1;
    bx = bx * f0c(i, 7);
    sx0 = sx0 + f0c(i, 8) * bx;
    sx1 = sx1 + f0c(i, 9) * bx;
    x2 = x1;
    x1 = bx;
    return bx;
  };
  C ring[SHW_N];  // ring[k]: the eliminated row r + k
  {
    QT v[SHW_N];
#pragma unroll
    for (int k = 0; k < SHW_N; k++) v[k] = q[(size_t)(k + 2) * nc + col];
#pragma unroll
    for (int k = 0; k < SHW_N; k++) ring[k] = fwd(k, wide(v[k]));
  }
  // R0..R3 = x'(i - 4 .. i - 1) when x'(i) arrives; row i - 2 of q is then complete
  C R0{0, 0}, R1{0, 0}, R2{0, 0}, R3{0, 0};
  auto push = [&](int i, const C &xp) {
    if (i >= SHW_W + 4) {
      const int j = i - 2;
      C acc{0, 0};
      acc = acc + tap(j, -2) * R0;
      acc = acc + tap(j, -1) * R1;
      acc = acc + tap(j, 0) * R2;
      acc = acc + tap(j, 1) * R3;
      acc = acc + tap(j, 2) * xp;
      st(q + (size_t)(j + 2) * nc + col, acc);
    }
    R0 = R1; R1 = R2; R2 = R3; R3 = xp;
  };
  // the window's lowest SHW_B rows, substituted from its top with a zero start
  auto emit = [&](int r, C (&hb)[SHW_B]) {
    C h1{0, 0}, h2{0, 0};
#pragma unroll
    for (int k = SHW_N - 1; k >= 0; k--) {
      const int i = r + k;
      const C h = ring[k] - f0c(i, 10) * h1 - f0c(i, 11) * h2;
      if (k < SHW_B) hb[k] = h;
      h2 = h1;
      h1 = h;
    }
  };
  auto on_rows = [&](int r, const C (&hb)[SHW_B], const C (&px)[SHW_B], const C (&pz)[SHW_B]) {
#pragma unroll
    for (int k = 0; k < SHW_B; k++) {
      const int i = r + k;
      if (i < SHW_S) st(xs + (size_t)i * nc + col, hb[k]);
      if (i >= SHW_W) push(i, hb[k] * phase(px[k], pz[k]));
    }
  };
  int r = 0;
  for (; r + SHW_N + SHW_B <= mi; r += SHW_B) {
    QT v[SHW_B];
    C px[SHW_B], pz[SHW_B], hb[SHW_B];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) v[k] = q[(size_t)(r + SHW_N + k + 2) * nc + col];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) {
      px[k] = ld(PX + (size_t)(r + k) * D.NXm + ix);
      pz[k] = ld(PZ + (size_t)(r + k) * D.NZm + iz + D.nz);
    }
    emit(r, hb);
    on_rows(r, hb, px, pz);
#pragma unroll
    for (int k = 0; k < SHW_W; k++) ring[k] = ring[k + SHW_B];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) ring[SHW_W + k] = fwd(r + SHW_N + k, wide(v[k]));
  }
  // the last, partial block: rem < SHW_B rows are left to eliminate (loads
  // clamped to the last row rather than under a branch)
  const int rem = mi - (r + SHW_N);
  {
    QT v[SHW_B];
    C px[SHW_B], pz[SHW_B], hb[SHW_B];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) v[k] = q[(size_t)(min(r + SHW_N + k, mi - 1) + 2) * nc + col];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) {
      px[k] = ld(PX + (size_t)(r + k) * D.NXm + ix);
      pz[k] = ld(PZ + (size_t)(r + k) * D.NZm + iz + D.nz);
    }
    emit(r, hb);
    on_rows(r, hb, px, pz);
#pragma unroll
    for (int k = 0; k < SHW_W; k++) ring[k] = ring[k + SHW_B];
#pragma unroll
    for (int k = 0; k < SHW_B; k++)
      if (k < rem) ring[SHW_W + k] = fwd(r + SHW_N + k, wide(v[k]));
    r += SHW_B;
  }
  // the border system; q's border rows are not written yet (every row written
  // so far lies below r - 2)
  const C r25 = wide(q[(size_t)(m - 2 + 2) * nc + col]), r27 = wide(q[(size_t)(m - 1 + 2) * nc + col]);
  const C ph0 = cexpi(-(D.alfa0 * ix * sx0s + D.beta0 * iz * sz0s));
  C b1, b2;
  d0_border(fc, ph0, x1, x2, sx0, sx1, r25, r27, b1, b2);
  const C cph = conj(ph0);
  const C yb1 = cph * b1, yb2 = cph * b2;
  // the rows r .. mi - 1 still in the window: the exact back substitution
  const int top = mi - r;  // SHW_W + rem of them
  {
    C xk1 = b1, xk2 = b2;
#pragma unroll
    for (int k = SHW_N - 1; k >= 0; k--) {
      if (k < top) {
        const int i = r + k;
        const C xk = ring[k] - F0(i, 5) * yb1 - F0(i, 6) * yb2 - F0(i, 3) * xk1 - F0(i, 4) * xk2;
        ring[k] = xk;
        xk2 = xk1;
        xk1 = xk;
      }
    }
  }
  // their x', then the border rows' (x(m - 2) = b1, x(m - 1) = b2), in order
#pragma unroll
  for (int k0 = 0; k0 < SHW_N; k0 += SHW_B) {
    C px[SHW_B], pz[SHW_B];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) {
      const int i = min(r + k0 + k, mi - 1);
      px[k] = ld(PX + (size_t)i * D.NXm + ix);
      pz[k] = ld(PZ + (size_t)i * D.NZm + iz + D.nz);
    }
#pragma unroll
    for (int k = 0; k < SHW_B; k++)
      if (k0 + k < top) push(r + k0 + k, ring[k0 + k] * phase(px[k], pz[k]));
  }
  push(m - 2, b1 * phase(ld(PX + (size_t)(m - 2) * D.NXm + ix), ld(PZ + (size_t)(m - 2) * D.NZm + iz + D.nz)));
  push(m - 1, b2 * phase(ld(PX + (size_t)(m - 1) * D.NXm + ix), ld(PZ + (size_t)(m - 1) * D.NZm + iz + D.nz)));
  // R0..R3 = x'(m - 4 .. m - 1)
  const C ph1 = cexpi(-(D.alfa0 * ix * sx1s + D.beta0 * iz * sz1s));
  const C T0 = R0, T1 = R1, T2 = R2, T3 = R3;
  // the first SHW_S rows: x from the stash plus, for the first SHW_W, the border
  // terms; rows 0 .. SHW_W + 1 of q as their x' come, with x'(-2) = x'(m - 2)
  // conj(ph1) and x'(-1) = x'(m - 1) conj(ph1) ahead of them in the window
  R2 = T2 * conj(ph1);
  R3 = T3 * conj(ph1);
  C f0x{0, 0}, f1x{0, 0};
#pragma unroll
  for (int i0 = 0; i0 < SHW_S; i0 += SHW_B) {
    C z[SHW_B], px[SHW_B], pz[SHW_B];
#pragma unroll
    for (int k = 0; k < SHW_B; k++) {
      const int i = i0 + k;
      z[k] = ld(xs + (size_t)i * nc + col);
      px[k] = ld(PX + (size_t)i * D.NXm + ix);
      pz[k] = ld(PZ + (size_t)i * D.NZm + iz + D.nz);
    }
#pragma unroll
    for (int k = 0; k < SHW_B; k++) {
      const int i = i0 + k;
      C x = z[k];
      if (i < SHW_W) x = x + sig[2 * i] * yb1 + sig[2 * i + 1] * yb2;
      const C xp = x * phase(px[k], pz[k]);
      if (i == 0) f0x = xp;
      if (i == 1) f1x = xp;
      if (i >= 2) {
        const int j = i - 2;
        C acc{0, 0};
        acc = acc + tap(j, -2) * R0;
        acc = acc + tap(j, -1) * R1;
        acc = acc + tap(j, 0) * R2;
        acc = acc + tap(j, 1) * R3;
        acc = acc + tap(j, 2) * xp;
        st(q + (size_t)(j + 2) * nc + col, acc);
      }
      R0 = R1; R1 = R2; R2 = R3; R3 = xp;
    }
  }
  // rows m - 2, m - 1, with x'(m) = x'(0) ph1, x'(m + 1) = x'(1) ph1
  const C gp0 = f0x * ph1, gp1 = f1x * ph1;
  {
    C acc{0, 0};
    acc = acc + tap(m - 2, -2) * T0;
    acc = acc + tap(m - 2, -1) * T1;
    acc = acc + tap(m - 2, 0) * T2;
    acc = acc + tap(m - 2, 1) * T3;
    acc = acc + tap(m - 2, 2) * gp0;
    st(q + (size_t)(m - 2 + 2) * nc + col, acc);
    acc = C{0, 0};
    acc = acc + tap(m - 1, -2) * T1;
    acc = acc + tap(m - 1, -1) * T2;
    acc = acc + tap(m - 1, 0) * T3;
    acc = acc + tap(m - 1, 1) * gp0;
    acc = acc + tap(m - 1, 2) * gp1;
    st(q + (size_t)(m - 1 + 2) * nc + col, acc);
  }
}

template <int CL>
__global__ void __launch_bounds__(256, 2)
    k_shift_window(Dims D, D0Const fc, QList ql, const double *__restrict__ f0, const double *__restrict__ der,
                   const double *__restrict__ sig, const double2 *__restrict__ PX, const double2 *__restrict__ PZ,
                   double sx0s, double sz0s, double sx1s, double sz1s) {
  const int m = D.ny, mi = m - 2;
  extern __shared__ double cfs[];
  if constexpr (CL) {
    // [row j < m][SHC]: D0 taps -2..2, then F0 columns 0, 1, 2, 7, 8, 3, 4 (rows < mi)
    const int n = m * SHC;
    for (int b0 = 0; b0 < n; b0 += 16 * 256) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int k = min(b0 + u * 256 + (int)threadIdx.x, n - 1), j = k / SHC, c = k - j * SHC;
        const int fj = c == 5 ? 0 : (c == 6 ? 1 : (c == 7 ? 2 : (c == 8 ? 7 : (c == 9 ? 8 : (c == 10 ? 3 : 4)))));
        v[u] = c < 5 ? DER(der, j, 0, c - 2) : (j < mi ? F0(j, fj) : 0.0);
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int k = b0 + u * 256 + threadIdx.x;
        if (k < n) cfs[k] = v[u];
      }
    }
    __syncthreads();
  }
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  const int k = blockIdx.y;
  if (ql.f32[k])
    shift_window_column<float2, CL>(D, fc, (float2 *)ql.q[k], ql.xr[k], col, f0, cfs, der, sig, PX, PZ, sx0s, sz0s,
                                    sx1s, sz1s);
  else
    shift_window_column<double2, CL>(D, fc, (double2 *)ql.q[k], ql.xr[k], col, f0, cfs, der, sig, PX, PZ, sx0s,
                                     sz0s, sx1s, sz1s);
}

int launch_shift_window(const Dims &D, int nq, void *const *q, const int *f32, double2 *const *XR,
                        const double *f0rows, const D0Const &fc, const double *der, const double *sig,
                        const double2 *PX, const double2 *PZ, double sx0, double sz0, double sx1, double sz1,
                        hipStream_t s) {
  if (D.ny - 2 < 2 * SHW_N + SHW_W || nq < 1 || nq > 4) return hipErrorInvalidValue;
  QList ql;
  for (int k = 0; k < nq; k++) { ql.q[k] = q[k]; ql.xr[k] = XR[k]; ql.f32[k] = f32[k]; }
  dim3 grid((D.ncol + 255) / 256, nq);
  const size_t lds = (size_t)D.ny * SHC * sizeof(double);
  if (lds <= 65536)
    hipLaunchKernelGGL(k_shift_window<1>, grid, dim3(256), lds, s, D, fc, ql, f0rows, der, sig, PX, PZ, sx0, sz0,
                       sx1, sz1);
  else
    hipLaunchKernelGGL(k_shift_window<0>, grid, dim3(256), 0, s, D, fc, ql, f0rows, der, sig, PX, PZ, sx0, sz0, sx1,
                       sz1);
  return hipGetLastError();
}

// ------------------------------------------------------------------ exact shift + line solve

// The exact shift of one right-hand side (shift_unweighted,
// src/hst_equations.f90:263-293) and its line solve (line_solve,
// src/hst_linsolve.f90:282-483) in one pass per column, with the shifted
// right-hand side handed from one to the other in registers.  The D0 solve
// of the shift runs over the rows in reverse order (the reversed line is the
// same cyclic system with the stencils mirrored and the wrap phase
// conjugated: f0r, fcr), so that its back substitution yields x in ascending
// rows; the re-weighted rows q = D0 x' (x' = x times the node phase, ghosts at
// the new shifts) then come out ascending, the order the line solve's
// elimination takes them.  Rows 0, 1 of x come from the reversed border
// system, rows ny-2, ny-1 first from the record (for the ghost rows below row
// 0), then from the back substitution.
// Loads two rows ahead in each sweep, the phase factors of the second sweep
// with them (multiplied only where used: 5 % faster than carrying their
// product, which made each prefetch wait for its loads).  Three wavefronts
// per SIMD (168 VGPRs).  Deeper prefetch ran slower.
// One system alone (half the launch) takes 0.48 of the time of both: the
// kernel is bound by HBM bandwidth, not by the second, partly filled round.
constexpr int SP_WAVES = 3, SP_PF = 2;
#ifndef SP_LDS
#define SP_LDS 1
#endif
template <int K>
__device__ __forceinline__ void shift_penta_line(const Dims &D, int col, int ix, int iz, int g,
                                                 const D0Const &fcr,
                                                 const double *__restrict__ f0, const double *__restrict__ der,
                                                 const double2 *__restrict__ q, double2 *__restrict__ XRs,
                                                 const double2 *__restrict__ PX, const double2 *__restrict__ PZ,
                                                 double sx0s, double sz0s, double sx1s, double sz1s, double lam,
                                                 double2 *__restrict__ XRp, const double4 *__restrict__ tab,
                                                 double2 *__restrict__ dst) {
  const size_t nc = D.ncol, ws = D.ncs;  // field rows, workspace rows
  const int m = D.ny, mi = m - 2;
  if (ix == 0 && iz == 0 && K == KD2V) {  // the (0,0) D2V line is zero (src/hst_linsolve.f90:199,221-224)
    for (int i = 0; i < m; i++) st(dst + (size_t)(i + 2) * nc + col, C{0.0, 0.0});
    return;
  }
  // 1. the reversed D0 elimination: reversed row k is row m-1-k
  C x1{0, 0}, x2{0, 0}, sx0{0, 0}, sx1{0, 0};
  rows_ascending<SP_PF, C>(
      mi, [&](int k) { return ld(q + (size_t)(m - 1 - k + 2) * nc + col); },
      [&](int k, const C &v) {
        C bx = v;
        if (k >= 2) bx = bx - F0(k, 0) * x2;
        if (k >= 1) bx = bx - F0(k, 1) * x1;
        bx = bx * F0(k, 2);
        st(XRs + (size_t)k * ws + col, bx);
        sx0 = sx0 + F0(k, 7) * bx;
        sx1 = sx1 + F0(k, 8) * bx;
        x2 = x1;
        x1 = bx;
      });
  const C r25 = ld(q + (size_t)(1 + 2) * nc + col), r27 = ld(q + (size_t)(0 + 2) * nc + col);
  const C ph0 = cexpi(-(D.alfa0 * ix * sx0s + D.beta0 * iz * sz0s));
  C b1, b2;  // reversed rows m-2, m-1: x(1), x(0)
  d0_border(fcr, conj(ph0), x1, x2, sx0, sx1, r25, r27, b1, b2);
  const C yb1 = ph0 * b1, yb2 = ph0 * b2;
  const double *rr = fcr.rec;
  const C xr0 = sx0 - ph0 * (rr[3] * b1 + rr[4] * b2) - (rr[5] * b1 + rr[6] * b2);    // x(m-1)
  const C xr1 = sx1 - ph0 * (rr[9] * b1 + rr[10] * b2) - (rr[11] * b1 + rr[12] * b2);  // x(m-2)
  // the z phase of -k is the conjugate of that of k: both halves of a
  // wavefront load the same row of PZ
  const int az = iz < 0 ? -iz : iz;
  auto phase = [&](int i) {
    const C pz = ld(PZ + (size_t)i * D.NZm + az + D.nz);
    return ld(PX + (size_t)i * D.NXm + ix) * (iz < 0 ? conj(pz) : pz);
  };
  const C ph1 = cexpi(-(D.alfa0 * ix * sx1s + D.beta0 * iz * sz1s));
  const C xp0 = b2 * phase(0), xp1 = b1 * phase(1);
  // 2. the line solve's elimination, fed row by row
  const double al = D.alfa0 * ix, be = D.beta0 * iz;
  const double kk = al * al + be * be, ni = D.ni;
  PentaFactor f;
  C px1{0, 0}, px2{0, 0}, psx0{0, 0}, psx1{0, 0}, r25p{0, 0}, r27p{0, 0};
  // row i (< mi) of the elimination
  auto feed = [&](int i, const C &bin) {
    double l2, l1, rp, t1, t2, p, qq;
    f.step<K>(i, lam, ni, kk, der, l2, l1, rp, t1, t2, p, qq);
    C bx = bin;
    if (i >= 2) bx = bx - l2 * px2;
    if (i >= 1) bx = bx - l1 * px1;
    bx = bx * rp;
    st(XRp + (size_t)i * ws + col, bx);
    psx0 = psx0 + f.ww * bx;
    psx1 = psx1 + f.vv * bx;
    px2 = px1;
    px1 = bx;
  };
  // window W = x'(i-2 .. i+1) for the row i about to be emitted; below row 0
  // the ghosts x'(-k) = x'(m-k) conj(ph1)
  C W0 = xr1 * phase(m - 2) * conj(ph1), W1 = xr0 * phase(m - 1) * conj(ph1), W2 = xp0, W3 = xp1;
  auto reweight = [&](int i, const C &xnext) {  // row i of D0 x'; xnext = x'(i + 2)
    C acc{0, 0};
    acc = acc + DER(der, i, 0, -2) * W0;
    acc = acc + DER(der, i, 0, -1) * W1;
    acc = acc + DER(der, i, 0, 0) * W2;
    acc = acc + DER(der, i, 0, 1) * W3;
    acc = acc + DER(der, i, 0, 2) * xnext;
    W0 = W1; W1 = W2; W2 = W3; W3 = xnext;
    return acc;
  };
  // the row's phase factors travel with it unmultiplied: a product formed at
  // load time would wait for them there, and the loads issued ahead would be
  // waited for one row after another
  struct Row {
    C xr, px, pz;
  };
  C xk1 = b1, xk2 = b2;
  rows_descending<SP_PF, Row>(
      mi,
      [&](int k) {
        const int i = m - 1 - k;
        return Row{ld(XRs + (size_t)k * ws + col), ld(PX + (size_t)i * D.NXm + ix),
                   ld(PZ + (size_t)i * D.NZm + az + D.nz)};
      },
      [&](int k, const Row &r) {
        const C xk = r.xr - F0(k, 5) * yb1 - F0(k, 6) * yb2 - F0(k, 3) * xk1 - F0(k, 4) * xk2;
        xk2 = xk1;
        xk1 = xk;
        const int i = m - 1 - k - 2;  // x'(m-1-k) completes row m-3-k < mi
        feed(i, reweight(i, xk * (r.px * (iz < 0 ? conj(r.pz) : r.pz))));
      });
  // the border rows' right-hand sides (the loop above covers the rows < mi,
  // so that its body has no branch on the row)
  r25p = reweight(m - 2, xp0 * ph1);  // x'(m) = x'(0) ph1
  r27p = reweight(m - 1, xp1 * ph1);  // x'(m+1) = x'(1) ph1
  // 3. the line solve's border system and back substitution (as penta_line)
  double rq[25];
  rq[3] = f.sp0; rq[4] = f.sq0; rq[5] = f.w2 * f.u22 + f.w1 * f.u11; rq[6] = f.w1 * f.u12;
  rq[9] = f.sp1; rq[10] = f.sq1; rq[11] = f.v2 * f.u22 + f.v1 * f.u11; rq[12] = f.v1 * f.u12;
  rq[15] = f.p2 - f.u21 * f.p1; rq[16] = f.q2 - f.u21 * f.q1; rq[17] = f.u22 - f.u21 * f.u11; rq[18] = -f.u21 * f.u12;
  rq[21] = f.p1; rq[22] = f.q1; rq[23] = f.u11; rq[24] = f.u12;
  const C xm2 = px2 - f.u21 * px1, xm1 = px1;
  const int ga = m - 2;
  C c1, c2;
  border_solve(rq, coefk<K>(lam, ni, kk, der, ga, -2), coefk<K>(lam, ni, kk, der, ga, -1),
               coefk<K>(lam, ni, kk, der, ga, 0), coefk<K>(lam, ni, kk, der, ga, 1),
               coefk<K>(lam, ni, kk, der, ga, 2), coefk<K>(lam, ni, kk, der, ga + 1, -2),
               coefk<K>(lam, ni, kk, der, ga + 1, -1), coefk<K>(lam, ni, kk, der, ga + 1, 0),
               coefk<K>(lam, ni, kk, der, ga + 1, 1), coefk<K>(lam, ni, kk, der, ga + 1, 2), ph1, psx0, psx1, xm2,
               xm1, r25p, r27p, c1, c2);
  st(dst + (size_t)(m - 2 + 2) * nc + col, c1);
  st(dst + (size_t)(m - 1 + 2) * nc + col, c2);
  const C cph1 = conj(ph1);
  const C zb1 = cph1 * c1, zb2 = cph1 * c2;
  penta_back<K>(
      D, g, lam, kk, der, tab, zb1, zb2, c1, c2, [&](int i) { return ld(XRp + (size_t)i * ws + col); },
      [&](int i, const C &x) { st(dst + (size_t)(i + 2) * nc + col, x); });
}

struct ShiftPentaArgs {
  const double2 *q[2];
  double2 *xrs[2], *xrp[2], *dst[2];
  const double4 *tab[2];
};

// blockIdx.y: system 0 (D2V on rhs2) or 1 (ETA on rhs1)
// CL: the reversed D0 factor rows read from a copy in LDS (9 doubles a row)
template <int CL>
__global__ void __launch_bounds__(256, SP_WAVES)
    k_shift_penta(Dims D, ShiftPentaArgs a, D0Const fcr, const double *__restrict__ f0g,
                  const double *__restrict__ der, const double2 *__restrict__ PX, const double2 *__restrict__ PZ,
                  double sx0s, double sz0s, double sx1s, double sz1s, double lam) {
  extern __shared__ double f0s[];
  if constexpr (CL) {
    const int n = (D.ny - 2) * 9;
    for (int b0 = 0; b0 < n; b0 += 16 * 256) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = f0g[min(b0 + u * 256 + (int)threadIdx.x, n - 1)];
#pragma unroll
      for (int u = 0; u < 16; u++)
        if (b0 + u * 256 + (int)threadIdx.x < n) f0s[b0 + u * 256 + threadIdx.x] = v[u];
    }
    __syncthreads();
  }
  const double *const f0r = CL ? f0s : f0g;
  // one column per lane in column order: a wavefront's row is one run of 64
  // columns (the +iz, -iz pairs of pair_lane share the table rows but read two
  // runs a row, measured 1.5 % slower here)
  const int l = xcd_linear();
  const int sys = l / gridDim.x;
  const int col = (l % gridDim.x) * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const int g = penta_group(D, ix, iz);
  if (sys == 0)
    shift_penta_line<KD2V>(D, col, ix, iz, g, fcr, f0r, der, a.q[0], a.xrs[0], PX, PZ, sx0s, sz0s, sx1s, sz1s, lam, a.xrp[0],
                           a.tab[0], a.dst[0]);
  else
    shift_penta_line<KETA>(D, col, ix, iz, g, fcr, f0r, der, a.q[1], a.xrs[1], PX, PZ, sx0s, sz0s, sx1s, sz1s, lam, a.xrp[1],
                           a.tab[1], a.dst[1]);
}

int launch_shift_penta(const Dims &D, const double2 *q2, const double2 *q1, double2 *const xrs[2],
                       double2 *const xrp[2], const void *const tabs[2], double2 *v2, double2 *eta,
                       const double *f0r, const D0Const &fcr, const double *der, const double2 *PX,
                       const double2 *PZ, double sx0, double sz0, double sx1, double sz1, double lambda,
                       hipStream_t s) {
  ShiftPentaArgs a;
  a.q[0] = q2; a.q[1] = q1;
  a.dst[0] = v2; a.dst[1] = eta;
  for (int k = 0; k < 2; k++) {
    a.xrs[k] = xrs[k];
    a.xrp[k] = xrp[k];
    a.tab[k] = (const double4 *)tabs[k];
  }
  dim3 grid((D.ncol + 255) / 256, 2);
  const size_t lds = (size_t)(D.ny - 2) * 9 * sizeof(double);
  if (SP_LDS && lds <= 49152)
    hipLaunchKernelGGL(k_shift_penta<1>, grid, dim3(256), lds, s, D, a, fcr, f0r, der, PX, PZ, sx0, sz0, sx1, sz1,
                       lambda);
  else
    hipLaunchKernelGGL(k_shift_penta<0>, grid, dim3(256), 0, s, D, a, fcr, f0r, der, PX, PZ, sx0, sz0, sx1, sz1,
                       lambda);
  return hipGetLastError();
}

// ------------------------------------------------------------------ ghosts

struct FList {
  double2 *f[8];
};

// wrap_rows (src/hst_mpi.f90:340-355): f(ny) = f(0) ph, f(ny+1) = f(1) ph,
// f(-1) = f(ny-1) conj(ph), f(-2) = f(ny-2) conj(ph).
__global__ void __launch_bounds__(256) k_ghosts(Dims D, FList fl, double sx, double sz) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  double2 *f = fl.f[blockIdx.y];
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const C ph = cexpi(-(D.alfa0 * ix * sx + D.beta0 * iz * sz));
  const size_t nc = D.ncol;
  const int ny = D.ny;
  const C a0 = ld(f + (size_t)2 * nc + col), a1 = ld(f + (size_t)3 * nc + col);
  const C bm1 = ld(f + (size_t)(ny + 1) * nc + col), bm2 = ld(f + (size_t)ny * nc + col);
  st(f + (size_t)(ny + 2) * nc + col, a0 * ph);
  st(f + (size_t)(ny + 3) * nc + col, a1 * ph);
  st(f + (size_t)1 * nc + col, bm1 * conj(ph));
  st(f + (size_t)0 * nc + col, bm2 * conj(ph));
}

int launch_ghosts(const Dims &D, int nf, double2 *const *f, double sx, double sz, hipStream_t s) {
  FList fl;
  for (int k = 0; k < nf; k++) fl.f[k] = f[k];
  dim3 grid((D.ncol + 255) / 256, nf);
  hipLaunchKernelGGL(k_ghosts, grid, dim3(256), 0, s, D, fl, sx, sz);
  return hipGetLastError();
}

// ------------------------------------------------------------------ statistics

// The first reduction of outstats (src/hst_stats.f90:153-198): per block the
// NSTAT sums uu vv ww uv vw uw eps mfx mfz q_in q_out e_in e_out ed pd
// ncol_damped; the column energies of the spectrum on the way (:224-238).
__global__ void __launch_bounds__(256)
    k_stats(Dims D, StatsArgs a, const double2 *__restrict__ V1, const double2 *__restrict__ V2,
            const double2 *__restrict__ V3, const double *__restrict__ dyl, double *__restrict__ part,
            double *__restrict__ exz) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  double s[NSTAT];
#pragma unroll
  for (int k = 0; k < NSTAT; k++) s[k] = 0.0;
  if (col < D.ncol) {
    int ix, iz;
    col_ix_iz(D, col, ix, iz);
    const double al = D.alfa0 * ix, be = D.beta0 * iz, kk = al * al + be * be;
    const size_t nc = D.ncol;
    const double gd = a.gdamp ? a.gdamp[col] : 0.0;
    if (gd > 0.0) s[15] = 1.0;
    // the column energy sum_y w |u|^2 of the spectrum, every column
    double qc = 0.0;
    for (int iy = 0; iy < D.ny; iy++) {
      const size_t o = (size_t)(iy + 2) * nc + col;
      const C cu = ld(V1 + o), cv = ld(V2 + o), cw = ld(V3 + o);
      const double dy = dyl[iy];
      const double uu = cu.x * cu.x + cu.y * cu.y, vv = cv.x * cv.x + cv.y * cv.y, wv = cw.x * cw.x + cw.y * cw.y;
      qc = qc + (ix == 0 ? dy : 2.0 * dy) * (uu + vv + wv);
      if (ix == 0 && iz == 0) {
        s[7] += cu.x * dy;
        s[8] += cw.x * dy;
        if (gd > 0.0) {  // damped mean profiles (damp_mean)
          const double q = dy * (uu + wv);
          s[13] += q;
          s[14] += 2.0 * a.gsoft * gd * q;
        }
        if (a.inc00) {
          s[0] += dy * uu;
          s[2] += dy * wv;
          s[5] += dy * (cu.x * cw.x + cu.y * cw.y);
        }
        continue;
      }
      const double w = ix == 0 ? dy : 2.0 * dy;
      const double q = w * (uu + vv + wv);
      if (gd > 0.0) {
        s[13] += q;
        s[14] += 2.0 * a.gsoft * gd * q;
      }
      s[0] += w * uu;
      s[1] += w * vv;
      s[2] += w * wv;
      s[3] += w * (cu.x * cv.x + cu.y * cv.y);
      s[4] += w * (cw.x * cv.x + cw.y * cv.y);
      s[5] += w * (cu.x * cw.x + cu.y * cw.y);
      s[6] += w * kk * (uu + vv + wv);
      if (a.inlayer && a.inlayer[iy]) {
        s[9] += q;
        s[11] += w * kk * (uu + vv + wv);
      } else {
        s[10] += q;
        s[12] += w * kk * (uu + vv + wv);
      }
    }
    if (exz) exz[col] += qc;
  }
  __shared__ double red[4][NSTAT];
#pragma unroll
  for (int k = 0; k < NSTAT; k++) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NSTAT) {
    const int k = threadIdx.x;
    part[(size_t)blockIdx.x * NSTAT + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
  }
}

int launch_stats(const Dims &D, const StatsArgs &a, const double2 *V1, const double2 *V2, const double2 *V3,
                 const double *dyl, double *part, int *n_part, double *exz_acc, hipStream_t s) {
  dim3 grid((D.ncol + 255) / 256);
  if (n_part) *n_part = grid.x;
  hipLaunchKernelGGL(k_stats, grid, dim3(256), 0, s, D, a, V1, V2, V3, dyl, part, exz_acc);
  return hipGetLastError();
}

// ------------------------------------------------------------------ physics hooks

// The layer profile W and body force f (src/hst_stokes.f90:63-88) at y, t.
__device__ __forceinline__ double stokes_s(const StokesParams &p, double yy) {
  const double e = (yy - 0.5 * p.ly) / p.delta;
  return sqrt(e * e + 0.1 * 0.1);  // eps = 0.1: the smoothed |eta| / delta
}

__global__ void k_stokes_force(StokesParams p, int ny, const double *__restrict__ y, double t, double ramp,
                               double *__restrict__ fy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;  // row + 2
  if (r >= ny + 4) return;
  const int iy = ((r - 2) % ny + ny) % ny;              // ghosts: the periodic copies
  const double yy = y[iy + 2];
  const double omega = 6.283185307179586 / p.period;
  const double eta = yy - 0.5 * p.ly;
  const double sv = stokes_s(p, yy);
  const double s1 = (eta / (p.delta * p.delta)) / sv;
  const double s2 = ((0.1 * 0.1) / (p.delta * p.delta)) / (sv * sv * sv);
  const double th = omega * (t - p.start) - sv;
  double sn, cs;
  sincos(th, &sn, &cs);
  fy[r] = ramp * p.amp * exp(-sv) * ((2.0 * p.ni * s1 * s1 - omega) * sn - p.ni * s2 * (sn - cs));
}

int launch_stokes_force(const StokesParams &p, int ny, const double *y, double t, double ramp, double *fy,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_stokes_force, dim3((ny + 4 + 255) / 256), dim3(256), 0, s, p, ny, y, t, ramp, fy);
  return hipGetLastError();
}

// One block: the interior rows of the (0,0) column of w, then its ghosts.
__global__ void __launch_bounds__(256)
    k_stokes_apply(Dims D, StokesParams p, const double *__restrict__ y, double t, double ramp, int whole,
                   double2 *__restrict__ V3) {
  const size_t c00 = (size_t)D.nz, nc = D.ncol;
  const int ny = D.ny;
  const double omega = 6.283185307179586 / p.period;
  for (int iy = threadIdx.x; iy < ny; iy += blockDim.x) {
    if (whole || iy == 0 || iy == ny - 1) {
      const double sv = stokes_s(p, y[iy + 2]);
      V3[(size_t)(iy + 2) * nc + c00] = make_double2(ramp * p.amp * exp(-sv) * cos(omega * (t - p.start) - sv), 0.0);
    }
  }
  __syncthreads();
  if (threadIdx.x < 4) {  // rows -2, -1, ny, ny+1 from ny-2, ny-1, 0, 1 (the phase of the (0,0) column is 1)
    const int g = threadIdx.x < 2 ? (int)threadIdx.x - 2 : ny + (int)threadIdx.x - 2;
    const int src = threadIdx.x < 2 ? ny + g : g - ny;
    V3[(size_t)(g + 2) * nc + c00] = V3[(size_t)(src + 2) * nc + c00];
  }
}

int launch_stokes_apply(const Dims &D, const StokesParams &p, const double *y, double t, double ramp, int whole,
                        double2 *V3, hipStream_t s) {
  hipLaunchKernelGGL(k_stokes_apply, dim3(1), dim3(256), 0, s, D, p, y, t, ramp, whole, V3);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
    k_damping_table(Dims D, DampParams p, double c, double sn, double *__restrict__ gdamp) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const double kx = D.alfa0 * ix, kz = D.beta0 * iz;
  const double kpar = kx * c + kz * sn, kperp = -kx * sn + kz * c;
  const double r = pow(pow(fabs(kpar) / p.kpar, p.norm) + pow(fabs(kperp) / p.kperp, p.norm), 1.0 / p.norm);
  double g = 0.0;
  if (r < 1.0 - 1.0e-6) {
    if (p.hard) {
      g = 1.0;
    } else if (r <= 1.0 - p.edge) {
      g = p.rate;
    } else {
      const double x = (1.0 - r) / p.edge;
      g = p.rate * x * x * (3.0 - 2.0 * x);
    }
  }
  if (ix == 0 && iz == 0 && !p.mean) g = 0.0;
  gdamp[col] = g;
}

int launch_damping_table(const Dims &D, const DampParams &p, double theta, double *gdamp, hipStream_t s) {
  hipLaunchKernelGGL(k_damping_table, dim3((D.ncol + 255) / 256), dim3(256), 0, s, D, p, cos(theta), sin(theta),
                     gdamp);
  return hipGetLastError();
}

template <typename OT>
__global__ void __launch_bounds__(256)
    k_damping_mask(Dims D, const double *__restrict__ gdamp, double2 *R1, double2 *R2, OT *O1, OT *O2) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol || !(gdamp[col] > 0.0)) return;
  const size_t nc = D.ncol;
  for (int iy = blockIdx.y; iy < D.ny; iy += gridDim.y) {
    const size_t o = (size_t)(iy + 2) * nc + col;
    R1[o] = make_double2(0.0, 0.0);
    R2[o] = make_double2(0.0, 0.0);
    O1[o] = OT{0, 0};
    O2[o] = OT{0, 0};
  }
}

int launch_damping_mask(const Dims &D, int mixed, const double *gdamp, double2 *R1, double2 *R2, void *O1, void *O2,
                        hipStream_t s) {
  dim3 grid((D.ncol + 255) / 256, std::min(D.ny, 64));
  if (mixed)
    hipLaunchKernelGGL(k_damping_mask<float2>, grid, dim3(256), 0, s, D, gdamp, R1, R2, (float2 *)O1, (float2 *)O2);
  else
    hipLaunchKernelGGL(k_damping_mask<double2>, grid, dim3(256), 0, s, D, gdamp, R1, R2, (double2 *)O1,
                       (double2 *)O2);
  return hipGetLastError();
}

// ------------------------------------------------------------------ pressure

// The right-hand side of the pressure equation of one column (compute_pressure,
// src/hst_pressure.f90:35-88):
//   D0 [alfa^2 uu + beta^2 ww + 2 alfa beta uw] - D2 vv - 2 i alfa D1 uv
//   - 2 i beta D1 vw - 2 i (S alfa + S2 beta) D0 v,
// the (0,0) column -<vv>, the raw product.  Stencils of the products in PR.
template <typename PR>
__global__ void __launch_bounds__(256)
    k_pressure_rhs(Dims D, const typename PTraits<PR>::T2 *__restrict__ P0, size_t pstride,
                   const double2 *__restrict__ V2, double2 *__restrict__ R, const double *__restrict__ der,
                   double s2now) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const double al = D.alfa0 * ix, be = D.beta0 * iz;
  const bool m00 = (ix == 0 && iz == 0);
  const size_t nc = D.ncol;
  const int iy = blockIdx.y;
  auto sten = [&](int q, int n) {
    Cx<PR> a{0, 0};
#pragma unroll
    for (int j = -2; j <= 2; j++) a = a + (PR)DER(der, iy, n, j) * ldp<PR>(P0 + q * pstride + (size_t)(iy + j + 2) * nc + col);
    return C{(double)a.x, (double)a.y};
  };
  C d0v{0, 0};
#pragma unroll
  for (int j = -2; j <= 2; j++) d0v = d0v + DER(der, iy, 0, j) * ld(V2 + (size_t)(iy + j + 2) * nc + col);
  C r = (al * al) * sten(0, 0) - imul(2.0 * (D.S * al + s2now * be), d0v);
  r = r - sten(1, 2);
  if (m00) {
    const Cx<PR> vv = ldp<PR>(P0 + 1 * pstride + (size_t)(iy + 2) * nc + col);
    r = C{-(double)vv.x, 0.0};
  }
  r = r + (be * be) * sten(2, 0);
  r = r - imul(2.0 * al, sten(3, 1));
  r = r - imul(2.0 * be, sten(4, 1));
  r = r + (2.0 * al * be) * sten(5, 0);
  st(R + (size_t)(iy + 2) * nc + col, r);
}

int launch_pressure_rhs(const Dims &D, int mixed, const void *P0, const double2 *V2, double2 *R, const double *der,
                        double s2now, hipStream_t s) {
  dim3 grid((D.ncol + 255) / 256, D.ny);
  const size_t pstride = (size_t)D.NYg * D.ncol;
  if (mixed)
    hipLaunchKernelGGL(k_pressure_rhs<float>, grid, dim3(256), 0, s, D, (const float2 *)P0, pstride, V2, R, der, s2now);
  else
    hipLaunchKernelGGL(k_pressure_rhs<double>, grid, dim3(256), 0, s, D, (const double2 *)P0, pstride, V2, R, der,
                       s2now);
  return hipGetLastError();
}

// The singular (0,0) mode (src/hst_pressure.f90:94-111): p00 = the right-hand
// side itself, minus its y mean sum(Re p00 dyl)/ly; one block, a fixed order.
__global__ void __launch_bounds__(256)
    k_pressure_mean(Dims D, const double2 *__restrict__ R, double2 *__restrict__ p, const double *__restrict__ dyl,
                    double ly) {
  __shared__ double red[256];
  const size_t c00 = (size_t)D.nz, nc = D.ncol;
  double a = 0.0;
  for (int iy = threadIdx.x; iy < D.ny; iy += 256) a += R[(size_t)(iy + 2) * nc + c00].x * dyl[iy];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double pmean = red[0] / ly;
  for (int iy = threadIdx.x; iy < D.ny; iy += 256)
    p[(size_t)(iy + 2) * nc + c00] = make_double2(R[(size_t)(iy + 2) * nc + c00].x - pmean, 0.0);
}

int launch_pressure_mean(const Dims &D, const double2 *R, double2 *p, const double *dyl, double ly, hipStream_t s) {
  hipLaunchKernelGGL(k_pressure_mean, dim3(1), dim3(256), 0, s, D, R, p, dyl, ly);
  return hipGetLastError();
}

// ------------------------------------------------------------------ restart repair

// project_restart (src/hst_io.f90:213-250) of one column per lane: the
// discrete divergence i alfa D0 u + D1 v + i beta D0 w of each interior row
// taken off u (ix > 0) or w (ix = 0, iz != 0), v = 0 in the (0,0) column.
// Every row's divergence is that of the field before the repair: the lane
// keeps the original rows in a five-row window while it overwrites its column.
__global__ void __launch_bounds__(256)
    k_project(Dims D, double2 *__restrict__ V1, double2 *__restrict__ V2, double2 *__restrict__ V3,
              const double *__restrict__ der) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.ncol) return;
  int ix, iz;
  col_ix_iz(D, col, ix, iz);
  const size_t nc = D.ncol;
  const double al = D.alfa0 * ix, be = D.beta0 * iz;
  if (ix == 0 && iz == 0) {
    for (int iy = 0; iy < D.ny; iy++) V2[(size_t)(iy + 2) * nc + col] = make_double2(0.0, 0.0);
    return;
  }
  C u[5], v[5], w[5];
  for (int k = 1; k < 5; k++) {
    const size_t r = (size_t)(k - 1) * nc + col;  // rows -2..1
    u[k] = ld(V1 + r); v[k] = ld(V2 + r); w[k] = ld(V3 + r);
  }
  for (int iy = 0; iy < D.ny; iy++) {
    for (int k = 0; k < 4; k++) { u[k] = u[k + 1]; v[k] = v[k + 1]; w[k] = w[k + 1]; }
    const size_t r = (size_t)(iy + 4) * nc + col;
    u[4] = ld(V1 + r); v[4] = ld(V2 + r); w[4] = ld(V3 + r);
    C du{0, 0}, dv{0, 0}, dw{0, 0};
#pragma unroll
    for (int j = 0; j < 5; j++) {
      du = du + DER(der, iy, 0, j - 2) * u[j];
      dv = dv + DER(der, iy, 1, j - 2) * v[j];
      dw = dw + DER(der, iy, 0, j - 2) * w[j];
    }
    const size_t o = (size_t)(iy + 2) * nc + col;
    if (ix != 0) {
      const C div = dv + imul(be, dw) + imul(al, du);
      // u - div / (i alfa) = (u.x - div.y / alfa, u.y + div.x / alfa)
      st(V1 + o, C{u[2].x - div.y / al, u[2].y + div.x / al});
    } else {
      const C div = dv + imul(be, dw);
      st(V3 + o, C{w[2].x - div.y / be, w[2].y + div.x / be});
    }
  }
}

int launch_project(const Dims &D, double2 *V1, double2 *V2, double2 *V3, const double *der, hipStream_t s) {
  hipLaunchKernelGGL(k_project, dim3((D.ncol + 255) / 256), dim3(256), 0, s, D, V1, V2, V3, der);
  return hipGetLastError();
}

// ------------------------------------------------------------------ layouts

// reference layout (col-major f(iy, iz, ix): row fastest) <-> device [row][col]
__global__ void k_transpose_c(const double2 *__restrict__ A, double2 *__restrict__ B, int n1, int n2) {
  // B[j][i] = A[i][j], A is n1 x n2 (row-major, j fastest)
  __shared__ double2 t[32][33];
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int i = i0 + k, j = j0 + threadIdx.x;
    if (i < n1 && j < n2) t[k][threadIdx.x] = A[(size_t)i * n2 + j];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int j = j0 + k, i = i0 + threadIdx.x;
    if (i < n1 && j < n2) B[(size_t)j * n1 + i] = t[threadIdx.x][k];
  }
}

int launch_to_device_layout(const Dims &D, const double2 *ref, double2 *dev, hipStream_t s) {
  // ref: [col][row] (row fastest)  ->  dev: [row][col]
  dim3 grid((D.NYg + 31) / 32, (D.ncol + 31) / 32);
  hipLaunchKernelGGL(k_transpose_c, grid, dim3(32, 8), 0, s, ref, dev, D.ncol, D.NYg);
  return hipGetLastError();
}

int launch_to_ref_layout(const Dims &D, const double2 *dev, double2 *ref, hipStream_t s) {
  dim3 grid((D.ncol + 31) / 32, (D.NYg + 31) / 32);
  hipLaunchKernelGGL(k_transpose_c, grid, dim3(32, 8), 0, s, dev, ref, D.NYg, D.ncol);
  return hipGetLastError();
}

template <typename PR>
__global__ void k_product_to_field(const typename PTraits<PR>::T2 *__restrict__ P, double2 *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const Cx<PR> v = ldp<PR>(P + i);
    dst[i] = make_double2((double)v.x, (double)v.y);
  }
}

int launch_product_to_field(const Dims &D, int mixed, const void *P, double2 *dst, hipStream_t s) {
  const size_t n = (size_t)D.NYg * D.ncol;
  dim3 grid((unsigned)((n + 255) / 256));
  if (mixed)
    hipLaunchKernelGGL(k_product_to_field<float>, grid, dim3(256), 0, s, (const float2 *)P, dst, n);
  else
    hipLaunchKernelGGL(k_product_to_field<double>, grid, dim3(256), 0, s, (const double2 *)P, dst, n);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_narrow(const double2 *__restrict__ src, float2 *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

__global__ void __launch_bounds__(256) k_widen(const float2 *__restrict__ src, double2 *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = make_double2((double)src[i].x, (double)src[i].y);
}

int launch_narrow(const double2 *src, float2 *dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_narrow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

int launch_widen(const float2 *src, double2 *dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

}  // namespace hst
