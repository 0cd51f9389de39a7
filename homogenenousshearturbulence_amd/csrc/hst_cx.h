// Complex numbers of the kernels: a pair (x = real, y = imaginary) of float
// or double with the operations the transforms and the stencils need.
#pragma once

#include <hip/hip_runtime.h>

namespace hst {

template <typename R>
struct alignas(2 * sizeof(R)) Cx {
  R x, y;
};

#define HSTD __host__ __device__ __forceinline__

template <typename R> HSTD Cx<R> operator+(Cx<R> a, Cx<R> b) { return Cx<R>{a.x + b.x, a.y + b.y}; }
template <typename R> HSTD Cx<R> operator-(Cx<R> a, Cx<R> b) { return Cx<R>{a.x - b.x, a.y - b.y}; }
template <typename R> HSTD Cx<R> operator-(Cx<R> a) { return Cx<R>{-a.x, -a.y}; }
template <typename R> HSTD Cx<R> operator*(Cx<R> a, Cx<R> b) {
  return Cx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R> HSTD Cx<R> operator*(Cx<R> a, R s) { return Cx<R>{a.x * s, a.y * s}; }
template <typename R> HSTD Cx<R> operator*(R s, Cx<R> a) { return Cx<R>{s * a.x, s * a.y}; }
template <typename R> HSTD Cx<R> conj(Cx<R> a) { return Cx<R>{a.x, -a.y}; }
template <typename R> HSTD Cx<R> mul_i(Cx<R> a) { return Cx<R>{-a.y, a.x}; }    // i a
template <typename R> HSTD Cx<R> mul_mi(Cx<R> a) { return Cx<R>{a.y, -a.x}; }   // -i a
// (i b) a for real b: the product of the purely imaginary i b with a
template <typename R> HSTD Cx<R> imul(R b, Cx<R> a) { return Cx<R>{-(b * a.y), b * a.x}; }
template <typename R> HSTD Cx<R> cdiv(Cx<R> a, Cx<R> b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    R r = b.y / b.x, d = b.x + r * b.y;
    return Cx<R>{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  R r = b.x / b.y, d = b.y + r * b.x;
  return Cx<R>{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

template <typename R> HSTD Cx<R> cx_from(double2 v) { return Cx<R>{(R)v.x, (R)v.y}; }
HSTD Cx<double> ld(const double2 *p) { double2 v = *p; return Cx<double>{v.x, v.y}; }
HSTD void st(double2 *p, Cx<double> v) { *p = make_double2(v.x, v.y); }
HSTD Cx<double> ld(const float2 *p) { float2 v = *p; return Cx<double>{(double)v.x, (double)v.y}; }
HSTD void st(float2 *p, Cx<double> v) { *p = make_float2((float)v.x, (float)v.y); }
HSTD Cx<float> ldf(const float2 *p) { float2 v = *p; return Cx<float>{v.x, v.y}; }
// Nontemporal stores, for the right-hand sides: k_buildrhs 4-6 % faster at
// nx = nz = 256, ny = 513 (the solve that reads rhs next loses about a third
// of that, its first rows no longer in the last-level cache).  The solves'
// workspace, read back by the same kernel, the solved fields and dv/dy keep
// the ordinary stores (measured 2-4 % slower with these).
typedef double hst_dv2 __attribute__((ext_vector_type(2)));
typedef float hst_fv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stn(double2 *p, Cx<double> v) {
  const hst_dv2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, (hst_dv2 *)p);
}
__device__ __forceinline__ void stn(float2 *p, Cx<double> v) {
  const hst_fv2 t = {(float)v.x, (float)v.y};
  __builtin_nontemporal_store(t, (hst_fv2 *)p);
}

// u and w of a Fourier column from x = dv/dy and eta (linsolve,
// src/hst_equations.f90:303-324): u = (i alfa x - i beta eta) / k^2,
// w = (i beta x + i alfa eta) / k^2; the (0,0) column: u = Re eta, w = Im eta.
HSTD void recover_uw(double al, double be, double kk, bool m00, Cx<double> x, Cx<double> eta, Cx<double> &u,
                     Cx<double> &w) {
  if (m00) {
    u = Cx<double>{eta.x, 0.0};
    w = Cx<double>{eta.y, 0.0};
  } else {
    const double ik = 1.0 / kk;
    const Cx<double> t = imul(al, x) - imul(be, eta);
    const Cx<double> s = imul(be, x) + imul(al, eta);
    u = Cx<double>{t.x * ik, t.y * ik};
    w = Cx<double>{s.x * ik, s.y * ik};
  }
}

#undef HSTD

// Workgroups are dealt to the 8 XCDs round-robin by their linear id and every
// XCD has its own L2.  This renumbers the linear id so that each XCD receives
// a contiguous range: neighbouring blocks, which share cache lines, then run
// on the same XCD.  Returns the new linear id of the calling block.
__device__ __forceinline__ int xcd_linear() {
  const int nx = gridDim.x, ny = gridDim.y, n = nx * ny * gridDim.z;
  const int lid = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const int per = n / 8, rem = n % 8, x = lid % 8, s = lid / 8;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + s;
}

}  // namespace hst
