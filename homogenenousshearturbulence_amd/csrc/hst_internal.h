// Internal declarations of libhstgpu (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hst_gpu.h"

namespace hst {

void set_error(const std::string &msg);
std::string check_deck(const hst_deck &d);
void grid_setup(const hst_deck &d, hst_grid_info &g, std::vector<double> &y, std::vector<double> &dyl,
                std::vector<double> &der);
void initial_field(const hst_deck &d, const std::vector<double> &y, std::complex<double> *V);

// The factorisation of the D0 matrix of one line (bordering, one slab): per
// interior row i = 0..ny-3 the multipliers of rows i-2 and i-1, the inverse
// pivot, the upper entries, the border columns and the weights of rows 0
// and 1 of the inverse unit upper factor; and the real half of the record.
struct D0Factor {
  enum { L2 = 0, L1, RP, U1, U2, B1, B2, W, V, NCOL };
  std::vector<double> rows;  // [(ny-2) * NCOL]
  double rec[25];            // record entries by their 1-based index; rec[0] = U1(ny-4)
  double c1m4, c1m3, d11, d12, c10, c2m3, d21, d22, c20, c21;  // border-row coefficients
};
void d0_factor(const std::vector<double> &der, int ny, D0Factor &f);

// ---------------------------------------------------------------- device side

// Sizes and constants every kernel takes by value.
struct Dims {
  int nx, ny, nz, NXm, NZm, NYg, ncol, nxd, nzd, NXp;
  // the row stride of the solves' workspace rows: ncol rounded up to 256
  // columns, so that every row starts on a 4 KB boundary (column sweeps over
  // such rows measured 7-10 % faster than over rows 16 B off the lines)
  int ncs;
  double alfa0, beta0, ni, S, factor, dx, dz;
};

// Border constants of the D0 solve, by value.
struct D0Const {
  double rec[25];
  double c1m4, c1m3, d11, d12, c10, c2m3, d21, d22, c20, c21;
};

// Workspace of one cyclic pentadiagonal sweep: [(ny-2) rows][ncol] each.
struct PentaWS {
  double *U1, *U2, *B1, *B2;
  double2 *XR;
};

// Kernel launchers (hst_kernels.hip); all enqueue on `stream` and return
// hipError_t as int.  Field pointers are double2 [NYg][ncol] arrays.
struct FftTables;  // twiddles, per precision

bool fft_size_supported(int n);
int make_fft_tables(const Dims &D, int mixed, void **tables);
void free_fft_tables(void *tables);

// The transform round trip for rows [row0, row0 + nrows) of the chunk:
// k_zift (z inverse of u, v, w into Hv), k_xpass (x inverse, the products
// or, mode 1, the velocities times factor, or, mode 2, nothing but the CFL
// partial maxima; x forward into Hp), k_zfft (z forward of np spectra into
// P0 + p*pstride, the spectral layout).
// compact: V[0] holds eta and V[2] dv/dy; u and w are recovered on the way in
int launch_zift(const Dims &D, int mixed, const double2 *const V[3], void *Hv, int row0, int nrows,
                const void *tables, hipStream_t s, int compact = 0);
int launch_xpass(const Dims &D, int mixed, int mode, const void *Hv, void *Hp, int row0, int nrows,
                 const void *tables, const double *dyl, double *cfl_part, int want_cfl, int *n_part, hipStream_t s);
int launch_zfft(const Dims &D, int mixed, int np, const void *Hp, void *P0, size_t pstride, int row0, int nrows,
                const void *tables, hipStream_t s);
double bytes_zift(const Dims &D, int mixed, int nrows);
double bytes_xpass(const Dims &D, int mixed, int mode, int nrows);
double bytes_zfft(const Dims &D, int mixed, int np, int nrows);

// buildrhs_prepare + buildrhs(1) + buildrhs(2) (+ the default shift when shift).
// nonlinear: 0 buildrhs_prepare only, 1 or 2 buildrhs(ODE, g) alone, 3 all.
struct RhsArgs {
  double ode0, ode1, ode2, deltat, s2now, dgx, dgz;
  int nonlinear, shift;
  // the physics hooks (all off: gdamp = fy = null, hard = no_mean_vw = 0)
  const double *gdamp = nullptr;  // per-column damping rates [ncol]
  double gsoft = 0.0;             // 1: the explicit damping term, 0: none or the hard mask
  int hard = 0;                   // the masked columns' rhs and oldrhs stored as zero
  int no_mean_vw = 0;             // the (0,0) column drops the vw divergence (layer with body force)
  const double *fy = nullptr;     // the layer's body force at the rows -2..ny+1
};
// oldrhs (O1, O2) is FP32 (float2) in the mixed form, FP64 otherwise.
int launch_phase4(const Dims &D, int mixed, double2 *R1, double2 *R2, void *O1, void *O2, const double *y,
                  double dgx, double dgz, hipStream_t s);
// u, w from eta (V1) and dv/dy (V3) in place, every row
int launch_uw_from_eta(const Dims &D, double2 *V1, double2 *V3, hipStream_t s);
int launch_buildrhs(const Dims &D, int mixed, const RhsArgs &a, const double2 *V1, const double2 *V2,
                    const double2 *V3, const void *const P[6], double2 *R1, double2 *R2, void *O1, void *O2,
                    const double *der, const double *y, hipStream_t s);

// Cyclic pentadiagonal solves with per-column rows (D2V, ETA, POISSON).
// nsys systems in one launch: system k solves kinds[k] from src[k] into dst[k].
// tabs (may be null): per system the factor table of launch_penta_factor for
// its kind and lambda; then the workspace holds only the eliminated rhs.
int launch_penta(const Dims &D, int nsys, const int *kinds, double lambda, const double2 *const *src,
                 double2 *const *dst, const PentaWS *ws, const void *const *tabs, const double *der, double sx,
                 double sz, hipStream_t s);
// The matrix part of the D2V or ETA lines for lambda: per (ix, |iz|) group
// the factorisation's state every 8 rows, from which the solves recompute
// the rows (penta_table_bytes of device memory).
size_t penta_table_bytes(const Dims &D);
int launch_penta_factor(const Dims &D, int kind, double lambda, const double *der, void *tab, hipStream_t s);

// D0-type solves with the shared factorisation.  rhs_mode 0: D0 x = src;
// 1: D0 x = D1 src (src ghost rows filled).  out_mode 0: x into dst;
// 1: u, w recovered from x = dv/dy and eta (dst = V1 holding eta, dst3 = V3);
// 2: sums of w |x|^2 into part, three doubles per block: all rows, the rows
// with inlayer set, the others (inlayer may be null: every row outside); the
// (0,0) column is left out when skip00 is set.
int launch_d0solve(const Dims &D, int rhs_mode, int out_mode, const double2 *src, double2 *dst, double2 *dst3,
                   double2 *XR, const double *f0rows, const D0Const &fc, const double *der, const double *dyl,
                   double sx, double sz, double *part, int *n_part, hipStream_t s, const int *inlayer = nullptr,
                   int skip00 = 0);
// D0 x = D1 src in one sweep (OUT 0 of launch_d0solve, the time step's dv/dy)
// with the back substitution windowed over dy_window_rows() rows; sig holds
// the back substitution of -B1, -B2 over the first dy_window_rows() rows
// ([i][2]).  The caller checks that the factorisation decays fast enough.
int launch_dy_window(const Dims &D, const double2 *src, double2 *dst, const double *f0rows, const D0Const &fc,
                     const double *der, const double *sig, double sx, double sz, hipStream_t s);
int dy_window_rows();
// batched FP64 line transforms of the per-field entry points (hst_lines.hip)
int launch_cfft_lines(int n, int sign, double2 *data, int64_t nlines, const void *tw, hipStream_t s);
int launch_rfft_lines(int m, int dir, double2 *data, int64_t nlines, const void *tw, const void *twh,
                      hipStream_t s);
int launch_products(const double *vel, double *prod, int64_t plane, int lw, int used, int g, double factor,
                    hipStream_t s);
// line_solve over a range of columns held in arrays of their own ([NYg][ncols]),
// workspace ws of line_cols_ws_bytes(D, kind, ncols) bytes
size_t line_cols_ws_bytes(const Dims &D, int kind, int ncols);
int launch_line_cols(const Dims &D, int kind, double lambda, int col0, int ncols, const double2 *src, double2 *dst,
                     void *ws, const double *f0rows, const D0Const &fc, const double *der, double sx, double sz,
                     hipStream_t s);

// The exact shift of q[0..nq) (rhs1, rhs2, oldrhs1, oldrhs2): D0 solve at the
// old shifts, node phase, ghosts at the new shifts, D0 re-weight.
// q[k] is float2 where f32[k] is set.  XR[k]: workspace of ny - 2 rows.
// PX, PZ: the node phases of launch_shift_phases for the substep's (dgx, dgz).
int launch_exact_shift(const Dims &D, int nq, void *const *q, const int *f32, double2 *const *XR,
                       const double *f0rows, const D0Const &fc, const double *der, const double2 *PX,
                       const double2 *PZ, double sx0, double sz0, double sx1, double sz1, hipStream_t s);
// The same shift in one sweep per array (the D0 solve windowed as in
// launch_dy_window, sig its border terms): q read and written once, XR holding
// the first dy_window_rows() + 4 rows in FP64; needs ny - 2 >= 104.
int launch_shift_window(const Dims &D, int nq, void *const *q, const int *f32, double2 *const *XR,
                        const double *f0rows, const D0Const &fc, const double *der, const double *sig,
                        const double2 *PX, const double2 *PZ, double sx0, double sz0, double sx1, double sz1,
                        hipStream_t s);
// The exact shift of rhs2 and rhs1 and their D2V and ETA line solves at lambda
// in one launch (v2, eta out; the shifted right-hand sides are not stored):
// f0r, fcr the D0 factorisation of the reversed line (d0_factor of the
// mirrored stencils); xrs, xrp the two sweeps' workspaces per system
// (distinct); tabs the D2V and ETA factor tables.
int launch_shift_penta(const Dims &D, const double2 *q2, const double2 *q1, double2 *const xrs[2],
                       double2 *const xrp[2], const void *const tabs[2], double2 *v2, double2 *eta,
                       const double *f0r, const D0Const &fcr, const double *der, const double2 *PX,
                       const double2 *PZ, double sx0, double sz0, double sx1, double sz1, double lambda,
                       hipStream_t s);
// PX[i][ix] = exp(-i alfa0 ix dgx y_i) (ny x (nx+1)), PZ[i][iz+nz] = exp(-i beta0 iz dgz y_i) (ny x (2nz+1)).
int launch_shift_phases(const Dims &D, const double *y, double dgx, double dgz, double2 *PX, double2 *PZ,
                        hipStream_t s);
// float2 <-> double2 copies of n complex values
int launch_narrow(const double2 *src, float2 *dst, size_t n, hipStream_t s);
int launch_widen(const float2 *src, double2 *dst, size_t n, hipStream_t s);

// Ghost rows of nf fields at the shifts (sx, sz).
int launch_ghosts(const Dims &D, int nf, double2 *const *f, double sx, double sz, hipStream_t s);

// First reduction of the statistics: per block NSTAT sums, in the order
// uu vv ww uv vw uw eps mfx mfz q_in q_out e_in e_out ed pd ncol_damped.
// exz_acc (may be null): the column energies sum_y w |u|^2 are added to it.
constexpr int NSTAT = 16;
struct StatsArgs {
  const double *gdamp = nullptr;  // damping rates (null: none)
  double gsoft = 0.0;
  const int *inlayer = nullptr;   // rows inside the Stokes layer (null: none)
  int inc00 = 1;                  // the (0,0) column enters uu, ww, uw
};
int launch_stats(const Dims &D, const StatsArgs &a, const double2 *V1, const double2 *V2, const double2 *V3,
                 const double *dyl, double *part, int *n_part, double *exz_acc, hipStream_t s);

// The Stokes layer (src/hst_stokes.f90): the profile and body force as
// closed forms of y and t; ramp: the factor of the switch-on at t.
struct StokesParams {
  double amp, period, delta, start, ly, ni;
};
// fy[row] for the rows -2..ny+1 (ghosts: the periodic copies of the rows
// ny-2, ny-1, 0, 1) at time t (stokes_force, :90-101).
int launch_stokes_force(const StokesParams &p, int ny, const double *y, double t, double ramp, double *fy,
                        hipStream_t s);
// stokes_apply (:103-118): the (0,0) mode of w set to W(y, t) on every row
// (whole) or on rows 0 and ny-1, then its four ghost rows (phase 1).
int launch_stokes_apply(const Dims &D, const StokesParams &p, const double *y, double t, double ramp, int whole,
                        double2 *V3, hipStream_t s);

// The damping rates g(kx, kz) for the direction theta (damping_update,
// src/hst_damping.f90:78-115).
struct DampParams {
  double kpar, kperp, rate, edge, norm;
  int hard, mean;
};
int launch_damping_table(const Dims &D, const DampParams &p, double theta, double *gdamp, hipStream_t s);
// damping_mask (:119-133): rhs and oldrhs of the columns with g > 0 set to zero.
int launch_damping_mask(const Dims &D, int mixed, const double *gdamp, double2 *R1, double2 *R2, void *O1, void *O2,
                        hipStream_t s);
// The pressure equation's right-hand side from the six product spectra and v
// into R (interior rows), and the (0,0) column of p from it.
int launch_pressure_rhs(const Dims &D, int mixed, const void *P0, const double2 *V2, double2 *R, const double *der,
                        double s2now, hipStream_t s);
int launch_pressure_mean(const Dims &D, const double2 *R, double2 *p, const double *dyl, double ly, hipStream_t s);

// project_restart's divergence repair of u, v, w (interior rows; ghosts filled before).
int launch_project(const Dims &D, double2 *V1, double2 *V2, double2 *V3, const double *der, hipStream_t s);

// Layout conversion between the reference layout (host) and the device layout.
int launch_to_device_layout(const Dims &D, const double2 *ref, double2 *dev, hipStream_t s);
int launch_to_ref_layout(const Dims &D, const double2 *dev, double2 *ref, hipStream_t s);
int launch_product_to_field(const Dims &D, int mixed, const void *P, double2 *dst, hipStream_t s);


// ---------------------------------------------------------------- shared helpers

Dims make_dims(const hst_deck &d, const hst_grid_info &g);
D0Const d0_const(const D0Factor &f);
// The shear shifts (sx, sz) of a deck at time t (src/hst_derivatives.f90:123-128).
void deck_shifts(const hst_deck &d, double lx, double lz, double t, double &sx, double &sz);

// ---------------------------------------------------------------- the live state, as the closure sees it

struct StateView {
  int mixed;
  Dims D;
  hst_deck deck;
  double2 *F[9];  // fields by id (HST_V1 .. HST_SCRATCH)
  void *P;        // the six product spectra [6][NYg][ncol], Cx<float> in the mixed form
  hipStream_t stream;
};
bool state_view(StateView &v);
// Called by hst_gpu_timestep after the product pass of every substep, with
// the time of the substep; a nonzero return stops the step.
typedef int (*StressHook)(void *ctx, double time);
void set_stress_hook(StressHook f, void *ctx);
// Runs a launch on the state's stream, timed when the state's profiling is on.
int profiled(const char *name, double bytes, const std::function<int()> &launch);

}  // namespace hst
