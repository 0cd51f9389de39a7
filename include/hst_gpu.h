/*
 * hst_gpu.h -- C ABI of libhstgpu.so, the MI355X (gfx950) implementation of
 * the RK3 substep of hst (homogeneous shear turbulence DNS).
 *
 * The Fortran host binds these with iso_c_binding interface blocks (module
 * hst_gpu_api); Python binds them with ctypes.  Every entry point returns 0
 * on success and a non-zero status on failure; the text of the last failure
 * is available from hst_gpu_last_error().  The reference's convention (every
 * error is fatal, `error stop`) is kept by the Fortran wrappers, which stop
 * on a non-zero status.
 *
 * Device state is process-global (one GPU per process, like one device per
 * rank in the reference): hst_gpu_init allocates it, hst_gpu_free releases
 * it.  Nothing is re-entrant; all work is in order on one HIP stream.
 *
 * Field identifiers name the device-resident arrays of the reference's
 * module hst_params (src/hst_params.f90:101-103):
 *   HST_V1..HST_V3      V(:, :, :, 1:3)      u, v, w
 *   HST_RHS1, HST_RHS2  rhs(:, :, :, 1:2)    eta and d2v right-hand sides
 *   HST_OLD1, HST_OLD2  oldrhs(:, :, :, 1:2) carried explicit terms
 *   HST_SCRATCH         a spare field of the same shape (tests, pressure)
 * Host copies of a field use the reference layout field(iy, iz, ix),
 * iy = -2..ny+1 fastest, then iz = -nz..nz, then ix = 0..nx (complex128).
 */
#ifndef HST_GPU_H
#define HST_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The deck values the substep needs: the namelist groups &mesh, &physics,
 * &init and &damping of hst.in (src/hst_input.f90:19-24, defaults :26-42,
 * checks :62-82).  Logicals are 0/1.  The Stokes layer (src/hst_stokes.f90)
 * is on when sl_amplitude != 0; the large-scale damping (src/hst_damping.f90)
 * when damp_kpar, damp_kperp > 0 and (damp_hard or damp_rate > 0). */
typedef struct hst_deck {
  int32_t nx, ny, nz;              /* &mesh: modes in x, points in y, modes in z */
  int32_t linear, exact_shift;     /* &physics logicals */
  int32_t seed, tg_mode;           /* &init */
  int32_t initial_condition;       /* 0: 'random', 1: 'taylor_green' */
  double alfa0, beta0, ly, ystretch;
  double re, S, s2_mean, s2_amplitude, s2_period, s2_start;
  double amplitude, kpeak, tg_amplitude;
  double sl_amplitude;             /* Stokes layer amplitude A (0: off) */
  double damp_kpar, damp_kperp;    /* damping region semi-axes (both 0: off) */
  double sl_period, sl_delta, sl_start;                /* layer period T (> 0), width, start t0 */
  double damp_rate, damp_edge, damp_norm, damp_angle;  /* soft rate, edge width, norm p >= 1, angle (deg) */
  int32_t sl_bodyforce, sl_ramp;                       /* &physics logicals of the layer */
  int32_t damp_hard, damp_mean, damp_follow;           /* &damping logicals */
} hst_deck;

/* Derived sizes and grid tables (src/hst_input.f90:92-115, src/hst_setup.f90:42-64). */
typedef struct hst_grid_info {
  int32_t nxd, nzd;                /* 3/2-padded sizes: 2*nxd real x points, nzd z points */
  int32_t nxm, nzm, nyg;           /* nx+1 x modes, 2nz+1 z modes, ny+4 rows with ghosts */
  int32_t ncol;                    /* nxm*nzm Fourier columns */
  double lx, lz, dx, dy, dz, factor, ni;
} hst_grid_info;

enum {
  HST_V1 = 1, HST_V2 = 2, HST_V3 = 3,
  HST_RHS1 = 4, HST_RHS2 = 5, HST_OLD1 = 6, HST_OLD2 = 7,
  HST_SCRATCH = 8
};

/* System kinds of line_solve (src/hst_linsolve.f90:83). */
enum { HST_KIND_D2V = 1, HST_KIND_ETA = 2, HST_KIND_POISSON = 3, HST_KIND_D0 = 4, HST_KIND_DY = 5 };

/* Arithmetic of the transforms, products and derivative right-hand sides. */
enum { HST_FP64 = 0, HST_MIXED = 1 };

/* ------------------------------------------------------------------ host-only (no GPU) */

/* Derived parameters, y grid, dyl and the compact weights der(0:ny-1, 0:3, -2:2)
 * (Fortran order, iy fastest): read_input's derived block (src/hst_input.f90:92-115),
 * allocate_fields' grid (src/hst_setup.f90:42-53), setup_derivatives
 * (src/hst_derivatives.f90:30-56).  y has ny+4 entries (iy = -2..ny+1), dyl ny,
 * der 20*ny.  Any output pointer may be NULL. */
int hst_grid_setup(const hst_deck *deck, hst_grid_info *info, double *y, double *dyl, double *der);

/* The seeded random field or the Taylor-Green field into a host array in the
 * reference layout, ghost rows included: generate_initial_field
 * (src/hst_initial.f90:27-67, 76-108). */
int hst_initial_field(const hst_deck *deck, double *V_host /* complex128[3][nx+1][2nz+1][ny+4] */);

/* <u_i u_i> of the seeded random field of the deck on a uniform y grid,
 * sum over the modes of w |u|^2 (w = 1 at ix = 0, else 2), which the discrete
 * orthogonality of the y modes |m| <= ny/3 makes exact; the amplitude that
 * gives <u_i u_i> = q2 is amplitude * sqrt(q2 / result). */
int hst_initial_energy(const hst_deck *deck, double *q2);

/* ------------------------------------------------------------------ device state */

/* allocate_fields + init_fft + init_linsolve + setup_derivatives
 * (src/hst_setup.f90:42, src/hst_fft.f90:75, src/hst_linsolve.f90:99,
 *  src/hst_derivatives.f90:30).  precision: HST_FP64 or HST_MIXED. */
int hst_gpu_init(const hst_deck *deck, int precision);
/* free_fields + free_fft + free_linsolve (src/hst_setup.f90:73, hst_fft.f90:317). */
int hst_gpu_free(void);
/* device_sync (src/hst_fft.f90:343). */
int hst_gpu_sync(void);
/* Text of the last failure (NUL-terminated, truncated to len). */
int hst_gpu_last_error(char *buf, int len);
/* Grid tables as computed at init (same as hst_grid_setup). */
int hst_gpu_grid(hst_grid_info *info, double *y, double *dyl, double *der);

/* `!$omp target update to/from` of one field (src/hst.f90:237,271; hst_io.f90:173). */
int hst_gpu_upload_field(int field, const double *host);
int hst_gpu_download_field(int field, double *host);
/* field = 0 (zero-fill) helper for rhs/oldrhs at set-up (src/hst_setup.f90:67-69). */
int hst_gpu_zero_field(int field);

/* ------------------------------------------------------------------ the substep's pieces */

/* fill_ghosts / fill_ghosts_field -> exchange_ghost_rows -> wrap_rows with
 * npy = 1 (src/hst_derivatives.f90:134-157, src/hst_mpi.f90:273,340-355). */
int hst_gpu_fill_ghosts(int field, double shift_x, double shift_z);

/* transform_to_physical (src/hst_transforms.f90:74-90), then the six products
 * (build_products, :97-121, both groups) and products_to_spectral (:125-139):
 * the product spectra of both groups are left on the device.  When cfl is
 * non-NULL the CFL number of the physical velocity is computed on the way
 * (compute_cfl, :163-179) and returned. */
int hst_gpu_nonlinear(double *cfl);
/* transform_to_physical + compute_cfl only (the linear=.true. substep 3). */
int hst_gpu_compute_cfl(double *cfl);
/* vvdz_to_field (src/hst_transforms.f90:142-157): product p (1..6, in the order
 * uu vv ww uv vw uw) of the last hst_gpu_nonlinear into a field. */
int hst_gpu_product_to_field(int p, int field);
/* The round trip of test_roundtrip (tests/test_roundtrip.f90:52-77): V -> physical
 * -> (times factor) -> spectral, left as products 1..3. */
int hst_gpu_roundtrip(void);

/* ---- per-field transforms on the reference's host buffers (module hst_fft's
 * FFT, IFT, RFT, HFT, src/hst_fft.f90:181-237, and build_products,
 * src/hst_transforms.f90:97-121), FP64 on the device whatever the precision of
 * the state; each call moves its lines to the device and back.  For callers
 * that run the transforms field by field; the time step does not use them. */

/* IFT (sign +1, exp(+i)) or FFT (sign -1) of nlines z lines of nzd complex,
 * contiguous (VVdz(:, :, :, m): nlines = (nx+1)(ny+4)), in place, unnormalised. */
int hst_gpu_fft_z(double *lines, int64_t nlines, int sign);
/* RFT (dir +1, c2r: modes 0..nxd -> 2 nxd real points, the imaginary parts of
 * modes 0 and nxd ignored, the two spare reals of a line set to 0) or HFT
 * (dir -1, r2c) of nlines x lines stored as nxd+1 complex (VVdx, VVdp: the
 * padded in-place layout), in place, unnormalised. */
int hst_gpu_fft_x(double *lines, int64_t nlines, int dir);
/* build_products(g): vel = rVVdx(:, :, :, 1:3), prod = products(:, :, :, 1:3),
 * each field nlines x lines of 2(nxd+1) reals; group 1: uu, vv, ww, 2: uv,
 * vw, uw, times factor = 1/(2 nxd nzd), on the 2 nxd points of a line (the
 * spare reals of prod are left as they are). */
int hst_gpu_products(int g, int64_t nlines, const double *vel, double *prod);
/* The product spectra VVdz(nzd, nx+1, ny+4, 1:3) of group g (as
 * products_to_spectral leaves them) into the device state, where
 * hst_gpu_buildrhs(ode, g) reads them: the kept modes iz = -nz..nz
 * (VVdz(izd(iz) + 1, ...)) of every row. */
int hst_gpu_set_products(int g, const double *vvdz);

/* buildrhs_prepare (src/hst_equations.f90:92-141) at `time` with step deltat. */
int hst_gpu_buildrhs_prepare(const double ode[3], double deltat, double time);
/* buildrhs(ODE, g) (src/hst_equations.f90:151-213) from the product spectra of group g. */
int hst_gpu_buildrhs(const double ode[3], int g);
/* shear_shift(dt_sub) at `time` (src/hst_equations.f90:226-293): default node
 * phase or, with exact_shift, the D0-unweighted shift. */
int hst_gpu_shear_shift(double dt_sub, double time);
/* line_solve(sys, lambda, src, dst, shift_x, shift_z) (src/hst_linsolve.f90:147-233). */
int hst_gpu_line_solve(int kind, double lambda, int src, int dst, double shift_x, double shift_z);
/* The line solve of hst_gpu_line_solve on caller-owned device arrays, with no
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight solver loops,
 * harnesses that forbid hipMalloc).  Solves the lines of the columns
 * col0 .. col0 + ncols - 1 (column order (iz + nz) + (2nz + 1) ix): src and
 * dst are device arrays of complex128 [ny + 4 rows][ncols] (row stride ncols,
 * rows iy = -2 .. ny + 1) holding just those columns; the interior rows of dst
 * receive the solution (the ghost rows are not written; KIND_DY reads the
 * ghost rows of src).  The workspace is a fixed device array of
 * HST_NOALLOC_MAX_BYTES inside the library (per device, held until the
 * process ends): ncols * (ny - 2) * 48 bytes for D2V, ETA, POISSON and
 * * 16 bytes for D0, DY must fit in it, else the call fails (solve fewer
 * columns per call).  Needs hst_gpu_init (the grid and its tables); calls no
 * hipMalloc or hipFree.  Enqueued on the library's stream (hst_gpu_sync
 * waits for it).  Every call on a device shares that workspace: calls that
 * overlap in time on one device race.  Nothing in the workspace is read
 * before the call writes it. */
#define HST_NOALLOC_MAX_BYTES (128u * 1024u * 1024u)
int hst_gpu_line_solve_noalloc(int kind, double lambda, int col0, int ncols, const double *src, double *dst,
                               double shift_x, double shift_z);
/* linsolve(lambda) at `time` (src/hst_equations.f90:298-326): solves, ghosts, u and w. */
int hst_gpu_linsolve(double lambda, double time);

/* timestep() (src/hst_equations.f90:56-87): one RK3 step of length deltat.
 * *time is advanced in FP64 by 2/RK_rai(1,i)*deltat per substep, as the
 * reference does; *cfl receives the substep-3 CFL number (may be NULL).
 * Without a closure, Stokes layer or damping (and not linear) the step keeps
 * eta and dv/dy in the device arrays of u and w between substeps and between
 * consecutive calls; every other entry point that touches the velocity
 * (download, outstats, the solves, ...) restores u, w first, so the state seen
 * through this interface is the reference's.  hst_gpu_sync, hst_gpu_grid and
 * hst_gpu_profile leave it as it is. */
int hst_gpu_timestep(double deltat, double *time, double *cfl);

/* outstats (src/hst_stats.f90:136-270) at `time` (the wrap phase of its dv/dy
 * solves), npy = 1, no Stokes layer or damping: out[0..9] = energy, diss, uv,
 * vw, uu, vv, ww, uw (box means, as written to Runtimedata and
 * variances_runtime.dat), mfx, mfz. */
int hst_gpu_outstats(double time, double out[10]);

/* stokes_apply (src/hst_stokes.f90:103-118) at `time`: the (0,0) mode of w
 * set to the layer profile W(y, time) on every row (without the body force or
 * within two periods of sl_start) or on the rows 0 and ny-1, its ghost rows
 * refilled.  hst_gpu_timestep calls it after each substep's solve; a host
 * calls it once after loading a field (src/hst.f90:76-77).  No-op without
 * the layer. */
int hst_gpu_stokes_apply(double time);

/* project_restart (src/hst_io.f90:213-250) with the ghost fills around it
 * (:184-190), at `time`: after a restart_remap = 'mode_index' load, the
 * discrete divergence i alfa D0 u + D1 v + i beta D0 w of every interior row
 * is taken off u (ix > 0) or w (ix = 0), and v = 0 in the (0,0) column.  Each
 * row's divergence is that of the loaded field (the reference's loop updates
 * in place while neighbouring rows read; here no row sees another's update). */
int hst_gpu_project_divergence(double time);

/* damping_mask (src/hst_damping.f90:119-133): the right-hand sides and
 * carried terms of the hard-masked columns set to zero (for a host that runs
 * the pieces of a substep; hst_gpu_timestep does it inside buildrhs).  The
 * mask is the one of the last hst_gpu_buildrhs_prepare. */
int hst_gpu_damping_mask(void);

/* The extra lines of the last hst_gpu_outstats (src/hst_stats.f90:270-284):
 * out[0..3] = energy outside and inside |y - ly/2| < 8 delta, dissipation
 * outside and inside (statistics/stokes_runtime.dat; zero without the layer);
 * out[4..7] = the damping direction in degrees, the number of damped columns,
 * half their energy and half the power they lose (damping_runtime.dat; zero
 * without damping). */
int hst_gpu_outstats_extra(double out[8]);

/* The fluctuation pressure of the current velocity (ghost rows filled) at
 * `time` into field dst (compute_pressure, src/hst_pressure.f90:35-112):
 * the products once more, the right-hand side in HST_RHS1, a POISSON line
 * solve, the (0,0) mode -<vv> about its y mean.  dst: HST_RHS2, HST_SCRATCH or
 * (FP64 form) HST_OLD1/2; outside a time step only. */
int hst_gpu_pressure(double time, int dst);

/* The column energies E(ix, iz) = ly/2 sum_y |u|^2 (the (0,0) column about
 * its box mean) averaged over the hst_gpu_outstats calls since the previous
 * call (spectrum_write, src/hst_stats.f90:110-134): E[ncol] in the column
 * order (iz + nz) + (2nz+1) ix, *nacc the number of samples; restarts the
 * average. */
int hst_gpu_spectrum(double *E, int32_t *nacc);

/* Rows per transform chunk (0: as many as memory allows; the physical-space
 * buffers hold one chunk of y rows).  Takes effect at the next hst_gpu_init. */
int hst_gpu_set_chunk_rows(int rows);
/* Device bytes held by the library and the rows per transform chunk in use. */
int hst_gpu_memory(int64_t *bytes, int *chunk_rows);

/* ------------------------------------------------------------------ measurement */

/* Per-kernel timing with HIP events on the library's stream (hst_timer's
 * phases, src/hst_timer.f90, at kernel granularity), every kernel timed alone:
 * while recording, the oldrhs shift that a time step otherwise runs on a second
 * stream beside the solves runs on the main stream (HST_SERIAL=1 in the
 * environment at hst_gpu_init does the same for an external profiler).
 * enable = 1 starts recording (and clears the table), 0 stops. */
int hst_gpu_profile(int enable);
/* Number of kernels in the table; name, launches, total milliseconds and the
 * algorithmic HBM bytes of all launches of entry i. */
int hst_gpu_profile_count(void);
int hst_gpu_profile_entry(int i, char *name, int len, int64_t *launches, double *ms, double *bytes);

/* The device's practical HBM bandwidth, measured (no counterpart in the
 * reference: the yardstick of the per-kernel GB/s).  Streaming kernels with no
 * arithmetic over buffers of `gib` GiB (0.25..64; far above the 256 MB last-level
 * cache for gib >= 4), four independent 16-byte loads in flight per lane,
 * non-temporal stores, several grid sizes, best of `reps` timed launches each:
 * out[0..2] = read-only, write-only and copy GB/s (copy counts bytes read plus
 * bytes written), out[3..5] the grid (blocks of 256) of each best.  Needs no
 * hst_gpu_init; allocates and frees its own buffers. */
int hst_gpu_bandwidth_probe(double gib, int reps, double out[6]);

#ifdef __cplusplus
}
#endif

#endif /* HST_GPU_H */
