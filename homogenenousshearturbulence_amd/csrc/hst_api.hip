// The C ABI of libhstgpu (include/hst_gpu.h): device state, the entry points
// of the RK3 substep and the per-kernel event timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hst_cx.h"
#include "hst_internal.h"

namespace hst {

static std::string g_error;
static std::mutex g_error_mutex;
void set_error(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_error_mutex);
  g_error = msg;
}

// RK3 (Rai-Moin) coefficients, one row per substep: (unknown/deltat, new
// explicit, old explicit); substep lengths 2/RK(i,0) deltat = 8/15, 2/15, 1/3.
static const double RK[3][3] = {{120.0 / 32.0, 2.0, 0.0},
                                {120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0},
                                {120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0}};
static const double TWOPI = 6.283185307179586;

// ------------------------------------------------------------------ profiling

struct ProfEntry {
  int64_t launches = 0;
  double ms = 0.0, bytes = 0.0;
};
struct Pending {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};

struct State {
  bool live = false;
  hst_deck deck;
  int mixed = 0;
  hst_grid_info g;
  Dims D;
  std::vector<double> y, dyl, der;
  D0Factor f0;
  D0Const fc;
  double *d_der = nullptr, *d_y = nullptr, *d_dyl = nullptr, *d_f0 = nullptr, *d_part = nullptr;
  // the D0 factorisation of the reversed line (the exact shift fused with the solves)
  D0Factor f0r;
  D0Const fcr;
  double *d_f0r = nullptr;
  size_t part_cap = 0;
  double *h_part = nullptr;  // pinned host copy of d_part (the CFL maxima of a time step)
  int n_part_pending = 0;
  double2 *F[9] = {nullptr};
  void *P = nullptr, *Hv = nullptr, *Hp = nullptr, *ws = nullptr;
  size_t ws_bytes = 0;
  int chunk = 0;
  PentaWS pws[2];
  double2 *XR[4];
  // factor tables of the D2V and ETA lines for up to three lambdas (the RK3
  // substeps at a fixed step), built on first use when memory allows
  void *ptab[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  double ptab_lam[3] = {0.0, 0.0, 0.0};
  int ptab_next = 0;
  bool ptab_off = false;
  void *tables = nullptr;
  hipStream_t stream = nullptr;
  // the oldrhs half of the exact shift runs on a second stream, beside the
  // solves and the next substep's transforms, and is joined before the next
  // buildrhs (or before any other entry point)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool side_pending = false;
  double2 *XRo[2] = {nullptr, nullptr};
  double2 *PX = nullptr, *PZ = nullptr;  // node phases of the exact shift
  double *exz_acc = nullptr;             // column energies summed over the outstats calls
  int exz_n = 0;
  double mean_acc = 0.0;                 // summed squared box means of the (0,0) column
  // physics hooks: the damping rates per column (rebuilt when the direction
  // changes), the layer's body force rows and the rows inside the layer
  double *d_gdamp = nullptr, *d_fy = nullptr;
  int *d_inlayer = nullptr;
  std::vector<int> inlayer;
  double theta_last = 0.0;
  bool theta_set = false;
  double extra[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // the last outstats' layer and damping lines
  int64_t bytes_held = 0;
  bool prof = false;
  bool serial = false;  // the oldrhs shift on the main stream (HST_SERIAL=1)
  // the time step's dv/dy in one sweep (k_dy_window) when the D0 factor decays
  // fast enough on this grid; HST_NO_DYWIN=1: the stored sweep of k_d0solve
  bool dyw = false;
  double *d_sig = nullptr;
  // the exact shift of oldrhs in one sweep (k_shift_window) under the same
  // condition; HST_NO_SHWIN=1: the chunked two-sweep k_exact_shift
  bool shw = false;
  // a time step leaves eta in V1 and dv/dy in V3 (every row) instead of u
  // and w: the next step's transforms and right-hand sides read that form,
  // and every other call restores u, w first (restore_uw)
  bool compact = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, ProfEntry> table;
  std::vector<std::string> order;
};

static State S;
static int g_chunk_request = 0;

// ------------------------------------------------------------------ the no-allocation workspace

// The workspace of hst_gpu_line_solve_noalloc: a file-scope device array, so
// that the entry point allocates nothing.  Every device the library runs on
// holds one copy until the process ends.  Host code never passes the array's
// name to a kernel (on the host it only stands for the symbol): it takes the
// device address once per device (scratch_attach) and passes that.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)HST_NOALLOC_MAX_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static void *g_scratch_dev[MAX_DEVICES] = {nullptr};

static bool field_f32(int f) { return S.mixed && (f == HST_OLD1 || f == HST_OLD2); }

static int fail(const std::string &what, int rc = 1) {
  set_error(what);
  return rc ? rc : 1;
}

static int hip_fail(const char *where, hipError_t e) {
  return fail(std::string(where) + ": " + hipGetErrorString(e), (int)e);
}

#define HIPCHK(call)                                          \
  do {                                                        \
    hipError_t _e = (call);                                   \
    if (_e != hipSuccess) return hip_fail(#call, _e);         \
  } while (0)

// The device address of the current device's copy of scratch, taken once
// per device with that device current (at hst_gpu_init, or at the first
// call on a device that init did not run on).
static int scratch_attach(void **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(std::string("hipGetDevice: ") + hipGetErrorString(e), (int)e);
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device ordinal out of range for the workspace table");
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), e ? (int)e : 1);
    g_scratch_dev[dev] = p;
  }
  if (ptr) *ptr = g_scratch_dev[dev];
  return 0;
}

static hipEvent_t take_event() {
  if (!S.pool.empty()) {
    hipEvent_t e = S.pool.back();
    S.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

static void flush_profile() {
  if (S.pending.empty()) return;
  (void)hipStreamSynchronize(S.stream);
  if (S.side) (void)hipStreamSynchronize(S.side);
  for (auto &p : S.pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    if (!S.table.count(p.name)) S.order.push_back(p.name);
    ProfEntry &e = S.table[p.name];
    e.launches++;
    e.ms += ms;
    e.bytes += p.bytes;
    S.pool.push_back(p.a);
    S.pool.push_back(p.b);
  }
  S.pending.clear();
}

// Runs one launcher; with timing on, brackets it with events on its stream.
template <class F>
static int run_on(hipStream_t st, const char *name, double bytes, F launch) {
  hipEvent_t a = nullptr, b = nullptr;
  if (S.prof) {
    a = take_event();
    b = take_event();
    (void)hipEventRecord(a, st);
  }
  int rc = launch();
  if (rc) return fail(std::string("launch of ") + name + " failed: " + hipGetErrorString((hipError_t)rc), rc);
  if (S.prof) {
    (void)hipEventRecord(b, st);
    S.pending.push_back(Pending{name, a, b, bytes});
    if (S.pending.size() > 4096) flush_profile();
  }
  return 0;
}

template <class F>
static int run(const char *name, double bytes, F launch) {
  return run_on(S.stream, name, bytes, launch);
}

int profiled(const char *name, double bytes, const std::function<int()> &launch) {
  return run(name, bytes, [&] { return launch(); });
}

// ------------------------------------------------------------------ clock

static double s2_of_deck(const hst_deck &d, double t) {
  if ((d.s2_amplitude == 0.0 && d.s2_mean == 0.0) || t < d.s2_start) return 0.0;
  if (d.s2_period > 0.0) return d.s2_mean + d.s2_amplitude * std::sin(TWOPI / d.s2_period * (t - d.s2_start));
  return d.s2_mean + d.s2_amplitude;
}

static double s2_integral_deck(const hst_deck &d, double t1, double t2) {
  if ((d.s2_amplitude == 0.0 && d.s2_mean == 0.0) || t2 <= d.s2_start) return 0.0;
  const double a = std::max(t1, d.s2_start), b = t2;
  if (d.s2_period > 0.0) {
    const double omega = TWOPI / d.s2_period;
    return d.s2_mean * (b - a) +
           d.s2_amplitude / omega * (std::cos(omega * (a - d.s2_start)) - std::cos(omega * (b - d.s2_start)));
  }
  return (d.s2_mean + d.s2_amplitude) * (b - a);
}

static double s2_of(double t) { return s2_of_deck(S.deck, t); }
static double s2_integral(double t1, double t2) { return s2_integral_deck(S.deck, t1, t2); }

// Fortran MODULO of reals (fmod, then the sign fix)
static double modulo_r(double a, double p) {
  double r = std::fmod(a, p);
  if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r += p;
  return r;
}

void deck_shifts(const hst_deck &d, double lx, double lz, double t, double &sx, double &sz) {
  sx = modulo_r(d.S * t * d.ly, lx);
  sz = modulo_r(s2_integral_deck(d, d.s2_start, t) * d.ly, lz);
}

static void shear_shifts(double t, double &sx, double &sz) { deck_shifts(S.deck, S.g.lx, S.g.lz, t, sx, sz); }

// ------------------------------------------------------------------ physics hooks

// stokes_active, damping_active, damping_soft_factor, damping_theta
// (src/hst_stokes.f90:48-50, src/hst_damping.f90:45-67)
static bool stokes_active() { return S.deck.sl_amplitude != 0.0; }
static bool damping_active() {
  return S.deck.damp_kpar > 0.0 && S.deck.damp_kperp > 0.0 && (S.deck.damp_hard || S.deck.damp_rate > 0.0);
}
static double damping_soft_factor() { return (damping_active() && !S.deck.damp_hard) ? 1.0 : 0.0; }
static double damping_theta(double t) {
  if (S.deck.damp_follow) return std::atan2(s2_of(t), S.deck.S);
  return S.deck.damp_angle * 3.141592653589793 / 180.0;
}

// the switch-on factor of the layer (src/hst_stokes.f90:52-62)
static double stokes_ramp(double t) {
  const hst_deck &d = S.deck;
  if (t < d.sl_start) return 0.0;
  if (d.sl_ramp) {
    const double tau = std::min((t - d.sl_start) / d.sl_period, 1.0);
    return 3.0 * tau * tau - 2.0 * tau * tau * tau;
  }
  return 1.0;
}

static StokesParams stokes_params() {
  return StokesParams{S.deck.sl_amplitude, S.deck.sl_period, S.deck.sl_delta, S.deck.sl_start, S.deck.ly, S.g.ni};
}

// damping_update (src/hst_damping.f90:78-115): the table for the direction at
// t, rebuilt only when the direction changed since the last build
static int damping_update(double t) {
  if (!damping_active()) return 0;
  const double th = damping_theta(t);
  if (S.theta_set && th == S.theta_last) return 0;
  S.theta_set = true;
  S.theta_last = th;
  const hst_deck &d = S.deck;
  const DampParams p{d.damp_kpar, d.damp_kperp, d.damp_rate, d.damp_edge, d.damp_norm, d.damp_hard, d.damp_mean};
  return run("damping_table", 0.0, [&] { return launch_damping_table(S.D, p, th, S.d_gdamp, S.stream); });
}

// stokes_force (src/hst_stokes.f90:90-101): the body force rows at t
static bool stokes_body() { return stokes_active() && S.deck.sl_bodyforce; }
static int stokes_force(double t) {
  if (!stokes_body()) return 0;
  return run("stokes_force", 0.0, [&] {
    return launch_stokes_force(stokes_params(), S.D.ny, S.d_y, t, stokes_ramp(t), S.d_fy, S.stream);
  });
}

// stokes_apply (src/hst_stokes.f90:103-118) at t, with the ghost rows of the column
static int stokes_apply(double t) {
  if (!stokes_active()) return 0;
  const int whole = (!S.deck.sl_bodyforce || (t - S.deck.sl_start <= 2.0 * S.deck.sl_period)) ? 1 : 0;
  return run("stokes_apply", 0.0, [&] {
    return launch_stokes_apply(S.D, stokes_params(), S.d_y, t, stokes_ramp(t), whole, S.F[HST_V3], S.stream);
  });
}

// ------------------------------------------------------------------ the closure's view of the state

static StressHook g_hook = nullptr;
static void *g_hook_ctx = nullptr;

void set_stress_hook(StressHook f, void *ctx) {
  g_hook = f;
  g_hook_ctx = ctx;
}

static int restore_uw();

bool state_view(StateView &v) {
  if (!S.live) return false;
  if (restore_uw()) return false;
  v.mixed = S.mixed;
  v.D = S.D;
  v.deck = S.deck;
  for (int i = 0; i < 9; i++) v.F[i] = S.F[i];
  v.P = S.P;
  v.stream = S.stream;
  return true;
}

// ------------------------------------------------------------------ pieces

static const void *P_of(int p) {  // product p = 0..5
  const size_t cs = S.mixed ? sizeof(float2) : sizeof(double2);
  return (const char *)S.P + (size_t)p * S.D.NYg * S.D.ncol * cs;
}

// mode 0: the six products; 1: the round trip; 2: the CFL number only;
// 3: the five combinations vv - uu, vv - ww, uv, vw, uw (a time step's own)
// defer: the CFL partial maxima are copied to host memory on the stream and
// reduced by cfl_finish() after the caller's next synchronisation, so that the
// rest of a time step is queued without a wait.
static int nonlinear_pass(int mode, bool want_cfl, double *cfl, bool defer = false) {
  const double2 *V[3] = {S.F[1], S.F[2], S.F[3]};
  const int np = (mode == 0 ? 6 : (mode == 3 ? 5 : 3));
  int npart_total = 0;
  for (int row0 = 0; row0 < S.D.NYg; row0 += S.chunk) {
    const int nrows = std::min(S.chunk, S.D.NYg - row0);
    int rc = run("zift", bytes_zift(S.D, S.mixed, nrows),
                 [&] { return launch_zift(S.D, S.mixed, V, S.Hv, row0, nrows, S.tables, S.stream, S.compact); });
    if (rc) return rc;
    int npart = 0;
    double *part = S.d_part + npart_total;
    rc = run(mode == 2 ? "xpass_cfl" : "xpass", bytes_xpass(S.D, S.mixed, mode, nrows), [&] {
      return launch_xpass(S.D, S.mixed, mode, S.Hv, S.Hp, row0, nrows, S.tables, S.d_dyl, part, want_cfl ? 1 : 0,
                          &npart, S.stream);
    });
    if (rc) return rc;
    if (want_cfl) npart_total += npart;
    if (mode == 2) continue;
    rc = run("zfft", bytes_zfft(S.D, S.mixed, np, nrows), [&] {
      return launch_zfft(S.D, S.mixed, np, S.Hp, S.P, (size_t)S.D.NYg * S.D.ncol, row0, nrows, S.tables, S.stream);
    });
    if (rc) return rc;
  }
  if (want_cfl && cfl && defer) {
    HIPCHK(hipMemcpyAsync(S.h_part, S.d_part, sizeof(double) * npart_total, hipMemcpyDeviceToHost, S.stream));
    S.n_part_pending = npart_total;
  } else if (want_cfl && cfl) {
    std::vector<double> h(npart_total);
    HIPCHK(hipMemcpyAsync(h.data(), S.d_part, sizeof(double) * npart_total, hipMemcpyDeviceToHost, S.stream));
    HIPCHK(hipStreamSynchronize(S.stream));
    double c = 0.0;
    for (double v : h) c = std::max(c, v);
    *cfl = c;
  }
  return 0;
}

// the CFL number of a deferred nonlinear_pass (the stream synchronised here)
static int cfl_finish(double *cfl) {
  HIPCHK(hipStreamSynchronize(S.stream));
  double c = 0.0;
  for (int k = 0; k < S.n_part_pending; k++) c = std::max(c, S.h_part[k]);
  S.n_part_pending = 0;
  *cfl = c;
  return 0;
}

static double bytes_field() { return (double)S.D.ny * S.D.ncol * 16.0; }

// buildrhs_prepare and/or buildrhs (mode 0 prepare, 1/2 one group, 3 all);
// mask: the hard damping mask fused into the stores (a whole substep only)
static int buildrhs(const double ode[3], double deltat, double time, int mode, bool shift, double dgx, double dgz,
                    bool mask = false) {
  RhsArgs a;
  a.ode0 = ode[0]; a.ode1 = ode[1]; a.ode2 = ode[2];
  a.deltat = deltat;
  a.s2now = s2_of(time);
  a.dgx = dgx; a.dgz = dgz;
  a.nonlinear = mode;
  a.shift = shift ? 1 : 0;
  if (damping_active()) {
    a.gdamp = S.d_gdamp;
    a.gsoft = damping_soft_factor();
    a.hard = (mask && S.deck.damp_hard) ? 1 : 0;
  }
  a.no_mean_vw = stokes_body() ? 1 : 0;
  if (mode == 0 || mode >= 3) {  // the modes with buildrhs_prepare
    if (int rc = stokes_force(time)) return rc;
    if (stokes_body()) a.fy = S.d_fy;
  }
  const void *P[6];
  for (int p = 0; p < 6; p++) P[p] = P_of(p);
  // prepare: u, v, w and oldrhs in, rhs and oldrhs out; a group adds its three
  // product spectra and reads and writes rhs, oldrhs.  Products and oldrhs are
  // FP32 in the mixed form.
  const double ps = S.mixed ? 8.0 : 16.0, pts = (double)S.D.ny * S.D.ncol;
  double bytes = pts * (5.0 * 16.0 + 4.0 * ps);
  if (mode == 3) bytes += 6.0 * pts * ps;
  else if (mode == 4) bytes += 5.0 * pts * ps;
  else if (mode == 5) bytes += 5.0 * pts * ps - pts * 16.0;  // no w
  else if (mode != 0) bytes = pts * (4.0 * 16.0 + 4.0 * ps + 3.0 * ps);
  return run(mode >= 3 ? "buildrhs" : (mode == 0 ? "buildrhs_prepare" : "buildrhs_group"), bytes, [&] {
    return launch_buildrhs(S.D, S.mixed, a, S.F[1], S.F[2], S.F[3], P, S.F[HST_RHS1], S.F[HST_RHS2], S.F[HST_OLD1],
                           S.F[HST_OLD2], S.d_der, S.d_y, S.stream);
  });
}

static int exact_shift(double dt_sub, double time) {
  const double dgx = S.deck.S * dt_sub, dgz = s2_integral(time, time + dt_sub);
  double sx0, sz0, sx1, sz1;
  shear_shifts(time, sx0, sz0);
  shear_shifts(time + dt_sub, sx1, sz1);
  if (int rc = run("shift_phases", 0.0, [&] {
        return launch_shift_phases(S.D, S.d_y, dgx, dgz, S.PX, S.PZ, S.stream);
      }))
    return rc;
  void *q[4] = {S.F[HST_RHS1], S.F[HST_RHS2], S.F[HST_OLD1], S.F[HST_OLD2]};
  const int f32[4] = {0, 0, S.mixed, S.mixed};
  // per array q in twice and out once (rhs FP64, oldrhs FP32 in the mixed form)
  const double qb = 2.0 * 3.0 * 16.0 + 2.0 * 3.0 * (S.mixed ? 8.0 : 16.0);
  return run("exact_shift", qb / 16.0 * bytes_field(), [&] {
    return launch_exact_shift(S.D, 4, q, f32, S.XR, S.d_f0, S.fc, S.d_der, S.PX, S.PZ, sx0, sz0, sx1, sz1,
                              S.stream);
  });
}

// The exact shift inside a time step: rhs on the main stream (the solves read
// it next), oldrhs on the side stream, joined before the next buildrhs.  With
// timing on both halves run on the main stream, so that every kernel's time
// is its own.
// fused: the rhs half is left to the fused shift-and-solve of the time step
static int exact_shift_split(double dt_sub, double time, bool fused = false) {
  const double dgx = S.deck.S * dt_sub, dgz = s2_integral(time, time + dt_sub);
  double sx0, sz0, sx1, sz1;
  shear_shifts(time, sx0, sz0);
  shear_shifts(time + dt_sub, sx1, sz1);
  void *qr[2] = {S.F[HST_RHS1], S.F[HST_RHS2]}, *qo[2] = {S.F[HST_OLD1], S.F[HST_OLD2]};
  const int fr[2] = {0, 0}, fo[2] = {S.mixed, S.mixed};
  double2 *xr[2] = {S.XR[0], S.XR[1]};
  const double os = S.mixed ? 8.0 : 16.0;
  if (int rc = run("shift_phases", 0.0, [&] {
        return launch_shift_phases(S.D, S.d_y, dgx, dgz, S.PX, S.PZ, S.stream);
      }))
    return rc;
  HIPCHK(hipEventRecord(S.ev_fork, S.stream));
  int rc = 0;
  if (!fused) {
    // per array q in twice (the forward sweep, then the recomputed chunks) and out once
    rc = run("exact_shift", 2.0 * 3.0 * bytes_field(), [&] {
      return launch_exact_shift(S.D, 2, qr, fr, xr, S.d_f0, S.fc, S.d_der, S.PX, S.PZ, sx0, sz0, sx1, sz1,
                                S.stream);
    });
    if (rc) return rc;
  }
  // serial (the kernel timing pass, or HST_SERIAL=1 in the environment for an
  // external profiler): on the main stream after the rhs half, each kernel
  // timed alone
  const bool serial = S.prof || S.serial;
  hipStream_t st = serial ? S.stream : S.side;
  if (!serial) HIPCHK(hipStreamWaitEvent(S.side, S.ev_fork, 0));
  // per array q in twice and out once, FP32 in the mixed form; in one sweep
  // (k_shift_window) in once and out once
  rc = run_on(st, "exact_shift_old", 2.0 * ((S.shw ? 2.0 : 3.0) * os) / 16.0 * bytes_field(), [&] {
    if (S.shw)
      return launch_shift_window(S.D, 2, qo, fo, S.XRo, S.d_f0, S.fc, S.d_der, S.d_sig, S.PX, S.PZ, sx0, sz0, sx1,
                                 sz1, st);
    return launch_exact_shift(S.D, 2, qo, fo, S.XRo, S.d_f0, S.fc, S.d_der, S.PX, S.PZ, sx0, sz0, sx1, sz1, st);
  });
  if (rc) return rc;
  if (!serial) {
    HIPCHK(hipEventRecord(S.ev_join, S.side));
    S.side_pending = true;
  }
  return 0;
}

static int shear_shift(double dt_sub, double time) {
  if (S.deck.exact_shift) return exact_shift(dt_sub, time);
  const double dgx = S.deck.S * dt_sub, dgz = s2_integral(time, time + dt_sub);
  return run("phase_shift", (4.0 * 16.0 + 4.0 * (S.mixed ? 8.0 : 16.0)) / 16.0 * bytes_field(), [&] {
    return launch_phase4(S.D, S.mixed, S.F[HST_RHS1], S.F[HST_RHS2], S.F[HST_OLD1], S.F[HST_OLD2], S.d_y, dgx, dgz,
                         S.stream);
  });
}

static int ghosts(int nf, double2 *const *f, double sx, double sz) {
  return run("ghosts", nf * 8.0 * S.D.ncol * 16.0, [&] { return launch_ghosts(S.D, nf, f, sx, sz, S.stream); });
}

// The factor tables for lambda (D2V, ETA), or null when they do not fit.
static const void *const *penta_tables(double lambda) {
  if (S.ptab_off) return nullptr;
  for (int k = 0; k < 3; k++)
    if (S.ptab[k][0] && S.ptab_lam[k] == lambda) return S.ptab[k];
  const int k = S.ptab_next;
  const size_t tb = penta_table_bytes(S.D);
  if (!S.ptab[k][0]) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < 2 * tb + ((size_t)4 << 30)) {
      S.ptab_off = true;
      return nullptr;
    }
    if (hipMalloc(&S.ptab[k][0], tb) != hipSuccess || hipMalloc(&S.ptab[k][1], tb) != hipSuccess) {
      (void)hipGetLastError();
      S.ptab_off = true;
      return nullptr;
    }
    S.bytes_held += 2 * (int64_t)tb;
  }
  const int kinds[2] = {HST_KIND_D2V, HST_KIND_ETA};
  for (int j = 0; j < 2; j++) {
    int rc = run("penta_factor", (double)tb, [&] {
      return launch_penta_factor(S.D, kinds[j], lambda, S.d_der, S.ptab[k][j], S.stream);
    });
    if (rc) {
      S.ptab_off = true;
      return nullptr;
    }
  }
  S.ptab_lam[k] = lambda;
  S.ptab_next = (k + 1) % 3;
  return S.ptab[k];
}

// The one-sweep dv/dy (k_dy_window) leaves out, for a row emitted from its
// window, the eliminated rows more than dy_window_rows() above it, and for
// the rows it does not correct, the start (b1, b2) and border terms.  All
// three are set by the D0 factorisation alone, the same for every column:
// this computes the border terms' share s1, s2 of the first rows (the
// back substitution of -B1, -B2) and takes the kernel only where each left-out
// share is below 1e-16 of the values it multiplies.
static bool dy_window_setup(std::vector<double> &sig) {
  const int mi = S.D.ny - 2, W = dy_window_rows();
  if (mi < 2 * (W + 8) + W) return false;
  const std::vector<double> &f = S.f0.rows;
  auto U1 = [&](int i) { return f[(size_t)i * D0Factor::NCOL + D0Factor::U1]; };
  auto U2 = [&](int i) { return f[(size_t)i * D0Factor::NCOL + D0Factor::U2]; };
  // the back substitution of sg * add (or of zero) from the start (s1, s2)
  auto back = [&](std::vector<double> &x, double s1, double s2, const std::vector<double> *add, double sg) {
    x.assign((size_t)mi + 2, 0.0);
    x[mi] = s1;
    x[mi + 1] = s2;
    for (int i = mi - 1; i >= 0; i--)
      x[i] = (add ? sg * (*add)[i] : 0.0) - U1(i) * x[i + 1] - U2(i) * x[i + 2];
  };
  std::vector<double> b1(mi), b2(mi);
  for (int i = 0; i < mi; i++) {
    b1[i] = f[(size_t)i * D0Factor::NCOL + D0Factor::B1];
    b2[i] = f[(size_t)i * D0Factor::NCOL + D0Factor::B2];
  }
  std::vector<double> s1, s2, t1, t2;
  back(s1, 0.0, 0.0, &b1, -1.0);
  back(s2, 0.0, 0.0, &b2, -1.0);
  back(t1, 1.0, 0.0, nullptr, 0.0);
  back(t2, 0.0, 1.0, nullptr, 0.0);
  double worst = 0.0;
  // border terms beyond the corrected rows; start terms below the exact tail
  for (int i = W; i < mi; i++) worst = std::max(worst, std::fabs(s1[i]) + std::fabs(s2[i]));
  for (int i = 0; i < mi - W - 8; i++) worst = std::max(worst, std::fabs(t1[i]) + std::fabs(t2[i]));
  // an eliminated row's share of the rows more than W below it
  for (int j = 0; j < mi; j++) {
    double h1 = 1.0, h2 = 0.0, tail = 0.0;  // h at rows i + 1, i + 2 of the impulse at row j
    for (int i = j - 1, d = 1; i >= 0 && d <= W + 64; i--, d++) {
      const double h = -U1(i) * h1 - U2(i) * h2;
      if (d > W) tail += std::fabs(h);
      h2 = h1;
      h1 = h;
    }
    worst = std::max(worst, tail);
  }
  if (!(worst < 1e-16)) return false;
  sig.resize(2 * (size_t)W);
  for (int i = 0; i < W; i++) {
    sig[2 * i] = s1[i];
    sig[2 * i + 1] = s2[i];
  }
  return true;
}

// after the solves: dv/dy, and u, w from it and eta (compact: dv/dy into V3,
// eta left in V1), then the ghost rows of both at the shifts (sx, sz)
static int recover(bool compact, double sx, double sz) {
  int rc;
  if (compact) {
    if (S.dyw)
      rc = run("dvdy", 2.0 * bytes_field(), [&] {
        return launch_dy_window(S.D, S.F[HST_V2], S.F[HST_V3], S.d_f0, S.fc, S.d_der, S.d_sig, sx, sz, S.stream);
      });
    else
      rc = run("dvdy", 4.0 * bytes_field(), [&] {
        return launch_d0solve(S.D, 1, 0, S.F[HST_V2], S.F[HST_V3], nullptr, S.XR[0], S.d_f0, S.fc, S.d_der,
                              S.d_dyl, sx, sz, nullptr, nullptr, S.stream);
      });
  } else {
    rc = run("dy_recover", 6.0 * bytes_field(), [&] {
      return launch_d0solve(S.D, 1, 1, S.F[HST_V2], S.F[HST_V1], S.F[HST_V3], S.XR[0], S.d_f0, S.fc, S.d_der,
                            S.d_dyl, sx, sz, nullptr, nullptr, S.stream);
    });
  }
  if (rc) return rc;
  double2 *g13[2] = {S.F[HST_V1], S.F[HST_V3]};
  if ((rc = ghosts(2, g13, sx, sz))) return rc;
  S.compact = compact;
  return 0;
}

static int restore_uw() {
  if (!S.compact) return 0;
  S.compact = false;
  return run("uw_from_eta", 4.0 * S.D.NYg * S.D.ncol * 16.0,
             [&] { return launch_uw_from_eta(S.D, S.F[HST_V1], S.F[HST_V3], S.stream); });
}

static int linsolve(double lambda, double time, bool compact = false) {
  double sx, sz;
  shear_shifts(time, sx, sz);
  const int kinds[2] = {HST_KIND_D2V, HST_KIND_ETA};
  const double2 *src[2] = {S.F[HST_RHS2], S.F[HST_RHS1]};
  double2 *dst[2] = {S.F[HST_V2], S.F[HST_V1]};
  const void *const *tabs = penta_tables(lambda);
  // algorithmic bytes of a line solve: rhs in, the eliminated rhs out and back in, solution out
  int rc = run("penta_d2v_eta", 2.0 * 4.0 * bytes_field(), [&] {
    return launch_penta(S.D, 2, kinds, lambda, src, dst, S.pws, tabs, S.d_der, sx, sz, S.stream);
  });
  if (rc) return rc;
  double2 *g2[1] = {S.F[HST_V2]};
  if ((rc = ghosts(1, g2, sx, sz))) return rc;
  return recover(compact, sx, sz);
}

// The exact shift of rhs1, rhs2 from t0 to t1 and linsolve at t1 with the
// factor tables tabs: one fused pass for both (the shifted right-hand sides
// stay in registers), then the ghosts of v, dv/dy with u and w, their ghosts.
static int shift_and_linsolve(double lambda, double t0, double t1, const void *const *tabs, bool compact) {
  double sx0, sz0, sx1, sz1;
  shear_shifts(t0, sx0, sz0);
  shear_shifts(t1, sx1, sz1);
  double2 *xrs[2] = {S.XR[1], S.XR[0]}, *xrp[2] = {S.pws[0].XR, S.pws[1].XR};
  const void *tb[2] = {tabs[0], tabs[1]};
  // per system: q in, the shift's eliminated rows out and in, the solve's out and in, the solution out
  int rc = run("shift_penta", 2.0 * 6.0 * bytes_field(), [&] {
    return launch_shift_penta(S.D, S.F[HST_RHS2], S.F[HST_RHS1], xrs, xrp, tb, S.F[HST_V2], S.F[HST_V1], S.d_f0r,
                              S.fcr, S.d_der, S.PX, S.PZ, sx0, sz0, sx1, sz1, lambda, S.stream);
  });
  if (rc) return rc;
  double2 *g2[1] = {S.F[HST_V2]};
  if ((rc = ghosts(1, g2, sx1, sz1))) return rc;
  return recover(compact, sx1, sz1);
}

// the main stream waits for the oldrhs shift of the side stream
static int join_side() {
  if (!S.side_pending) return 0;
  S.side_pending = false;
  HIPCHK(hipStreamWaitEvent(S.stream, S.ev_join, 0));
  return 0;
}

static int check_live_nojoin() {
  if (!S.live) return fail("hst_gpu_init has not been called");
  return 0;
}

// the state as the reference holds it: u, v, w in V1..V3
static int check_live() {
  if (int rc = check_live_nojoin()) return rc;
  if (int rc = join_side()) return rc;
  return restore_uw();
}

// for calls that do not touch the velocity: the time step's form is kept
static int check_live_keep() {
  if (int rc = check_live_nojoin()) return rc;
  return join_side();
}

static int check_field(int f) {
  if (f < 1 || f > 8) return fail("field id out of range (1..8)");
  return 0;
}

// a field the kernels take as FP64 complex (every one but oldrhs in the mixed form)
static int check_field64(int f) {
  if (int rc = check_field(f)) return rc;
  if (field_f32(f)) return fail("oldrhs (fields 6, 7) is held in FP32 in the mixed form; use it through the step");
  return 0;
}

static void release() {
  if (S.stream) (void)hipStreamSynchronize(S.stream);
  if (S.side) (void)hipStreamSynchronize(S.side);
  for (int i = 1; i <= 8; i++) if (S.F[i]) (void)hipFree(S.F[i]);
  void *ptrs[] = {S.P, S.Hv, S.Hp, S.ws, S.d_der, S.d_y, S.d_dyl, S.d_f0, S.d_part, S.d_f0r, S.d_sig};
  for (void *p : ptrs) if (p) (void)hipFree(p);
  for (auto &t : S.ptab)
    for (void *p : t) if (p) (void)hipFree(p);
  if (S.tables) free_fft_tables(S.tables);
  for (auto &p : S.pending) { S.pool.push_back(p.a); S.pool.push_back(p.b); }
  for (auto e : S.pool) (void)hipEventDestroy(e);
  if (S.stream) (void)hipStreamDestroy(S.stream);
  if (S.side) (void)hipStreamDestroy(S.side);
  if (S.ev_fork) (void)hipEventDestroy(S.ev_fork);
  if (S.ev_join) (void)hipEventDestroy(S.ev_join);
  for (double2 *p : S.XRo) if (p) (void)hipFree(p);
  if (S.PX) (void)hipFree(S.PX);
  if (S.exz_acc) (void)hipFree(S.exz_acc);
  for (void *p : {(void *)S.d_gdamp, (void *)S.d_fy, (void *)S.d_inlayer})
    if (p) (void)hipFree(p);
  if (S.PZ) (void)hipFree(S.PZ);
  if (S.h_part) (void)hipHostFree(S.h_part);
  S = State();
  // an attached closure belongs to the state released here: the next state
  // (perhaps on another grid) steps without it until it is attached again
  set_stress_hook(nullptr, nullptr);
}

}  // namespace hst

using namespace hst;

extern "C" {

int hst_gpu_last_error(char *buf, int len) {
  std::lock_guard<std::mutex> lk(g_error_mutex);
  if (!buf || len <= 0) return 1;
  std::snprintf(buf, (size_t)len, "%s", g_error.c_str());
  return 0;
}

int hst_gpu_set_chunk_rows(int rows) {
  if (rows < 0) return fail("chunk rows must be >= 0");
  g_chunk_request = rows;
  return 0;
}

int hst_gpu_memory(int64_t *bytes, int *chunk_rows) {
  if (bytes) *bytes = S.bytes_held;
  if (chunk_rows) *chunk_rows = S.chunk;
  return 0;
}

int hst_gpu_init(const hst_deck *deck, int precision) {
  if (!deck) return fail("deck is NULL");
  if (S.live) release();
  std::string err = check_deck(*deck);
  if (!err.empty()) return fail(err);
  if (precision != HST_FP64 && precision != HST_MIXED) return fail("precision must be HST_FP64 or HST_MIXED");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device");
  S.deck = *deck;
  S.mixed = precision == HST_MIXED;
  grid_setup(S.deck, S.g, S.y, S.dyl, S.der);
  if (!fft_size_supported(S.g.nzd)) return fail("unsupported z transform size nzd = " + std::to_string(S.g.nzd));
  if (!fft_size_supported(S.g.nxd)) return fail("unsupported x transform size nxd = " + std::to_string(S.g.nxd));
  S.D = make_dims(*deck, S.g);
  Dims &D = S.D;
  d0_factor(S.der, D.ny, S.f0);
  S.fc = d0_const(S.f0);
  {
    // the reversed line: row i of it is row ny-1-i, its stencils mirrored
    std::vector<double> dr(S.der.size());
    const int ny = D.ny;
    for (int n = 0; n < 4; n++)
      for (int j = -2; j <= 2; j++)
        for (int i = 0; i < ny; i++)
          dr[(size_t)i + (size_t)ny * (n + 4 * (j + 2))] = S.der[(size_t)(ny - 1 - i) + (size_t)ny * (n + 4 * (-j + 2))];
    d0_factor(dr, ny, S.f0r);
    S.fcr = d0_const(S.f0r);
  }

  HIPCHK(hipStreamCreate(&S.stream));
  HIPCHK(hipStreamCreate(&S.side));
  HIPCHK(hipEventCreateWithFlags(&S.ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&S.ev_join, hipEventDisableTiming));
  if (const char *e = std::getenv("HST_SERIAL")) S.serial = std::atoi(e) != 0;
  // HST_NO_TABLES=1: no factor tables (the path taken when they do not fit in
  // device memory): the exact shift of rhs runs apart from the line solves,
  // which eliminate their rows through the full workspace
  if (const char *e = std::getenv("HST_NO_TABLES")) S.ptab_off = std::atoi(e) != 0;
  const size_t nf = (size_t)D.NYg * D.ncol;
  const size_t cs = S.mixed ? sizeof(float2) : sizeof(double2);
  int64_t held = 0;
  for (int i = 1; i <= 8; i++) {
    // oldrhs holds the explicit nonlinear terms: FP32 in the mixed form, like the products
    const size_t es = field_f32(i) ? sizeof(float2) : sizeof(double2);
    HIPCHK(hipMalloc(&S.F[i], nf * es));
    HIPCHK(hipMemset(S.F[i], 0, nf * es));
    held += nf * es;
  }
  HIPCHK(hipMalloc(&S.P, 6 * nf * cs));
  HIPCHK(hipMemset(S.P, 0, 6 * nf * cs));
  held += 6 * nf * cs;
  // solver workspace: two per-column sweeps (48 B per row each) or four D0
  // sweeps (16 B), laid out with the padded row stride D.ncs (the fused shift
  // and solve indexes its workspace with it; the other kernels, with ncol,
  // stay inside each view)
  const size_t rows = (size_t)(D.ny - 2) * D.ncs;
  S.ws_bytes = std::max(2 * rows * 48, nf * sizeof(double2));
  HIPCHK(hipMalloc(&S.ws, S.ws_bytes));
  held += S.ws_bytes;
  char *w = (char *)S.ws;
  for (int k = 0; k < 2; k++) {
    char *b = w + k * rows * 48;
    S.pws[k].U1 = (double *)b;
    S.pws[k].U2 = (double *)(b + rows * 8);
    S.pws[k].B1 = (double *)(b + rows * 16);
    S.pws[k].B2 = (double *)(b + rows * 24);
    S.pws[k].XR = (double2 *)(b + rows * 32);
  }
  for (int k = 0; k < 4; k++) S.XR[k] = (double2 *)(w + k * rows * 16);
  if (deck->exact_shift) {
    for (int k = 0; k < 2; k++) {
      HIPCHK(hipMalloc(&S.XRo[k], rows * sizeof(double2)));
      held += rows * sizeof(double2);
    }
  }
  HIPCHK(hipMalloc(&S.exz_acc, (size_t)D.ncol * sizeof(double)));
  HIPCHK(hipMemset(S.exz_acc, 0, (size_t)D.ncol * sizeof(double)));
  // physics hooks: gdamp (zero until the first damping_update), fy, and the
  // rows inside |y - ly/2| < 8 delta (stokes_setup, src/hst_stokes.f90:76-86)
  HIPCHK(hipMalloc(&S.d_gdamp, (size_t)D.ncol * sizeof(double)));
  HIPCHK(hipMemset(S.d_gdamp, 0, (size_t)D.ncol * sizeof(double)));
  HIPCHK(hipMalloc(&S.d_fy, (size_t)D.NYg * sizeof(double)));
  HIPCHK(hipMemset(S.d_fy, 0, (size_t)D.NYg * sizeof(double)));
  S.inlayer.assign(D.ny, 0);
  if (stokes_active())
    for (int iy = 0; iy < D.ny; iy++)
      S.inlayer[iy] = std::fabs(S.y[iy + 2] - 0.5 * S.deck.ly) < 8.0 * S.deck.sl_delta ? 1 : 0;
  HIPCHK(hipMalloc(&S.d_inlayer, (size_t)D.ny * sizeof(int)));
  HIPCHK(hipMemcpy(S.d_inlayer, S.inlayer.data(), (size_t)D.ny * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&S.PX, (size_t)D.ny * D.NXm * sizeof(double2)));
  HIPCHK(hipMalloc(&S.PZ, (size_t)D.ny * D.NZm * sizeof(double2)));
  // transform chunk: rows of the physical-space buffers Hv (3 fields) and Hp (6)
  const size_t per_row = 9 * (size_t)D.nzd * D.NXp * cs;
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  int chunk = D.NYg;
  if (g_chunk_request > 0) chunk = std::min(g_chunk_request, D.NYg);
  else {
    const size_t budget = free_b > (size_t)4 << 30 ? free_b - ((size_t)4 << 30) : free_b / 2;
    if ((size_t)chunk * per_row > budget) chunk = (int)std::max<size_t>(1, budget / per_row);
  }
  S.chunk = chunk;
  HIPCHK(hipMalloc(&S.Hv, 3 * (size_t)chunk * D.nzd * D.NXp * cs));
  HIPCHK(hipMalloc(&S.Hp, 6 * (size_t)chunk * D.nzd * D.NXp * cs));
  held += 9 * (size_t)chunk * D.nzd * D.NXp * cs;
  // device tables
  std::vector<double> derd((size_t)D.ny * 20);
  for (int iy = 0; iy < D.ny; iy++)
    for (int n = 0; n < 4; n++)
      for (int j = 0; j < 5; j++) derd[(size_t)iy * 20 + n * 5 + j] = S.der[(size_t)iy + (size_t)D.ny * (n + 4 * j)];
  HIPCHK(hipMalloc(&S.d_der, derd.size() * sizeof(double)));
  HIPCHK(hipMemcpy(S.d_der, derd.data(), derd.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&S.d_y, S.y.size() * sizeof(double)));
  HIPCHK(hipMemcpy(S.d_y, S.y.data(), S.y.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&S.d_dyl, S.dyl.size() * sizeof(double)));
  HIPCHK(hipMemcpy(S.d_dyl, S.dyl.data(), S.dyl.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&S.d_f0, S.f0.rows.size() * sizeof(double)));
  HIPCHK(hipMemcpy(S.d_f0, S.f0.rows.data(), S.f0.rows.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&S.d_f0r, S.f0r.rows.size() * sizeof(double)));
  HIPCHK(hipMemcpy(S.d_f0r, S.f0r.rows.data(), S.f0r.rows.size() * sizeof(double), hipMemcpyHostToDevice));
  {
    std::vector<double> sig;
    auto off = [](const char *name) {
      const char *e = std::getenv(name);
      return e && std::atoi(e) != 0;
    };
    if (dy_window_setup(sig)) {
      HIPCHK(hipMalloc(&S.d_sig, sig.size() * sizeof(double)));
      HIPCHK(hipMemcpy(S.d_sig, sig.data(), sig.size() * sizeof(double), hipMemcpyHostToDevice));
      S.dyw = !off("HST_NO_DYWIN");
      S.shw = !off("HST_NO_SHWIN");
    }
  }
  S.part_cap = std::max<size_t>((size_t)D.NYg * D.nzd, (size_t)((D.ncol + 255) / 256) * NSTAT + 64);
  HIPCHK(hipMalloc(&S.d_part, S.part_cap * sizeof(double)));
  HIPCHK(hipHostMalloc((void **)&S.h_part, S.part_cap * sizeof(double), hipHostMallocDefault));
  if (make_fft_tables(D, S.mixed, &S.tables)) return fail("twiddle tables");
  if (int rc = scratch_attach(nullptr)) return rc;
  S.bytes_held = held;
  S.live = true;
  return 0;
}

int hst_gpu_free(void) {
  release();
  return 0;
}

int hst_gpu_sync(void) {
  if (int rc = check_live_keep()) return rc;
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int hst_gpu_grid(hst_grid_info *info, double *y, double *dyl, double *der) {
  if (int rc = check_live_keep()) return rc;
  if (info) *info = S.g;
  if (y) std::memcpy(y, S.y.data(), S.y.size() * sizeof(double));
  if (dyl) std::memcpy(dyl, S.dyl.data(), S.dyl.size() * sizeof(double));
  if (der) std::memcpy(der, S.der.data(), S.der.size() * sizeof(double));
  return 0;
}

int hst_gpu_upload_field(int field, const double *host) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field(field)) return rc;
  const size_t nf = (size_t)S.D.NYg * S.D.ncol;
  double2 *stage = (double2 *)S.ws, *wide = stage + nf;
  HIPCHK(hipMemcpyAsync(stage, host, nf * sizeof(double2), hipMemcpyHostToDevice, S.stream));
  int rc = launch_to_device_layout(S.D, stage, field_f32(field) ? wide : S.F[field], S.stream);
  if (!rc && field_f32(field)) rc = launch_narrow(wide, (float2 *)S.F[field], nf, S.stream);
  if (rc) return hip_fail("layout", (hipError_t)rc);
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int hst_gpu_download_field(int field, double *host) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field(field)) return rc;
  const size_t nf = (size_t)S.D.NYg * S.D.ncol;
  double2 *stage = (double2 *)S.ws, *wide = stage + nf;
  int rc = 0;
  if (field_f32(field)) rc = launch_widen((const float2 *)S.F[field], wide, nf, S.stream);
  if (!rc) rc = launch_to_ref_layout(S.D, field_f32(field) ? wide : S.F[field], stage, S.stream);
  if (rc) return hip_fail("layout", (hipError_t)rc);
  HIPCHK(hipMemcpyAsync(host, stage, nf * sizeof(double2), hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int hst_gpu_zero_field(int field) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field(field)) return rc;
  const size_t es = field_f32(field) ? sizeof(float2) : sizeof(double2);
  HIPCHK(hipMemsetAsync(S.F[field], 0, (size_t)S.D.NYg * S.D.ncol * es, S.stream));
  return 0;
}

int hst_gpu_fill_ghosts(int field, double shift_x, double shift_z) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field64(field)) return rc;
  double2 *f[1] = {S.F[field]};
  return ghosts(1, f, shift_x, shift_z);
}

int hst_gpu_nonlinear(double *cfl) {
  if (int rc = check_live()) return rc;
  return nonlinear_pass(0, cfl != nullptr, cfl);
}

int hst_gpu_compute_cfl(double *cfl) {
  if (int rc = check_live()) return rc;
  double c = 0.0;
  int rc = nonlinear_pass(2, true, &c);
  if (cfl) *cfl = c;
  return rc;
}

int hst_gpu_roundtrip(void) {
  if (int rc = check_live()) return rc;
  return nonlinear_pass(1, false, nullptr);
}

int hst_gpu_product_to_field(int p, int field) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field64(field)) return rc;
  if (p < 1 || p > 6) return fail("product index out of range (1..6)");
  int rc = launch_product_to_field(S.D, S.mixed, P_of(p - 1), S.F[field], S.stream);
  if (rc) return hip_fail("product_to_field", (hipError_t)rc);
  return 0;
}

// ------------------------------------------------------------------ per-field transforms of host buffers

// A device buffer for one call of the per-field entry points (they run
// outside the time step, on the reference's buffer shapes).
struct TmpBuf {
  void *p = nullptr;
  ~TmpBuf() {
    if (p) (void)hipFree(p);
  }
};

// exp(-2 pi i k / den) for k = 0 .. n-1, FP64, into a new device buffer
static int twiddles64(TmpBuf &b, int n, double den) {
  std::vector<double2> h((size_t)n);
  for (int k = 0; k < n; k++) {
    const double a = -TWOPI * (double)k / den;
    h[(size_t)k] = make_double2(std::cos(a), std::sin(a));
  }
  HIPCHK(hipMalloc(&b.p, (size_t)n * sizeof(double2)));
  HIPCHK(hipMemcpy(b.p, h.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
  return 0;
}

// nlines lines of n complex (host), transformed on the device in place
static int lines_on_device(double *lines, int64_t nlines, size_t line_bytes,
                           const std::function<int(double2 *)> &launch) {
  if (!lines || nlines < 0) return fail("lines is NULL or nlines < 0");
  if (nlines == 0) return 0;
  const size_t bytes = (size_t)nlines * line_bytes;
  TmpBuf d;
  HIPCHK(hipMalloc(&d.p, bytes));
  HIPCHK(hipMemcpyAsync(d.p, lines, bytes, hipMemcpyHostToDevice, S.stream));
  const int rc = launch((double2 *)d.p);
  if (rc) return hip_fail("per-field transform", (hipError_t)rc);
  HIPCHK(hipMemcpyAsync(lines, d.p, bytes, hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int hst_gpu_fft_z(double *lines, int64_t nlines, int sign) {
  if (int rc = check_live_keep()) return rc;
  if (sign != 1 && sign != -1) return fail("fft_z: sign must be +1 (IFT) or -1 (FFT)");
  const int n = S.D.nzd;
  TmpBuf tw;
  if (int rc = twiddles64(tw, n, n)) return rc;
  return lines_on_device(lines, nlines, (size_t)n * sizeof(double2),
                         [&](double2 *d) { return launch_cfft_lines(n, sign, d, nlines, tw.p, S.stream); });
}

int hst_gpu_fft_x(double *lines, int64_t nlines, int dir) {
  if (int rc = check_live_keep()) return rc;
  if (dir != 1 && dir != -1) return fail("fft_x: dir must be +1 (RFT, c2r) or -1 (HFT, r2c)");
  const int m = S.D.nxd;
  TmpBuf tw, twh;
  if (int rc = twiddles64(tw, m, m)) return rc;
  if (int rc = twiddles64(twh, m + 1, 2.0 * m)) return rc;
  return lines_on_device(lines, nlines, (size_t)(m + 1) * sizeof(double2), [&](double2 *d) {
    return launch_rfft_lines(m, dir, d, nlines, tw.p, twh.p, S.stream);
  });
}

int hst_gpu_products(int g, int64_t nlines, const double *vel, double *prod) {
  if (int rc = check_live_keep()) return rc;
  if (g != 1 && g != 2) return fail("products: group must be 1 or 2");
  if (!vel || !prod || nlines < 0) return fail("products: NULL buffer or nlines < 0");
  if (nlines == 0) return 0;
  const int lw = 2 * (S.D.nxd + 1);
  const int64_t plane = nlines * lw;
  const size_t bytes = (size_t)plane * 3 * sizeof(double);
  TmpBuf dv, dp;
  HIPCHK(hipMalloc(&dv.p, bytes));
  HIPCHK(hipMalloc(&dp.p, bytes));
  HIPCHK(hipMemcpyAsync(dv.p, vel, bytes, hipMemcpyHostToDevice, S.stream));
  HIPCHK(hipMemcpyAsync(dp.p, prod, bytes, hipMemcpyHostToDevice, S.stream));
  const int rc = launch_products((const double *)dv.p, (double *)dp.p, plane, lw, 2 * S.D.nxd, g, S.D.factor,
                                 S.stream);
  if (rc) return hip_fail("products", (hipError_t)rc);
  HIPCHK(hipMemcpyAsync(prod, dp.p, bytes, hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int hst_gpu_set_products(int g, const double *vvdz) {
  if (int rc = check_live_keep()) return rc;
  if (g != 1 && g != 2) return fail("set_products: group must be 1 or 2");
  if (!vvdz) return fail("set_products: vvdz is NULL");
  const Dims &D = S.D;
  const size_t nf = (size_t)D.NYg * D.ncol;
  const double2 *src = (const double2 *)vvdz;
  for (int m = 0; m < 3; m++) {
    // VVdz(jz, ix + 1, row, m): jz = izd(iz) + 1, izd(iz) = iz mod nzd
    std::vector<double2> h(nf);
    for (int row = 0; row < D.NYg; row++)
      for (int ix = 0; ix < D.NXm; ix++)
        for (int iz = -D.nz; iz <= D.nz; iz++) {
          const int jz = iz >= 0 ? iz : D.nzd + iz;
          h[(size_t)row * D.ncol + (size_t)ix * D.NZm + (iz + D.nz)] =
              src[(((size_t)m * D.NYg + row) * D.NXm + ix) * D.nzd + jz];
        }
    void *dst = (void *)P_of(3 * (g - 1) + m);
    if (S.mixed) {
      std::vector<float2> f(nf);
      for (size_t i = 0; i < nf; i++) f[i] = make_float2((float)h[i].x, (float)h[i].y);
      HIPCHK(hipMemcpyAsync(dst, f.data(), nf * sizeof(float2), hipMemcpyHostToDevice, S.stream));
      HIPCHK(hipStreamSynchronize(S.stream));
    } else {
      HIPCHK(hipMemcpyAsync(dst, h.data(), nf * sizeof(double2), hipMemcpyHostToDevice, S.stream));
      HIPCHK(hipStreamSynchronize(S.stream));
    }
  }
  return 0;
}

int hst_gpu_buildrhs_prepare(const double ode[3], double deltat, double time) {
  if (int rc = check_live()) return rc;
  // the substep's damping_update precedes buildrhs_prepare (src/hst_equations.f90:64)
  if (int rc = damping_update(time)) return rc;
  return buildrhs(ode, deltat, time, 0, false, 0.0, 0.0);
}

int hst_gpu_damping_mask(void) {
  if (int rc = check_live()) return rc;
  if (!(damping_active() && S.deck.damp_hard)) return 0;
  if (!S.theta_set) return fail("damping_mask: no damping table yet (hst_gpu_buildrhs_prepare builds it)");
  return run("damping_mask", 4.0 * bytes_field(), [&] {
    return launch_damping_mask(S.D, S.mixed, S.d_gdamp, S.F[HST_RHS1], S.F[HST_RHS2], S.F[HST_OLD1], S.F[HST_OLD2],
                               S.stream);
  });
}

int hst_gpu_project_divergence(double time) {
  if (int rc = check_live()) return rc;
  double sx, sz;
  shear_shifts(time, sx, sz);
  double2 *f[3] = {S.F[HST_V1], S.F[HST_V3], S.F[HST_V2]};
  int rc = ghosts(3, f, sx, sz);
  if (rc) return rc;
  rc = run("project", 4.0 * bytes_field(), [&] {
    return launch_project(S.D, S.F[HST_V1], S.F[HST_V2], S.F[HST_V3], S.d_der, S.stream);
  });
  if (rc) return rc;
  return ghosts(3, f, sx, sz);
}

int hst_gpu_stokes_apply(double time) {
  if (int rc = check_live()) return rc;
  return stokes_apply(time);
}

int hst_gpu_buildrhs(const double ode[3], int g) {
  if (int rc = check_live()) return rc;
  if (g != 1 && g != 2) return fail("group must be 1 or 2");
  return buildrhs(ode, 1.0, 0.0, g, false, 0.0, 0.0);
}

int hst_gpu_shear_shift(double dt_sub, double time) {
  if (int rc = check_live()) return rc;
  return shear_shift(dt_sub, time);
}

int hst_gpu_line_solve(int kind, double lambda, int src, int dst, double shift_x, double shift_z) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field64(src)) return rc;
  if (int rc = check_field64(dst)) return rc;
  if (src == dst) return fail("line_solve: src and dst must be different fields");
  if (kind == HST_KIND_D0 || kind == HST_KIND_DY) {
    return run(kind == HST_KIND_D0 ? "d0_solve" : "dy_solve", 4.0 * bytes_field(), [&] {
      return launch_d0solve(S.D, kind == HST_KIND_DY ? 1 : 0, 0, S.F[src], S.F[dst], nullptr, S.XR[0], S.d_f0, S.fc,
                            S.d_der, S.d_dyl, shift_x, shift_z, nullptr, nullptr, S.stream);
    });
  }
  if (kind < 1 || kind > 3) return fail("unknown system kind");
  const int kinds[1] = {kind};
  const double2 *s[1] = {S.F[src]};
  double2 *d[1] = {S.F[dst]};
  return run("penta", 4.0 * bytes_field(), [&] {
    return launch_penta(S.D, 1, kinds, lambda, s, d, S.pws, nullptr, S.d_der, shift_x, shift_z, S.stream);
  });
}

// [p, p + bytes) inside one device allocation of this process
static int device_range(const void *p, size_t bytes, const char *what) {
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  hipError_t e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
  if (e != hipSuccess || !base) {
    // the failed query leaves its status as the thread's last HIP error: take
    // it back, or the caller's next launch check would report it as its own
    (void)hipGetLastError();
    return fail(std::string(what) + " is not device memory of this process (hipMemGetAddressRange: " +
                hipGetErrorString(e) + ")");
  }
  if ((const char *)p + bytes > (const char *)base + size)
    return fail(std::string(what) + ": the allocation ends before the ny+4 rows of ncols columns");
  return 0;
}

int hst_gpu_line_solve_noalloc(int kind, double lambda, int col0, int ncols, const double *src, double *dst,
                               double shift_x, double shift_z) {
  if (int rc = check_live_keep()) return rc;
  if (kind < HST_KIND_D2V || kind > HST_KIND_DY) return fail("line_solve_noalloc: unknown system kind");
  if (ncols < 1 || col0 < 0 || col0 > S.D.ncol - ncols)
    return fail("line_solve_noalloc: columns " + std::to_string(col0) + " + " + std::to_string(ncols) +
                " outside 0.." + std::to_string(S.D.ncol - 1));
  if (!src || !dst) return fail("line_solve_noalloc: src or dst is NULL");
  const size_t nbytes = (size_t)S.D.NYg * (size_t)ncols * sizeof(double2);
  if ((const char *)src < (const char *)dst + nbytes && (const char *)dst < (const char *)src + nbytes)
    return fail("line_solve_noalloc: src and dst overlap");
  const size_t need = line_cols_ws_bytes(S.D, kind, ncols);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail("line_solve_noalloc: " + std::to_string(ncols) + " columns need " + std::to_string(need) +
                " bytes of workspace, above the fixed " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                " (HST_NOALLOC_MAX_BYTES): solve fewer columns per call");
  if (int rc = device_range(src, nbytes, "line_solve_noalloc: src")) return rc;
  if (int rc = device_range(dst, nbytes, "line_solve_noalloc: dst")) return rc;
  void *ws = nullptr;
  if (int rc = scratch_attach(&ws)) return rc;
  const int rc = launch_line_cols(S.D, kind, lambda, col0, ncols, (const double2 *)src, (double2 *)dst, ws, S.d_f0,
                                  S.fc, S.d_der, shift_x, shift_z, S.stream);
  if (rc) return hip_fail("line_solve_noalloc", (hipError_t)rc);
  return 0;
}

int hst_gpu_linsolve(double lambda, double time) {
  if (int rc = check_live()) return rc;
  return linsolve(lambda, time);
}

int hst_gpu_timestep(double deltat, double *time, double *cfl) {
  if (int rc = check_live_nojoin()) return rc;
  if (!time) return fail("time is NULL");
  // the substep divides by deltat (lambda = RK(i,0)/deltat): a deck with
  // deltat = 0 takes its step from the CFL number (new_timestep) first
  if (!(std::isfinite(deltat) && deltat > 0.0))
    return fail("timestep: deltat must be finite and > 0 (got " + std::to_string(deltat) + ")");
  double t = *time, c = 0.0;
  // between substeps (and calls) eta and dv/dy stand for u and w unless a
  // hook reads the velocity (closure, Stokes layer, damping) or the run is linear
  const bool compact = !S.deck.linear && g_hook == nullptr && !stokes_active() && !damping_active();
  if (!compact)
    if (int rc = restore_uw()) return rc;
  for (int i = 0; i < 3; i++) {
    const double *ode = RK[i];
    int rc = 0;
    const bool want = (i == 2) && cfl != nullptr;
    // the five product combinations the right-hand sides need; the six
    // products themselves when a closure is attached (it adds its modelled
    // stresses to them)
    const bool six = g_hook != nullptr;
    if (!S.deck.linear) rc = nonlinear_pass(six ? 0 : 3, want, &c, true);
    else if (want) rc = nonlinear_pass(2, true, &c, true);  // linear: the transform only for the CFL number
    if (rc) return rc;
    if (!S.deck.linear && six && (rc = g_hook(g_hook_ctx, t))) return rc;
    const double dt_sub = 2.0 / ode[0] * deltat;
    const bool fuse_shift = !S.deck.exact_shift;
    const double dgx = S.deck.S * dt_sub, dgz = s2_integral(t, t + dt_sub);
    if ((rc = join_side())) return rc;  // the previous substep's oldrhs shift
    if ((rc = damping_update(t))) return rc;
    rc = buildrhs(ode, deltat, t, S.deck.linear ? 0 : (six ? 3 : (S.compact ? 5 : 4)), fuse_shift, dgx, dgz, true);
    if (rc) return rc;
    // with the factor tables the rhs half of the exact shift is fused with the solves
    const double lam = ode[0] / deltat;
    const void *const *tabs = S.deck.exact_shift ? penta_tables(lam) : nullptr;
    const bool fused = tabs != nullptr;
    if (S.deck.exact_shift && (rc = exact_shift_split(dt_sub, t, fused))) return rc;
    const double t0 = t;
    t = t + dt_sub;
    rc = fused ? shift_and_linsolve(lam, t0, t, tabs, compact) : linsolve(lam, t, compact);
    if (rc) return rc;
    if ((rc = stokes_apply(t))) return rc;
  }
  *time = t;
  if (cfl) {
    if (int rc = cfl_finish(&c)) return rc;
    *cfl = c;
  }
  return 0;
}

int hst_gpu_outstats(double time, double out[10]) {
  if (int rc = check_live()) return rc;
  if (!out) return fail("outstats: out is NULL");
  // the damping table as set up at the start, or as last rebuilt by a substep
  if (!S.theta_set)
    if (int rc = damping_update(time)) return rc;
  StatsArgs a;
  if (damping_active()) {
    a.gdamp = S.d_gdamp;
    a.gsoft = damping_soft_factor();
  }
  if (stokes_active()) {
    a.inlayer = S.d_inlayer;
    a.inc00 = 0;  // the (0,0) column is the imposed layer: mean flow
  }
  int nb = 0;
  int rc = run("stats", 3.0 * bytes_field(), [&] {
    return launch_stats(S.D, a, S.F[1], S.F[2], S.F[3], S.d_dyl, S.d_part, &nb, S.exz_acc, S.stream);
  });
  if (rc) return rc;
  std::vector<double> h((size_t)nb * NSTAT);
  HIPCHK(hipMemcpyAsync(h.data(), S.d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  double s[NSTAT] = {0};
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < NSTAT; k++) s[k] += h[(size_t)b * NSTAT + k];
  double sx, sz;
  shear_shifts(time, sx, sz);
  double eps = s[6], e_in = s[11], e_out = s[12];
  for (int c = 1; c <= 3; c++) {
    int np = 0;
    rc = run("dy_stats", 3.0 * bytes_field(), [&] {
      return launch_d0solve(S.D, 1, 2, S.F[c], nullptr, nullptr, S.XR[0], S.d_f0, S.fc, S.d_der, S.d_dyl, sx, sz,
                            S.d_part, &np, S.stream, a.inlayer, a.inc00 ? 0 : 1);
    });
    if (rc) return rc;
    std::vector<double> g((size_t)np * 3);
    HIPCHK(hipMemcpyAsync(g.data(), S.d_part, g.size() * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIPCHK(hipStreamSynchronize(S.stream));
    double grad = 0.0, g_in = 0.0, g_out = 0.0;
    for (int b = 0; b < np; b++) {
      grad += g[(size_t)b * 3];
      g_in += g[(size_t)b * 3 + 1];
      g_out += g[(size_t)b * 3 + 2];
    }
    eps += grad;
    e_in += g_in;
    e_out += g_out;
  }
  double uu = s[0], vv = s[1], ww = s[2], uv = s[3], vw = s[4], uw = s[5], mfx = s[7], mfz = s[8];
  if (a.inc00) {
    // squared box means off the (0,0) column's raw sums
    double ly_w = 0.0;
    for (double d : S.dyl) ly_w += d;
    uu = uu - mfx * mfx / ly_w;
    ww = ww - mfz * mfz / ly_w;
    uw = uw - mfx * mfz / ly_w;
    S.mean_acc += (mfx * mfx + mfz * mfz) / ly_w;
  }
  S.exz_n += 1;
  const double q2 = uu + vv + ww, inv_ly = 1.0 / S.deck.ly;
  out[0] = 0.5 * q2 * inv_ly;
  out[1] = S.g.ni * eps * inv_ly;
  out[2] = uv * inv_ly;
  out[3] = vw * inv_ly;
  out[4] = uu * inv_ly;
  out[5] = vv * inv_ly;
  out[6] = ww * inv_ly;
  out[7] = uw * inv_ly;
  out[8] = mfx;
  out[9] = mfz;
  // stokes_runtime.dat and damping_runtime.dat (src/hst_stats.f90:270-284)
  for (double &v : S.extra) v = 0.0;
  if (stokes_active()) {
    double l_in = 0.0, l_out = 0.0;
    for (int iy = 0; iy < S.D.ny; iy++) (S.inlayer[iy] == 1 ? l_in : l_out) += S.dyl[iy];
    S.extra[0] = 0.5 * s[10] / l_out;
    S.extra[1] = 0.5 * s[9] / l_in;
    S.extra[2] = S.g.ni * e_out / l_out;
    S.extra[3] = S.g.ni * e_in / l_in;
  }
  if (damping_active()) {
    S.extra[4] = damping_theta(time) * 180.0 / 3.141592653589793;
    S.extra[5] = s[15];
    S.extra[6] = 0.5 * s[13];
    S.extra[7] = 0.5 * s[14];
  }
  return 0;
}

int hst_gpu_outstats_extra(double out[8]) {
  if (int rc = check_live()) return rc;
  if (!out) return fail("outstats_extra: out is NULL");
  for (int k = 0; k < 8; k++) out[k] = S.extra[k];
  return 0;
}

int hst_gpu_spectrum(double *E, int32_t *nacc) {
  if (int rc = check_live()) return rc;
  if (!E || !nacc) return fail("spectrum: NULL argument");
  const size_t n = S.D.ncol;
  HIPCHK(hipMemcpyAsync(E, S.exz_acc, n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipMemsetAsync(S.exz_acc, 0, n * sizeof(double), S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  *nacc = S.exz_n;
  // average; the (0,0) column about its box mean; E = half the column energy
  // (spectrum_write, src/hst_stats.f90:110-134)
  if (S.exz_n > 0) {
    for (size_t c = 0; c < n; c++) E[c] /= S.exz_n;
    E[S.D.nz] -= S.mean_acc / S.exz_n;
  } else {
    for (size_t c = 0; c < n; c++) E[c] = 0.0;
  }
  for (size_t c = 0; c < n; c++) E[c] *= 0.5;
  S.exz_n = 0;
  S.mean_acc = 0.0;
  return 0;
}

int hst_gpu_pressure(double time, int dst) {
  if (int rc = check_live()) return rc;
  if (int rc = check_field64(dst)) return rc;
  if (dst == HST_V1 || dst == HST_V2 || dst == HST_V3 || dst == HST_RHS1)
    return fail("pressure: dst must not be u, v, w or rhs1 (the right-hand side's buffer)");
  if (S.deck.linear) return fail("pressure: the linear equations have no products");
  int rc = nonlinear_pass(0, false, nullptr);
  if (rc) return rc;
  const double ps = S.mixed ? 8.0 : 16.0;
  rc = run("pressure_rhs", (double)S.D.ny * S.D.ncol * (6.0 * ps + 16.0 + 16.0), [&] {
    return launch_pressure_rhs(S.D, S.mixed, S.P, S.F[HST_V2], S.F[HST_RHS1], S.d_der, s2_of(time), S.stream);
  });
  if (rc) return rc;
  double sx, sz;
  shear_shifts(time, sx, sz);
  const int kinds[1] = {HST_KIND_POISSON};
  const double2 *src[1] = {S.F[HST_RHS1]};
  double2 *dsts[1] = {S.F[dst]};
  rc = run("penta", 4.0 * bytes_field(), [&] {
    return launch_penta(S.D, 1, kinds, 0.0, src, dsts, S.pws, nullptr, S.d_der, sx, sz, S.stream);
  });
  if (rc) return rc;
  return run("pressure_mean", 0.0, [&] {
    return launch_pressure_mean(S.D, S.F[HST_RHS1], S.F[dst], S.d_dyl, S.deck.ly, S.stream);
  });
}

int hst_gpu_profile(int enable) {
  if (int rc = check_live_keep()) return rc;
  if (enable) {
    flush_profile();
    S.table.clear();
    S.order.clear();
    S.prof = true;
  } else {
    flush_profile();
    S.prof = false;
  }
  return 0;
}

int hst_gpu_profile_count(void) {
  flush_profile();
  return (int)S.order.size();
}

int hst_gpu_profile_entry(int i, char *name, int len, int64_t *launches, double *ms, double *bytes) {
  flush_profile();
  if (i < 0 || i >= (int)S.order.size()) return fail("profile entry out of range");
  const ProfEntry &e = S.table[S.order[i]];
  if (name && len > 0) std::snprintf(name, (size_t)len, "%s", S.order[i].c_str());
  if (launches) *launches = e.launches;
  if (ms) *ms = e.ms;
  if (bytes) *bytes = e.bytes;
  return 0;
}

}  // extern "C"
