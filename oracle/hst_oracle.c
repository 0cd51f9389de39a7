/*
 * hst_oracle.c -- FP64 CPU restatement of hst's RK3 step (one rank, npy = 1).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the reference the GPU library is compared
 * with; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it.  The product (libhstgpu.so) never calls it.
 *
 * It follows the reference's loops and operation order in the reference's own
 * field layout V(iy, iz, ix, c), iy = -2..ny+1 fastest.  Each function names
 * the file:line of the reference it restates (paths relative to the reference
 * repository root).  The FFTs are a plain Stockham radix-2/3/4 transform of our
 * own; the real transforms run as complex transforms of the Hermitian-extended
 * (c2r) or real (r2c) line, with FFTW's c2r semantics made explicit: the
 * imaginary part of the kx = 0 bin is ignored.
 *
 * Pinning (what checks each function; see tests/):
 *   whole step        tests/golden/reference/small.fld, small_s2x.fld (50 steps, 1e-10)
 *   line_solve        dense LU of the cyclic matrix on small lines; operator identity
 *   transforms        round trip; closed-form Taylor-Green decay
 *   shear_shift etc.  Kelvin-mode closed form (default and exact shift, S2 cases)
 *   der               polynomial exactness of each stencil
 *   initial field     integer hash restated exactly; divergence-free; the fixtures
 *   outstats          direct recomputation, box-mean removal (test_stats)
 *   pressure          Taylor-Green closed form (tests/test_pressure.f90: the two
 *                     nonzero modes to 1e-4, the rest to 1e-10); the shear term
 *                     against the inverse of the continuous operator on one mode
 *   column energy     sum over the columns = the energy of outstats; one mode
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/hst_gpu.h"

typedef double complex cplx;

#define PI 3.141592653589793
#define TWOPI 6.283185307179586

/* ------------------------------------------------------------------ FFT */

typedef struct {
  int n, nf, f[40];
  cplx *tw; /* tw[k] = exp(-2 pi i k / n) */
} fftplan;

static void plan_init(fftplan *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->f[p->nf++] = 3; m /= 3; }
  if (m != 1) abort();
  p->tw = (cplx *)malloc(sizeof(cplx) * n);
  for (int k = 0; k < n; k++) {
    double a = 2.0 * PI * (double)k / (double)n;
    p->tw[k] = cos(a) - I * sin(a);
  }
}

static void plan_free(fftplan *p) { free(p->tw); p->tw = NULL; }

/* Stockham, decimation in frequency.  sign -1: forward (FFTW_FORWARD),
 * +1: backward; unnormalised.  x holds the input and receives the output;
 * y is work space of the same length. */
static void fft_run(const fftplan *p, cplx *x, cplx *y, int sign) {
  const int n = p->n;
  int s = 1, len = n;
  cplx *a = x, *b = y;
  for (int st = 0; st < p->nf; st++) {
    const int r = p->f[st], L = len / r;
    for (int n1 = 0; n1 < L; n1++) {
      for (int q = 0; q < s; q++) {
        cplx c[4], d[4];
        for (int n2 = 0; n2 < r; n2++) c[n2] = a[q + s * (n1 + L * n2)];
        for (int k2 = 0; k2 < r; k2++) {
          cplx acc = 0;
          for (int n2 = 0; n2 < r; n2++) {
            int e = ((n2 * k2) % r) * (n / r);
            cplx w = p->tw[e];
            if (sign > 0) w = conj(w);
            acc += c[n2] * w;
          }
          d[k2] = acc;
        }
        for (int k2 = 0; k2 < r; k2++) {
          cplx w = p->tw[(n1 * k2 * s) % n];
          if (sign > 0) w = conj(w);
          b[q + s * k2 + s * r * n1] = d[k2] * w;
        }
      }
    }
    len = L;
    s *= r;
    cplx *t = a; a = b; b = t;
  }
  if (a != x) memcpy(x, a, sizeof(cplx) * n);
}

/* ------------------------------------------------------------------ state */

typedef struct hsto {
  hst_deck d;
  int nx, ny, nz, nxd, nzd, NXm, NZm, NYg;
  double lx, lz, dx, dy, dz, factor, ni;
  double *y;   /* y(-2:ny+1) -> y[iy+2] */
  double *dyl; /* dyl(0:ny-1) */
  double *der; /* der(0:ny-1, 0:3, -2:2) Fortran order */
  double *k2;  /* k2(-nz:nz, 0:nx) */
  cplx *f[9];  /* fields 1..8 as in hst_gpu.h, reference layout */
  cplx *prod;  /* 6 product spectra, reference layout */
  fftplan pz, px;
  double cfl;
  /* physics hooks (src/hst_stokes.f90, src/hst_damping.f90) */
  double *gdamp;    /* gdamp(-nz:nz, 0:nx) */
  double *fy;       /* fy(-2:ny+1) -> fy[iy+2] */
  int *inlayer;     /* inlayer(0:ny-1) */
  double theta_last;
  int theta_set;
  double extra[8];  /* stokes_runtime.dat, damping_runtime.dat values of the last outstats */
} hsto;

#define FIDX(o, ix, iz, iy) ((((size_t)(ix)) * (o)->NZm + (size_t)((iz) + (o)->nz)) * (o)->NYg + (size_t)((iy) + 2))
#define DER(o, iy, n, j) ((o)->der[(size_t)(iy) + (size_t)(o)->ny * ((size_t)(n) + 4 * (size_t)((j) + 2))])
#define K2(o, iz, ix) ((o)->k2[(size_t)((iz) + (o)->nz) + (size_t)(o)->NZm * (size_t)(ix)])
#define GD(o, iz, ix) ((o)->gdamp[(size_t)((iz) + (o)->nz) + (size_t)(o)->NZm * (size_t)(ix)])

/* RK_rai (src/hst_params.f90:108-111), column i = substep. */
static const double RK_rai[3][3] = {
    {120.0 / 32.0, 2.0, 0.0}, {120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0}, {120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0}};

/* fft_fit (src/hst_input.f90:107-115) */
static int fft_fit(int n) {
  int j = n;
  while (j % 2 == 0) j /= 2;
  return j == 1 || j == 3;
}

/* The power x**n of gfortran for a runtime integer exponent (binary powering). */
static double powi(double a, int b) {
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

/* ymap (src/hst_setup.f90:80-89) */
static double ymap(double ystretch, double xi) {
  if (ystretch <= 1.0e-10) return xi;
  if (xi <= 0.5) return 0.5 * tanh(2.0 * ystretch * xi) / tanh(ystretch);
  return 0.5 * (2.0 + tanh(2.0 * ystretch * (xi - 1.0)) / tanh(ystretch));
}

static int modi(int a, int p) {
  int r = a % p;
  return r < 0 ? r + p : r;
}

/* Fortran MODULO for reals, as gfortran evaluates it (fmod, then the sign fix). */
static double modulo_r(double a, double p) {
  double r = fmod(a, p);
  if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r += p;
  return r;
}

/* solve5 (src/hst_derivatives.f90:59-79): dense 5x5 with partial pivoting.
 * M is indexed M[i][j] as the Fortran M(i, j). */
static void solve5(const double Min[5][5], const double bin[5], double x[5]) {
  double M[5][5], b[5];
  memcpy(M, Min, sizeof(M));
  memcpy(b, bin, sizeof(b));
  for (int i = 0; i < 5; i++) {
    int p = i;
    double best = fabs(M[i][i]);
    for (int r = i + 1; r < 5; r++)
      if (fabs(M[r][i]) > best) { best = fabs(M[r][i]); p = r; }
    if (p != i) {
      for (int c = 0; c < 5; c++) { double t = M[i][c]; M[i][c] = M[p][c]; M[p][c] = t; }
      double t = b[i]; b[i] = b[p]; b[p] = t;
    }
    for (int j = i + 1; j < 5; j++) {
      double f = M[j][i] / M[i][i];
      for (int c = i; c < 5; c++) M[j][c] = M[j][c] - f * M[i][c];
      b[j] = b[j] - f * b[i];
    }
  }
  for (int i = 4; i >= 0; i--) {
    double s = 0.0;
    for (int c = i + 1; c < 5; c++) s += M[i][c] * x[c];
    x[i] = (b[i] - s) / M[i][i];
  }
}

/* setup_derivatives (src/hst_derivatives.f90:30-56) */
static void setup_derivatives(hsto *o) {
  for (int iy = 0; iy < o->ny; iy++) {
    double dyj[5], M[5][5], t[5], x[5];
    for (int j = 0; j < 5; j++) dyj[j] = o->y[iy + j - 2 + 2] - o->y[iy + 2];
    /* D4 */
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = powi(dyj[j], 4 - i);
    for (int i = 0; i < 5; i++) t[i] = 0.0;
    t[0] = 24.0;
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(o, iy, 3, j - 2) = x[j];
    /* D0 */
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = (double)((5 - i) * (6 - i) * (7 - i) * (8 - i)) * powi(dyj[j], 4 - i);
    for (int i = 0; i < 5; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += DER(o, iy, 3, j - 2) * powi(dyj[j], 8 - i);
      t[i] = s;
    }
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(o, iy, 0, j - 2) = x[j];
    /* D2 */
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = powi(dyj[j], 4 - i);
    for (int i = 0; i < 5; i++) t[i] = 0.0;
    for (int i = 0; i <= 2; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += DER(o, iy, 0, j - 2) * (double)(4 - i) * (double)(3 - i) * powi(dyj[j], 2 - i);
      t[i] = s;
    }
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(o, iy, 2, j - 2) = x[j];
    /* D1 */
    for (int i = 0; i < 5; i++) t[i] = 0.0;
    for (int i = 0; i <= 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += DER(o, iy, 0, j - 2) * (double)(4 - i) * powi(dyj[j], 3 - i);
      t[i] = s;
    }
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(o, iy, 1, j - 2) = x[j];
  }
}

/* ------------------------------------------------------------------ clock helpers */

/* s2_of (src/hst_derivatives.f90:85-96) */
static double s2_of(const hsto *o, double t) {
  const hst_deck *d = &o->d;
  if ((d->s2_amplitude == 0.0 && d->s2_mean == 0.0) || t < d->s2_start) return 0.0;
  if (d->s2_period > 0.0) return d->s2_mean + d->s2_amplitude * sin(TWOPI / d->s2_period * (t - d->s2_start));
  return d->s2_mean + d->s2_amplitude;
}

/* s2_integral (src/hst_derivatives.f90:98-112) */
static double s2_integral(const hsto *o, double t1, double t2) {
  const hst_deck *d = &o->d;
  if ((d->s2_amplitude == 0.0 && d->s2_mean == 0.0) || t2 <= d->s2_start) return 0.0;
  double a = t1 > d->s2_start ? t1 : d->s2_start, b = t2;
  if (d->s2_period > 0.0) {
    double omega = TWOPI / d->s2_period;
    return d->s2_mean * (b - a) + d->s2_amplitude / omega * (cos(omega * (a - d->s2_start)) - cos(omega * (b - d->s2_start)));
  }
  return (d->s2_mean + d->s2_amplitude) * (b - a);
}

/* shear_shifts (src/hst_derivatives.f90:123-128) */
void hsto_shear_shifts(const hsto *o, double t, double *sx, double *sz) {
  *sx = modulo_r(o->d.S * t * o->d.ly, o->lx);
  *sz = modulo_r(s2_integral(o, o->d.s2_start, t) * o->d.ly, o->lz);
}

double hsto_gamma_y(const hsto *o, double t) { return s2_integral(o, o->d.s2_start, t); }
double hsto_s2_of(const hsto *o, double t) { return s2_of(o, t); }

/* ------------------------------------------------------------------ set-up */

void hsto_free(hsto *o);

/* read_input's derived block (src/hst_input.f90:92-103), allocate_fields
 * (src/hst_setup.f90:42-71), setup_derivatives, init_fft's sizes. */
hsto *hsto_create(const hst_deck *d) {
  hsto *o = (hsto *)calloc(1, sizeof(hsto));
  o->d = *d;
  o->nx = d->nx; o->ny = d->ny; o->nz = d->nz;
  o->ni = 1.0 / d->re;
  o->lx = 2.0 * PI / d->alfa0;
  o->lz = 2.0 * PI / d->beta0;
  o->nxd = 3 * (o->nx + 1) / 2;
  while (!fft_fit(o->nxd)) o->nxd++;
  o->nzd = 3 * o->nz;
  while (!fft_fit(o->nzd)) o->nzd++;
  o->dx = PI / (d->alfa0 * o->nxd);
  o->dz = 2.0 * PI / (d->beta0 * o->nzd);
  o->dy = d->ly / o->ny;
  o->factor = 1.0 / (2.0 * o->nxd * o->nzd);
  o->NXm = o->nx + 1; o->NZm = 2 * o->nz + 1; o->NYg = o->ny + 4;
  o->y = (double *)malloc(sizeof(double) * o->NYg);
  o->dyl = (double *)malloc(sizeof(double) * o->ny);
  o->der = (double *)malloc(sizeof(double) * 20 * o->ny);
  o->k2 = (double *)malloc(sizeof(double) * o->NZm * o->NXm);
  for (int iy = -2; iy <= o->ny + 1; iy++) {
    int m = modi(iy, o->ny);
    o->y[iy + 2] = d->ly * ymap(d->ystretch, (double)m / o->ny) + d->ly * (double)((iy - m) / o->ny);
  }
  for (int iy = 0; iy < o->ny; iy++) o->dyl[iy] = 0.5 * (o->y[iy + 1 + 2] - o->y[iy - 1 + 2]);
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      double a = d->alfa0 * ix, b = d->beta0 * iz;
      K2(o, iz, ix) = a * a + b * b;
    }
  setup_derivatives(o);
  size_t nf = (size_t)o->NXm * o->NZm * o->NYg;
  for (int i = 1; i <= 8; i++) o->f[i] = (cplx *)calloc(nf, sizeof(cplx));
  o->prod = (cplx *)calloc(6 * nf, sizeof(cplx));
  plan_init(&o->pz, o->nzd);
  plan_init(&o->px, 2 * o->nxd);
  o->gdamp = (double *)calloc((size_t)o->NZm * o->NXm, sizeof(double));
  o->fy = (double *)calloc(o->NYg, sizeof(double));
  o->inlayer = (int *)calloc(o->ny, sizeof(int));
  /* stokes_setup (src/hst_stokes.f90:76-86) */
  if (d->sl_amplitude != 0.0)
    for (int iy = 0; iy < o->ny; iy++) o->inlayer[iy] = fabs(o->y[iy + 2] - 0.5 * d->ly) < 8.0 * d->sl_delta;
  return o;
}

void hsto_free(hsto *o) {
  if (!o) return;
  free(o->y); free(o->dyl); free(o->der); free(o->k2);
  for (int i = 1; i <= 8; i++) free(o->f[i]);
  free(o->prod);
  free(o->gdamp); free(o->fy); free(o->inlayer);
  plan_free(&o->pz); plan_free(&o->px);
  free(o);
}

void hsto_info(const hsto *o, hst_grid_info *g, double *y, double *dyl, double *der) {
  if (g) {
    g->nxd = o->nxd; g->nzd = o->nzd; g->nxm = o->NXm; g->nzm = o->NZm; g->nyg = o->NYg;
    g->ncol = o->NXm * o->NZm;
    g->lx = o->lx; g->lz = o->lz; g->dx = o->dx; g->dy = o->dy; g->dz = o->dz;
    g->factor = o->factor; g->ni = o->ni;
  }
  if (y) memcpy(y, o->y, sizeof(double) * o->NYg);
  if (dyl) memcpy(dyl, o->dyl, sizeof(double) * o->ny);
  if (der) memcpy(der, o->der, sizeof(double) * 20 * o->ny);
}

/* Field access for the tests: complex128 in the reference layout. */
double *hsto_field(hsto *o, int id) { return (double *)o->f[id]; }
double *hsto_products(hsto *o) { return (double *)o->prod; }

/* ------------------------------------------------------------------ initial field */

/* uniform_from_key (src/hst_initial.f90:139-159): integer-exact. */
double hsto_uniform_from_key(int seed, int component, int iy, int iz, int ix) {
  const int64_t MASK32 = 4294967295LL, MIX_A = 2146121005LL, MIX_B = 2032289749LL;
  int64_t h = (int64_t)seed;
  h = h + 2654435761LL * (int64_t)component;
  h = h + 2246822519LL * (int64_t)iy;
  h = h + 3266489917LL * (int64_t)iz;
  h = h + 668265263LL * (int64_t)ix;
  h = h & MASK32;
  h = h ^ (h >> 16); /* ishft of a non-negative value: logical = arithmetic */
  h = (h * MIX_A) & MASK32;
  h = h ^ (h >> 13);
  h = (h * MIX_B) & MASK32;
  h = h ^ (h >> 16);
  return (double)h * (1.0 / 4294967296.0);
}

/* potential (src/hst_initial.f90:113-133) */
static void potential(int seed, int ix, int iz, int m, cplx a[3]) {
  int jz = iz, jm = m, mirror = 0;
  if (ix == 0 && (iz < 0 || (iz == 0 && m < 0))) { jz = -iz; jm = -m; mirror = 1; }
  for (int c = 1; c <= 3; c++) {
    double r = hsto_uniform_from_key(seed, c, jm, jz, ix);
    double phase = TWOPI * hsto_uniform_from_key(seed, c + 3, jm, jz, ix);
    a[c - 1] = r * (cos(phase) + I * sin(phase));
    if (mirror) a[c - 1] = conj(a[c - 1]);
  }
}

/* generate_initial_field and generate_taylor_green (src/hst_initial.f90:27-108) */
void hsto_initial_field(hsto *o) {
  const hst_deck *d = &o->d;
  size_t nf = (size_t)o->NXm * o->NZm * o->NYg;
  for (int c = 1; c <= 3; c++) memset(o->f[c], 0, nf * sizeof(cplx));
  if (d->initial_condition == 1) {
    double ky = 2.0 * PI * d->tg_mode / d->ly;
    int ix = d->tg_mode;
    for (int iz = -d->tg_mode; iz <= d->tg_mode; iz += 2 * d->tg_mode)
      for (int iy = -2; iy <= o->ny + 1; iy++) {
        o->f[1][FIDX(o, ix, iz, iy)] = -I * d->tg_amplitude * cos(ky * o->y[iy + 2]) / 4.0;
        o->f[2][FIDX(o, ix, iz, iy)] = -d->tg_amplitude * sin(ky * o->y[iy + 2]) / 4.0;
      }
    return;
  }
  const int mmax = o->ny / 3;
  /* exp(i ky y) for every (m, iy): the loop body of :58-63 evaluates the same values */
  cplx *ph = (cplx *)malloc(sizeof(cplx) * (2 * mmax + 1) * o->NYg);
  for (int m = -mmax; m <= mmax; m++) {
    double ky = 2.0 * PI * m / d->ly;
    for (int iy = 0; iy < o->NYg; iy++) {
      double a = ky * o->y[iy];
      ph[(m + mmax) * o->NYg + iy] = cos(a) + I * sin(a);
    }
  }
#pragma omp parallel for schedule(dynamic)
  for (int ix = 0; ix <= o->nx; ix++) {
    double kx = d->alfa0 * ix;
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      double kz = d->beta0 * iz;
      if (ix == 0 && iz == 0) continue;
      for (int m = -mmax; m <= mmax; m++) {
        double ky = 2.0 * PI * m / d->ly;
        double k = sqrt(kx * kx + ky * ky + kz * kz);
        if (k == 0.0) continue;
        double env = powi(k / d->kpeak, 2) * exp(1.0 - powi(k / d->kpeak, 2));
        if (env < 1.0e-8) continue;
        cplx a[3], u[3];
        potential(d->seed, ix, iz, m, a);
        for (int c = 0; c < 3; c++) a[c] = d->amplitude * env * a[c];
        u[0] = I * (ky * a[2] - kz * a[1]);
        u[1] = I * (kz * a[0] - kx * a[2]);
        u[2] = I * (kx * a[1] - ky * a[0]);
        const cplx *p = ph + (m + mmax) * o->NYg;
        cplx *v1 = o->f[1] + FIDX(o, ix, iz, -2), *v2 = o->f[2] + FIDX(o, ix, iz, -2), *v3 = o->f[3] + FIDX(o, ix, iz, -2);
        for (int iy = 0; iy < o->NYg; iy++) {
          v1[iy] = v1[iy] + u[0] * p[iy];
          v2[iy] = v2[iy] + u[1] * p[iy];
          v3[iy] = v3[iy] + u[2] * p[iy];
        }
      }
    }
  }
  free(ph);
}

/* ------------------------------------------------------------------ ghosts */

/* wrap_rows (src/hst_mpi.f90:340-355) */
void hsto_fill_ghosts(hsto *o, int id, double sx, double sz) {
  cplx *f = o->f[id];
  const int ny = o->ny;
#pragma omp parallel for
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      double a = -(o->d.alfa0 * ix * sx + o->d.beta0 * iz * sz);
      cplx ph = cos(a) + I * sin(a);
      cplx *c = f + FIDX(o, ix, iz, -2); /* c[iy + 2] = f(iy) */
      c[ny + 2] = c[0 + 2] * ph;         /* f(ny)   = f(0) ph   */
      c[ny + 1 + 2] = c[1 + 2] * ph;     /* f(ny+1) = f(1) ph   */
      c[-1 + 2] = c[ny - 1 + 2] * conj(ph);
      c[-2 + 2] = c[ny - 2 + 2] * conj(ph);
    }
}

/* fill_ghosts(c) at the current time (src/hst_derivatives.f90:134-146) */
static void fill_ghosts_t(hsto *o, int id, double t) {
  double sx, sz;
  hsto_shear_shifts(o, t, &sx, &sz);
  hsto_fill_ghosts(o, id, sx, sz);
}

/* ------------------------------------------------------------------ line solve */

/* coef (src/hst_linsolve.f90:236-252) */
static inline double coef(const hsto *o, int sys, double lambda, double kk, int iy, int j) {
  const double ni = o->ni;
  switch (sys) {
    case HST_KIND_D2V:
      return lambda * (DER(o, iy, 2, j) - kk * DER(o, iy, 0, j)) -
             ni * (DER(o, iy, 3, j) - 2.0 * kk * DER(o, iy, 2, j) + kk * kk * DER(o, iy, 0, j));
    case HST_KIND_ETA:
      return lambda * DER(o, iy, 0, j) - ni * (DER(o, iy, 2, j) - kk * DER(o, iy, 0, j));
    case HST_KIND_POISSON:
      return DER(o, iy, 2, j) - kk * DER(o, iy, 0, j);
    default:
      return DER(o, iy, 0, j);
  }
}

/* rhs_row (src/hst_linsolve.f90:257-271); src points at row -2 of the line */
static inline cplx rhs_row(const hsto *o, int sys, const cplx *src, int iy) {
  if (sys == HST_KIND_DY) {
    cplx r = 0.0;
    for (int j = -2; j <= 2; j++) r = r + DER(o, iy, 1, j) * src[iy + j + 2];
    return r;
  }
  return src[iy + 2];
}

/* penta_forward + penta_backward with one slab (src/hst_linsolve.f90:282-483,
 * nslab = 1, islab = 0, m = ny, ny0 = 0) for one line. */
static void penta_line(const hsto *o, int sys, double lambda, double kk, cplx ph, const cplx *src, cplx *dst,
                       double *U1, double *U2, double *B1, double *B2, cplx *XR) {
  const int m = o->ny, mi = m - 2;
  double a[5], l, rp, t1, t2, p, q;
  double u11 = 0, u12 = 0, p1 = 0, q1 = 0, u21 = 0, u22 = 0, p2 = 0, q2 = 0;
  double w, w1 = 0, w2 = 0, v, v1 = 0, v2 = 0;
  double sp0 = 0, sq0 = 0, sp1 = 0, sq1 = 0;
  cplx bx, x1 = 0, x2 = 0, sx0 = 0, sx1 = 0;
  double rec[29];
  for (int i = 0; i < mi; i++) {
    for (int j = -2; j <= 2; j++) a[j + 2] = coef(o, sys, lambda, kk, i, j);
    p = 0.0; q = 0.0;
    if (i == 0) { p = a[0]; q = a[1]; }
    if (i == 1) q = a[0];
    bx = rhs_row(o, sys, src, i);
    if (i >= 2) {
      l = a[0];
      a[1] = a[1] - l * u21;
      a[2] = a[2] - l * u22;
      bx = bx - l * x2; p = p - l * p2; q = q - l * q2;
    }
    if (i >= 1) {
      l = a[1];
      a[2] = a[2] - l * u11;
      a[3] = a[3] - l * u12;
      bx = bx - l * x1; p = p - l * p1; q = q - l * q1;
    }
    rp = 1.0 / a[2];
    t1 = a[3] * rp; t2 = a[4] * rp; bx = bx * rp; p = p * rp; q = q * rp;
    U1[i] = t1; U2[i] = t2; B1[i] = p; B2[i] = q; XR[i] = bx;
    w = -(u11 * w1 + u22 * w2); if (i == 0) w = 1.0;
    v = -(u11 * v1 + u22 * v2); if (i == 1) v = 1.0;
    sx0 = sx0 + w * bx; sp0 = sp0 + w * p; sq0 = sq0 + w * q;
    sx1 = sx1 + v * bx; sp1 = sp1 + v * p; sq1 = sq1 + v * q;
    w2 = w1; w1 = w; v2 = v1; v1 = v;
    u21 = u11; u22 = u12; x2 = x1; p2 = p1; q2 = q1;
    u11 = t1; u12 = t2; x1 = bx; p1 = p; q1 = q;
  }
  rec[1] = creal(sx0); rec[2] = cimag(sx0); rec[3] = sp0; rec[4] = sq0;
  rec[5] = w2 * u22 + w1 * u11; rec[6] = w1 * u12;
  rec[7] = creal(sx1); rec[8] = cimag(sx1); rec[9] = sp1; rec[10] = sq1;
  rec[11] = v2 * u22 + v1 * u11; rec[12] = v1 * u12;
  bx = x2 - u21 * x1;
  rec[13] = creal(bx); rec[14] = cimag(bx); rec[15] = p2 - u21 * p1; rec[16] = q2 - u21 * q1;
  rec[17] = u22 - u21 * u11; rec[18] = -u21 * u12;
  rec[19] = creal(x1); rec[20] = cimag(x1); rec[21] = p1; rec[22] = q1;
  rec[23] = u11; rec[24] = u12;
  bx = rhs_row(o, sys, src, m - 2); rec[25] = creal(bx); rec[26] = cimag(bx);
  bx = rhs_row(o, sys, src, m - 1); rec[27] = creal(bx); rec[28] = cimag(bx);

  /* the 2x2 border system (:384-453) with s = t = 0 */
  cplx A11 = 0, A12 = 0, A21 = 0, A22 = 0, b1, b2, f;
  const cplx cph = conj(ph), php = ph;
  const int ga = m - 2;
  double c1m4 = coef(o, sys, lambda, kk, ga, -2), c1m3 = coef(o, sys, lambda, kk, ga, -1);
  double d11 = coef(o, sys, lambda, kk, ga, 0), d12 = coef(o, sys, lambda, kk, ga, 1);
  double c10 = coef(o, sys, lambda, kk, ga, 2);
  double c2m3 = coef(o, sys, lambda, kk, ga + 1, -2), d21 = coef(o, sys, lambda, kk, ga + 1, -1);
  double d22 = coef(o, sys, lambda, kk, ga + 1, 0), c20 = coef(o, sys, lambda, kk, ga + 1, 1);
  double c21 = coef(o, sys, lambda, kk, ga + 1, 2);
  cplx xm2 = rec[13] + I * rec[14], xm1 = rec[19] + I * rec[20], x0 = rec[1] + I * rec[2];
  b1 = (rec[25] + I * rec[26]) - c1m4 * xm2 - c1m3 * xm1 - c10 * php * x0;
  A11 = A11 - cph * (c1m4 * rec[15] + c1m3 * rec[21]);
  A12 = A12 - cph * (c1m4 * rec[16] + c1m3 * rec[22]);
  A11 = A11 + d11 - (c1m4 * rec[17] + c1m3 * rec[23]);
  A12 = A12 + d12 - (c1m4 * rec[18] + c1m3 * rec[24]);
  f = c10 * php;
  f = f * conj(ph);
  A11 = A11 - f * rec[3];
  A12 = A12 - f * rec[4];
  A11 = A11 - c10 * php * rec[5];
  A12 = A12 - c10 * php * rec[6];
  cplx xx1 = rec[7] + I * rec[8];
  b2 = (rec[27] + I * rec[28]) - c2m3 * xm1 - php * (c20 * x0 + c21 * xx1);
  A21 = A21 - cph * c2m3 * rec[21];
  A22 = A22 - cph * c2m3 * rec[22];
  A21 = A21 + d21 - c2m3 * rec[23];
  A22 = A22 + d22 - c2m3 * rec[24];
  f = php;
  f = f * conj(ph);
  A21 = A21 - f * (c20 * rec[3] + c21 * rec[9]);
  A22 = A22 - f * (c20 * rec[4] + c21 * rec[10]);
  A21 = A21 - php * (c20 * rec[5] + c21 * rec[11]);
  A22 = A22 - php * (c20 * rec[6] + c21 * rec[12]);
  cplx piv = 1.0 / A11;
  f = A21 * piv;
  A22 = A22 - f * A12;
  b2 = b2 - f * b1;
  b2 = b2 / A22;
  b1 = (b1 - A12 * b2) / A11;
  const cplx xb1 = b1, xb2 = b2;
  dst[m - 2 + 2] = xb1;
  dst[m - 1 + 2] = xb2;
  const cplx yb1 = cph * b1, yb2 = cph * b2;
  cplx xk, xk1 = xb1, xk2 = xb2;
  for (int i = m - 3; i >= 0; i--) {
    xk = XR[i] - B1[i] * yb1 - B2[i] * yb2 - U1[i] * xk1 - U2[i] * xk2;
    dst[i + 2] = xk;
    xk2 = xk1; xk1 = xk;
  }
}

/* line_solve (src/hst_linsolve.f90:147-233) with explicit shifts */
void hsto_line_solve(hsto *o, int sys, double lambda, int src, int dst, double sx, double sz) {
  const cplx *S = o->f[src];
  cplx *D = o->f[dst];
  const int ny = o->ny;
#pragma omp parallel
  {
    double *U1 = (double *)malloc(sizeof(double) * ny * 4);
    double *U2 = U1 + ny, *B1 = U2 + ny, *B2 = B1 + ny;
    cplx *XR = (cplx *)malloc(sizeof(cplx) * ny);
#pragma omp for collapse(2) schedule(static)
    for (int ix = 0; ix <= o->nx; ix++)
      for (int iz = -o->nz; iz <= o->nz; iz++) {
        cplx *d = D + FIDX(o, ix, iz, -2);
        if (ix == 0 && iz == 0 && sys == HST_KIND_D2V) {
          for (int iy = 0; iy < ny; iy++) d[iy + 2] = 0.0;
          continue;
        }
        if (ix == 0 && iz == 0 && sys == HST_KIND_POISSON) continue;
        double a = -(o->d.alfa0 * ix * sx + o->d.beta0 * iz * sz);
        cplx ph = cos(a) + I * sin(a);
        penta_line(o, sys, lambda, K2(o, iz, ix), ph, S + FIDX(o, ix, iz, -2), d, U1, U2, B1, B2, XR);
      }
    free(U1);
    free(XR);
  }
}

/* ------------------------------------------------------------------ transforms */

/* One y row of the transform round trip, all three velocity components:
 * assemble_vvdz + IFT + transpose + zero_vvdx_padding + RFT
 * (src/hst_transforms.f90:33-90, src/hst_fft.f90:195-223), the products
 * (build_products, :97-121; both groups, or u, v, w times factor for the
 * round-trip test when mode = 1), and HFT + transpose with truncation + FFT
 * (:125-139, src/hst_fft.f90:181-237, src/hst_mpi.f90:774-784).  cfl: the
 * row's max of |u|/dx + |v|/dyl + |w|/dz when requested (compute_cfl :163-179). */
static void row_nonlinear(hsto *o, int iy, int mode, double *cflrow, cplx *buf, cplx *work, double *phys) {
  const int nzd = o->nzd, nxd = o->nxd, nx = o->nx, nz = o->nz, N = 2 * nxd;
  const size_t plane = (size_t)nzd * N;
  /* phys: [3][nzd][N] real velocities of this row */
  for (int c = 0; c < 3; c++) {
    const cplx *V = o->f[c + 1];
    /* z pencils: VVx[iz_p][ix] after the z transform */
    cplx *VVx = buf; /* [nzd][nx+1] */
    cplx *line = work, *w2 = work + (nzd > N ? nzd : N);
    for (int ix = 0; ix <= nx; ix++) {
      for (int k = 0; k < nzd; k++) {
        if (k <= nz) line[k] = V[FIDX(o, ix, k, iy)];
        else if (k < nzd - nz) line[k] = 0.0;
        else line[k] = V[FIDX(o, ix, k - nzd, iy)];
      }
      fft_run(&o->pz, line, w2, +1);
      for (int k = 0; k < nzd; k++) VVx[(size_t)k * (nx + 1) + ix] = line[k];
    }
    for (int k = 0; k < nzd; k++) {
      /* x c2r of length N: Hermitian extension, Im of the kx = 0 bin dropped */
      for (int j = 0; j < N; j++) line[j] = 0.0;
      line[0] = creal(VVx[(size_t)k * (nx + 1)]);
      for (int ix = 1; ix <= nx; ix++) {
        line[ix] = VVx[(size_t)k * (nx + 1) + ix];
        line[N - ix] = conj(line[ix]);
      }
      fft_run(&o->px, line, w2, +1);
      double *out = phys + (size_t)c * plane + (size_t)k * N;
      for (int j = 0; j < N; j++) out[j] = creal(line[j]);
    }
  }
  if (cflrow) {
    double cm = 0.0, dyl = o->dyl[iy];
    for (size_t i = 0; i < plane; i++) {
      double t = fabs(phys[i]) / o->dx + fabs(phys[plane + i]) / dyl + fabs(phys[2 * plane + i]) / o->dz;
      if (t > cm) cm = t;
    }
    *cflrow = cm;
  }
  const double factor = o->factor;
  for (int p = 0; p < 6; p++) {
    const double *u = phys, *v = phys + plane, *w = phys + 2 * plane;
    cplx *HX = buf; /* [nzd][nx+1] x modes after r2c */
    cplx *line = work, *w2 = work + (nzd > N ? nzd : N);
    if (mode == 1 && p >= 3) break;
    for (int k = 0; k < nzd; k++) {
      const size_t b = (size_t)k * N;
      for (int j = 0; j < N; j++) {
        double uu = u[b + j], vv = v[b + j], ww = w[b + j], r;
        if (mode == 1) r = (p == 0 ? uu : p == 1 ? vv : ww) * factor;
        else switch (p) {
            case 0: r = uu * uu * factor; break;
            case 1: r = vv * vv * factor; break;
            case 2: r = ww * ww * factor; break;
            case 3: r = uu * vv * factor; break;
            case 4: r = vv * ww * factor; break;
            default: r = uu * ww * factor; break;
          }
        line[j] = r;
      }
      fft_run(&o->px, line, w2, -1);
      for (int ix = 0; ix <= nx; ix++) HX[(size_t)k * (nx + 1) + ix] = line[ix];
    }
    cplx *P = o->prod + (size_t)p * o->NXm * o->NZm * o->NYg;
    for (int ix = 0; ix <= nx; ix++) {
      for (int k = 0; k < nzd; k++) line[k] = HX[(size_t)k * (nx + 1) + ix];
      fft_run(&o->pz, line, w2, -1);
      for (int iz = -nz; iz <= nz; iz++) P[FIDX(o, ix, iz, iy)] = line[iz >= 0 ? iz : nzd + iz];
    }
  }
}

/* transform_to_physical + build_products + products_to_spectral for both
 * groups, all rows -2..ny+1; returns the CFL number when want_cfl. */
double hsto_nonlinear(hsto *o, int mode, int want_cfl) {
  const int nzd = o->nzd, N = 2 * o->nxd;
  double cfl = 0.0;
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nzd * (o->nx + 1));
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(nzd > N ? nzd : N));
    double *phys = (double *)malloc(sizeof(double) * 3 * (size_t)nzd * N);
    double cl = 0.0;
#pragma omp for schedule(dynamic)
    for (int iy = -2; iy <= o->ny + 1; iy++) {
      double cr = 0.0;
      int interior = (iy >= 0 && iy < o->ny);
      row_nonlinear(o, iy, mode, (want_cfl && interior) ? &cr : NULL, buf, work, phys);
      if (cr > cl) cl = cr;
    }
#pragma omp critical
    if (cl > cfl) cfl = cl;
    free(buf); free(work); free(phys);
  }
  o->cfl = cfl;
  return cfl;
}

/* ------------------------------------------------------------------ equations */

/* ------------------------------------------------------------------ physics hooks */

/* stokes_active, damping_active, damping_soft_factor, damping_theta
 * (src/hst_stokes.f90:48-50, src/hst_damping.f90:45-67) */
static int stokes_active(const hsto *o) { return o->d.sl_amplitude != 0.0; }
static int damping_active(const hsto *o) {
  return o->d.damp_kpar > 0.0 && o->d.damp_kperp > 0.0 && (o->d.damp_hard || o->d.damp_rate > 0.0);
}
static double damping_soft_factor(const hsto *o) { return (damping_active(o) && !o->d.damp_hard) ? 1.0 : 0.0; }
static double damping_theta(const hsto *o, double t) {
  if (o->d.damp_follow) return atan2(s2_of(o, t), o->d.S);
  return o->d.damp_angle * PI / 180.0;
}

/* ramp, stokes_profile, stokes_force_at (src/hst_stokes.f90:52-88) */
static double sl_ramp(const hsto *o, double t) {
  const hst_deck *d = &o->d;
  if (t < d->sl_start) return 0.0;
  if (d->sl_ramp) {
    double tau = fmin((t - d->sl_start) / d->sl_period, 1.0);
    return 3.0 * tau * tau - 2.0 * tau * tau * tau;
  }
  return 1.0;
}
static double stokes_profile(const hsto *o, double yy, double t) {
  const hst_deck *d = &o->d;
  const double omega = TWOPI / d->sl_period, e = (yy - 0.5 * d->ly) / d->sl_delta;
  const double s = sqrt(e * e + 0.1 * 0.1);
  return sl_ramp(o, t) * d->sl_amplitude * exp(-s) * cos(omega * (t - d->sl_start) - s);
}
static double stokes_force_at(const hsto *o, double yy, double t) {
  const hst_deck *d = &o->d;
  const double omega = TWOPI / d->sl_period, eta = yy - 0.5 * d->ly, e = eta / d->sl_delta;
  const double s = sqrt(e * e + 0.1 * 0.1);
  const double s1 = (eta / (d->sl_delta * d->sl_delta)) / s;
  const double s2 = ((0.1 * 0.1) / (d->sl_delta * d->sl_delta)) / (s * s * s);
  const double theta = omega * (t - d->sl_start) - s;
  return sl_ramp(o, t) * d->sl_amplitude * exp(-s) *
         ((2.0 * o->ni * s1 * s1 - omega) * sin(theta) - o->ni * s2 * (sin(theta) - cos(theta)));
}

/* stokes_force (src/hst_stokes.f90:90-101): fy with periodic ghost rows */
static void stokes_force(hsto *o, double t) {
  if (!(stokes_active(o) && o->d.sl_bodyforce)) return;
  const int ny = o->ny;
  for (int iy = 0; iy < ny; iy++) o->fy[iy + 2] = stokes_force_at(o, o->y[iy + 2], t);
  o->fy[0] = o->fy[ny - 2 + 2]; o->fy[1] = o->fy[ny - 1 + 2];
  o->fy[ny + 2] = o->fy[2]; o->fy[ny + 3] = o->fy[3];
}

/* stokes_apply (src/hst_stokes.f90:103-118) at t, then fill_ghosts(3) */
void hsto_stokes_apply(hsto *o, double t) {
  if (!stokes_active(o)) return;
  const int whole = (!o->d.sl_bodyforce) || (t - o->d.sl_start <= 2.0 * o->d.sl_period);
  cplx *W = o->f[3];
  for (int iy = 0; iy < o->ny; iy++)
    if (whole || iy == 0 || iy == o->ny - 1) W[FIDX(o, 0, 0, iy)] = stokes_profile(o, o->y[iy + 2], t);
  fill_ghosts_t(o, 3, t);
}

/* damping_update (src/hst_damping.f90:78-115) */
static void damping_update(hsto *o, double t) {
  if (!damping_active(o)) return;
  const hst_deck *d = &o->d;
  const double theta = damping_theta(o, t);
  if (o->theta_set && theta == o->theta_last) return;
  o->theta_set = 1;
  o->theta_last = theta;
  const double c = cos(theta), sn = sin(theta);
  for (int ix = 0; ix <= o->nx; ix++) {
    const double kx = d->alfa0 * ix;
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const double kz = d->beta0 * iz;
      const double kpar = kx * c + kz * sn, kperp = -kx * sn + kz * c;
      const double r = pow(pow(fabs(kpar) / d->damp_kpar, d->damp_norm) + pow(fabs(kperp) / d->damp_kperp, d->damp_norm),
                           1.0 / d->damp_norm);
      double g = 0.0;
      if (r < 1.0 - 1.0e-6) {
        if (d->damp_hard) {
          g = 1.0;
        } else if (r <= 1.0 - d->damp_edge) {
          g = d->damp_rate;
        } else {
          const double x = (1.0 - r) / d->damp_edge;
          g = d->damp_rate * x * x * (3.0 - 2.0 * x);
        }
      }
      if (ix == 0 && iz == 0 && !d->damp_mean) g = 0.0;
      GD(o, iz, ix) = g;
    }
  }
}

/* damping_mask (src/hst_damping.f90:119-133) */
void hsto_damping_mask(hsto *o) {
  if (!(damping_active(o) && o->d.damp_hard)) return;
  cplx *R1 = o->f[HST_RHS1], *R2 = o->f[HST_RHS2], *O1 = o->f[HST_OLD1], *O2 = o->f[HST_OLD2];
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      if (!(GD(o, iz, ix) > 0.0)) continue;
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        R1[c] = 0.0; R2[c] = 0.0; O1[c] = 0.0; O2[c] = 0.0;
      }
    }
}

/* buildrhs_prepare (src/hst_equations.f90:92-141) at `time`, after the
 * substep's damping_update (:64) */
void hsto_buildrhs_prepare(hsto *o, const double ODE[3], double deltat, double time) {
  damping_update(o, time);
  stokes_force(o, time);
  const double s2now = s2_of(o, time), ni = o->ni, S = o->d.S, gsoft = damping_soft_factor(o);
  cplx *V1 = o->f[1], *V2 = o->f[2], *V3 = o->f[3];
  cplx *R1 = o->f[HST_RHS1], *R2 = o->f[HST_RHS2], *O1 = o->f[HST_OLD1], *O2 = o->f[HST_OLD2];
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const double kk = K2(o, iz, ix), g = gsoft * GD(o, iz, ix);
      const cplx ialfa = I * (ix * o->d.alfa0), ibeta = I * (iz * o->d.beta0);
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        cplx unkn = 0.0, impl = 0.0, expl = 0.0, f;
        double helm, biharm, force = 0.0;
        for (int j = -2; j <= 2; j++) {
          helm = DER(o, iy, 2, j) - kk * DER(o, iy, 0, j);
          biharm = DER(o, iy, 3, j) - 2.0 * kk * DER(o, iy, 2, j) + kk * kk * DER(o, iy, 0, j);
          unkn = unkn + helm * V2[c + j];
          impl = impl + ni * biharm * V2[c + j];
        }
        R2[c] = ODE[0] * unkn / deltat + impl - ODE[2] * O2[c] - ODE[1] * g * unkn;
        O2[c] = -g * unkn;
        unkn = 0.0; impl = 0.0;
        for (int j = -2; j <= 2; j++) {
          helm = DER(o, iy, 2, j) - kk * DER(o, iy, 0, j);
          if (ix == 0 && iz == 0) {
            f = creal(V1[c + j]) + I * creal(V3[c + j]);
            unkn = unkn + DER(o, iy, 0, j) * f;
            impl = impl + ni * DER(o, iy, 2, j) * f;
            force = force + DER(o, iy, 0, j) * o->fy[iy + j + 2];
          } else {
            f = ibeta * V1[c + j] - ialfa * V3[c + j];
            unkn = unkn + DER(o, iy, 0, j) * f;
            impl = impl + ni * helm * f;
          }
          expl = expl + (s2now * ialfa - S * ibeta) * DER(o, iy, 0, j) * V2[c + j];
        }
        if (ix == 0 && iz == 0) expl = expl + I * force;
        expl = expl - g * unkn;
        R1[c] = ODE[0] * unkn / deltat + impl - ODE[2] * O1[c] + ODE[1] * expl;
        O1[c] = expl;
      }
    }
}

/* buildrhs(ODE, g) (src/hst_equations.f90:151-213), no Stokes layer */
void hsto_buildrhs(hsto *o, const double ODE[3], int g) {
  cplx *R1 = o->f[HST_RHS1], *R2 = o->f[HST_RHS2], *O1 = o->f[HST_OLD1], *O2 = o->f[HST_OLD2];
  /* with the layer's body force the mean w equation drops the Reynolds-stress divergence */
  const int no_mean_vw = stokes_active(o) && o->d.sl_bodyforce;
  const size_t nf = (size_t)o->NXm * o->NZm * o->NYg;
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const double kk = K2(o, iz, ix);
      const cplx ialfa = I * (ix * o->d.alfa0), ibeta = I * (iz * o->d.beta0);
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        cplx r1 = R1[c], r2 = R2[c], o1 = O1[c], o2 = O2[c];
        for (int p = 1; p <= 3; p++) {
          int m = 3 * (g - 1) + p;
          const cplx *P = o->prod + (size_t)(m - 1) * nf;
          cplx d0 = 0.0, d1 = 0.0, d2 = 0.0, rhsu, rhsw, expl, e;
          for (int j = -2; j <= 2; j++) {
            d0 = d0 + DER(o, iy, 0, j) * P[c + j];
            d1 = d1 + DER(o, iy, 1, j) * P[c + j];
            d2 = d2 + DER(o, iy, 2, j) * P[c + j];
          }
          switch (m) {
            case 1: rhsu = -(ialfa * d0); rhsw = 0.0; expl = ialfa * ialfa * d1; break;
            case 2: rhsu = 0.0; rhsw = 0.0; expl = kk * d1; break;
            case 3: rhsu = 0.0; rhsw = -(ibeta * d0); expl = ibeta * ibeta * d1; break;
            case 4: rhsu = -d1; rhsw = 0.0; expl = ialfa * d2 + ialfa * kk * d0; break;
            case 5:
              rhsu = 0.0; rhsw = -d1;
              if (no_mean_vw && ix == 0 && iz == 0) rhsw = 0.0;
              expl = ibeta * d2 + ibeta * kk * d0;
              break;
            default: rhsu = -(ibeta * d0); rhsw = -(ialfa * d0); expl = 2.0 * ialfa * ibeta * d1; break;
          }
          r2 = r2 + ODE[1] * expl;
          o2 = o2 + expl;
          if (ix == 0 && iz == 0) e = creal(rhsu) + I * creal(rhsw);
          else e = ibeta * rhsu - ialfa * rhsw;
          r1 = r1 + ODE[1] * e;
          o1 = o1 + e;
        }
        R1[c] = r1; R2[c] = r2; O1[c] = o1; O2[c] = o2;
      }
    }
}

/* shift_unweighted (src/hst_equations.f90:263-293) of field q */
static void shift_unweighted(hsto *o, int q, double dgx, double dgz, double sx0, double sz0, double sx1, double sz1) {
  hsto_line_solve(o, HST_KIND_D0, 0.0, q, HST_V1, sx0, sz0);
  cplx *V1 = o->f[HST_V1], *Q = o->f[q];
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++)
      for (int iy = 0; iy < o->ny; iy++) {
        double a = -(o->d.alfa0 * ix * dgx + o->d.beta0 * iz * dgz) * o->y[iy + 2];
        V1[FIDX(o, ix, iz, iy)] = V1[FIDX(o, ix, iz, iy)] * (cos(a) + I * sin(a));
      }
  hsto_fill_ghosts(o, HST_V1, sx1, sz1);
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++)
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        cplx acc = 0.0;
        for (int j = -2; j <= 2; j++) acc = acc + DER(o, iy, 0, j) * V1[c + j];
        Q[c] = acc;
      }
}

/* shear_shift (src/hst_equations.f90:226-258) */
void hsto_shear_shift(hsto *o, double dt_sub, double time) {
  const double dgx = o->d.S * dt_sub, dgz = s2_integral(o, time, time + dt_sub);
  if (o->d.exact_shift) {
    double sx0, sz0, sx1, sz1;
    hsto_shear_shifts(o, time, &sx0, &sz0);
    hsto_shear_shifts(o, time + dt_sub, &sx1, &sz1);
    shift_unweighted(o, HST_RHS1, dgx, dgz, sx0, sz0, sx1, sz1);
    shift_unweighted(o, HST_RHS2, dgx, dgz, sx0, sz0, sx1, sz1);
    shift_unweighted(o, HST_OLD1, dgx, dgz, sx0, sz0, sx1, sz1);
    shift_unweighted(o, HST_OLD2, dgx, dgz, sx0, sz0, sx1, sz1);
    return;
  }
  cplx *R1 = o->f[HST_RHS1], *R2 = o->f[HST_RHS2], *O1 = o->f[HST_OLD1], *O2 = o->f[HST_OLD2];
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++)
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        double a = -(o->d.alfa0 * ix * dgx + o->d.beta0 * iz * dgz) * o->y[iy + 2];
        cplx f = cos(a) + I * sin(a);
        R1[c] = R1[c] * f; R2[c] = R2[c] * f; O1[c] = O1[c] * f; O2[c] = O2[c] * f;
      }
}

/* linsolve(lambda) (src/hst_equations.f90:298-326) at time */
void hsto_linsolve(hsto *o, double lambda, double time) {
  double sx, sz;
  hsto_shear_shifts(o, time, &sx, &sz);
  hsto_line_solve(o, HST_KIND_D2V, lambda, HST_RHS2, HST_V2, sx, sz);
  hsto_line_solve(o, HST_KIND_ETA, lambda, HST_RHS1, HST_V1, sx, sz);
  hsto_fill_ghosts(o, HST_V2, sx, sz);
  hsto_line_solve(o, HST_KIND_DY, 0.0, HST_V2, HST_V3, sx, sz);
  cplx *V1 = o->f[1], *V3 = o->f[3];
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const double kk = K2(o, iz, ix);
      const cplx ialfa = I * (ix * o->d.alfa0), ibeta = I * (iz * o->d.beta0);
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        if (ix == 0 && iz == 0) {
          cplx t = V1[c];
          V1[c] = creal(t);
          V3[c] = cimag(t);
        } else {
          cplx t = (ialfa * V3[c] - ibeta * V1[c]) / kk;
          V3[c] = (ibeta * V3[c] + ialfa * V1[c]) / kk;
          V1[c] = t;
        }
      }
    }
  hsto_fill_ghosts(o, HST_V1, sx, sz);
  hsto_fill_ghosts(o, HST_V3, sx, sz);
}

/* timestep (src/hst_equations.f90:56-87) */
void hsto_timestep(hsto *o, double deltat, double *time, double *cfl) {
  for (int i = 0; i < 3; i++) {
    const double *ODE = RK_rai[i];
    int nonlin = !o->d.linear;
    if (nonlin) {
      double c = hsto_nonlinear(o, 0, i == 2);
      if (i == 2 && cfl) *cfl = c;
    } else if (i == 2) {
      double c = hsto_nonlinear(o, 1, 1); /* transform for the CFL number only */
      if (cfl) *cfl = c;
    }
    hsto_buildrhs_prepare(o, ODE, deltat, *time);
    if (nonlin) {
      hsto_buildrhs(o, ODE, 1);
      hsto_buildrhs(o, ODE, 2);
    }
    hsto_damping_mask(o);
    double dt_sub = 2.0 / ODE[0] * deltat;
    hsto_shear_shift(o, dt_sub, *time);
    *time = *time + dt_sub;
    hsto_linsolve(o, ODE[0] / deltat, *time);
    hsto_stokes_apply(o, *time);
  }
}

/* fill the ghost rows of u, v, w at time t (the start-up fill, src/hst.f90:240-242) */
void hsto_fill_all_ghosts(hsto *o, double t) {
  for (int c = 1; c <= 3; c++) fill_ghosts_t(o, c, t);
}

/* ------------------------------------------------------------------ statistics */

/* The first reduction of outstats (src/hst_stats.f90:153-198) without Stokes
 * layer and damping: raw sums eps(spectral), uv, uu, vv, ww, vw, uw, mfx, mfz. */
/* The first reduction of outstats (src/hst_stats.f90:145-198):
 * out = eps uv uu vv ww vw uw mfx mfz q_in q_out e_in e_out ed pd ncol_damped */
static void outstats_raw(hsto *o, int inc00, double out[16]) {
  double eps = 0, uv = 0, uu = 0, vv = 0, ww = 0, vw = 0, uw = 0, mfx = 0, mfz = 0;
  double q_in = 0, q_out = 0, e_in = 0, e_out = 0, ed = 0, pd = 0, nd = 0;
  const double gsoft = damping_soft_factor(o);
  cplx *V1 = o->f[1], *V2 = o->f[2], *V3 = o->f[3];
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const double gd = GD(o, iz, ix);
      if (gd > 0.0) nd += 1.0;
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        cplx cu = V1[c], cv = V2[c], cw = V3[c];
        if (ix == 0 && iz == 0) {
          mfx += creal(cu) * o->dyl[iy];
          mfz += creal(cw) * o->dyl[iy];
          if (gd > 0.0) {
            double qc = o->dyl[iy] * creal(cu * conj(cu) + cw * conj(cw));
            ed += qc;
            pd += 2.0 * gsoft * gd * qc;
          }
          if (inc00) {
            uu += o->dyl[iy] * creal(cu * conj(cu));
            ww += o->dyl[iy] * creal(cw * conj(cw));
            uw += o->dyl[iy] * creal(cu * conj(cw));
          }
          continue;
        }
        double w = 2.0 * o->dyl[iy];
        if (ix == 0) w = o->dyl[iy];
        const double q = w * creal(cu * conj(cu) + cv * conj(cv) + cw * conj(cw));
        if (gd > 0.0) {
          ed += q;
          pd += 2.0 * gsoft * gd * q;
        }
        uu += w * creal(cu * conj(cu));
        vv += w * creal(cv * conj(cv));
        ww += w * creal(cw * conj(cw));
        uv += w * creal(cu * conj(cv));
        vw += w * creal(cw * conj(cv));
        uw += w * creal(cu * conj(cw));
        const double e = w * K2(o, iz, ix) * creal(cu * conj(cu) + cv * conj(cv) + cw * conj(cw));
        eps += e;
        if (o->inlayer[iy] == 1) {
          q_in += q;
          e_in += e;
        } else {
          q_out += q;
          e_out += e;
        }
      }
    }
  out[0] = eps; out[1] = uv; out[2] = uu; out[3] = vv; out[4] = ww; out[5] = vw; out[6] = uw; out[7] = mfx; out[8] = mfz;
  out[9] = q_in; out[10] = q_out; out[11] = e_in; out[12] = e_out; out[13] = ed; out[14] = pd; out[15] = nd;
}

/* outstats (src/hst_stats.f90:136-285) at `time` (the wrap of the DY solves):
 * out = energy, diss, uv, vw, uu, vv, ww, uw, mfx, mfz as written to
 * Runtimedata / variances_runtime.dat; the region and damping lines go to
 * o->extra (hsto_outstats_extra). */
void hsto_outstats_at(hsto *o, double time, double out[10]) {
  if (!o->theta_set) damping_update(o, time); /* damping_setup's table at the start */
  const int inc00 = !stokes_active(o);
  double raw[16];
  outstats_raw(o, inc00, raw);
  double eps = raw[0], uv = raw[1], uu = raw[2], vv = raw[3], ww = raw[4], vw = raw[5], uw = raw[6];
  double mfx = raw[7], mfz = raw[8], e_in = raw[11], e_out = raw[12];
  double sx, sz;
  hsto_shear_shifts(o, time, &sx, &sz);
  for (int c = 1; c <= 3; c++) {
    hsto_line_solve(o, HST_KIND_DY, 0.0, c, HST_SCRATCH, sx, sz);
    const cplx *R = o->f[HST_SCRATCH];
    double grad = 0.0, g_in = 0.0, g_out = 0.0;
    for (int ix = 0; ix <= o->nx; ix++)
      for (int iz = -o->nz; iz <= o->nz; iz++)
        for (int iy = 0; iy < o->ny; iy++) {
          if (ix == 0 && iz == 0 && !inc00) continue;
          double w = 2.0 * o->dyl[iy];
          if (ix == 0) w = o->dyl[iy];
          cplx dq = R[FIDX(o, ix, iz, iy)];
          const double q = w * creal(dq * conj(dq));
          grad += q;
          if (o->inlayer[iy] == 1) g_in += q;
          else g_out += q;
        }
    eps += grad;
    e_in += g_in;
    e_out += g_out;
  }
  if (inc00) {
    double ly_w = 0.0;
    for (int iy = 0; iy < o->ny; iy++) ly_w += o->dyl[iy];
    uu = uu - mfx * mfx / ly_w;
    ww = ww - mfz * mfz / ly_w;
    uw = uw - mfx * mfz / ly_w;
  }
  double q2 = uu + vv + ww, inv_ly = 1.0 / o->d.ly;
  out[0] = 0.5 * q2 * inv_ly;
  out[1] = o->ni * eps * inv_ly;
  out[2] = uv * inv_ly;
  out[3] = vw * inv_ly;
  out[4] = uu * inv_ly;
  out[5] = vv * inv_ly;
  out[6] = ww * inv_ly;
  out[7] = uw * inv_ly;
  out[8] = mfx;
  out[9] = mfz;
  for (int k = 0; k < 8; k++) o->extra[k] = 0.0;
  if (stokes_active(o)) {
    double l_in = 0.0, l_out = 0.0;
    for (int iy = 0; iy < o->ny; iy++) {
      if (o->inlayer[iy] == 1) l_in += o->dyl[iy];
      else l_out += o->dyl[iy];
    }
    o->extra[0] = 0.5 * raw[10] / l_out;
    o->extra[1] = 0.5 * raw[9] / l_in;
    o->extra[2] = o->ni * e_out / l_out;
    o->extra[3] = o->ni * e_in / l_in;
  }
  if (damping_active(o)) {
    o->extra[4] = damping_theta(o, time) * 180.0 / 3.141592653589793;
    o->extra[5] = raw[15];
    o->extra[6] = 0.5 * raw[13];
    o->extra[7] = 0.5 * raw[14];
  }
}

void hsto_outstats_extra(const hsto *o, double out[8]) { memcpy(out, o->extra, sizeof(double) * 8); }

/* project_restart (src/hst_io.f90:213-250) with the ghost fills around it
 * (:184-190) at `time`.  The reference updates V in place while the
 * stencils of neighbouring rows read it; here every row's divergence is that
 * of the field before the repair (the originals are copied first). */
void hsto_project_divergence(hsto *o, double time) {
  fill_ghosts_t(o, 1, time);
  fill_ghosts_t(o, 3, time);
  fill_ghosts_t(o, 2, time);
  const size_t nf = (size_t)o->NXm * o->NZm * o->NYg;
  cplx *U = (cplx *)malloc(sizeof(cplx) * nf * 3);
  memcpy(U, o->f[1], sizeof(cplx) * nf);
  memcpy(U + nf, o->f[2], sizeof(cplx) * nf);
  memcpy(U + 2 * nf, o->f[3], sizeof(cplx) * nf);
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const double al = o->d.alfa0 * ix, be = o->d.beta0 * iz;
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        cplx du = 0.0, dv = 0.0, dw = 0.0;
        for (int j = -2; j <= 2; j++) {
          du = du + DER(o, iy, 0, j) * U[c + j];
          dv = dv + DER(o, iy, 1, j) * U[nf + c + j];
          dw = dw + DER(o, iy, 0, j) * U[2 * nf + c + j];
        }
        if (ix != 0) {
          cplx div = dv + I * be * dw + I * al * du;
          o->f[1][c] = (creal(U[c]) - cimag(div) / al) + I * (cimag(U[c]) + creal(div) / al);
        } else if (iz != 0) {
          cplx div = dv + I * be * dw;
          o->f[3][c] = (creal(U[2 * nf + c]) - cimag(div) / be) + I * (cimag(U[2 * nf + c]) + creal(div) / be);
        } else {
          o->f[2][c] = 0.0;
        }
      }
    }
  free(U);
  fill_ghosts_t(o, 1, time);
  fill_ghosts_t(o, 3, time);
  fill_ghosts_t(o, 2, time);
}

/* gdamp(-nz:nz, 0:nx) as last built (zero before the first damping_update) */
double *hsto_gdamp(hsto *o) { return o->gdamp; }

/* compute_pressure (src/hst_pressure.f90:35-112) into field dst at `time`
 * (the velocity's ghost rows filled): the products once more, the pressure
 * right-hand side into HST_RHS1, a POISSON line solve, and the singular
 * (0,0) mode set to -<vv> about its y mean. */
void hsto_pressure(hsto *o, double time, int dst) {
  hsto_nonlinear(o, 0, 0);
  const double s2now = s2_of(o, time), S = o->d.S;
  const cplx *V2 = o->f[2];
  cplx *R = o->f[HST_RHS1];
  const size_t nf = (size_t)o->NXm * o->NZm * o->NYg;
#pragma omp parallel for collapse(2)
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      const cplx ialfa = I * (ix * o->d.alfa0), ibeta = I * (iz * o->d.beta0);
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        cplx r = 0.0;
        for (int m = 1; m <= 6; m++) {
          const cplx *P = o->prod + (size_t)(m - 1) * nf;
          cplx d0 = 0.0, d1 = 0.0, d2 = 0.0, term;
          for (int j = -2; j <= 2; j++) {
            d0 = d0 + DER(o, iy, 0, j) * P[c + j];
            d1 = d1 + DER(o, iy, 1, j) * P[c + j];
            d2 = d2 + DER(o, iy, 2, j) * P[c + j];
          }
          switch (m) {
            case 1: term = -(ialfa * ialfa * d0); break;
            case 2: term = -d2; break;
            case 3: term = -(ibeta * ibeta * d0); break;
            case 4: term = -2.0 * ialfa * d1; break;
            case 5: term = -2.0 * ibeta * d1; break;
            default: term = -2.0 * ialfa * ibeta * d0; break;
          }
          if (m == 1) {
            cplx v0 = 0.0;
            for (int j = -2; j <= 2; j++) v0 = v0 + DER(o, iy, 0, j) * V2[c + j];
            r = term - 2.0 * (S * ialfa + s2now * ibeta) * v0;
          } else {
            r = r + term;
          }
          if (ix == 0 && iz == 0 && m == 2) r = -creal(P[c]);
        }
        R[c] = r;
      }
    }
  double sx, sz;
  hsto_shear_shifts(o, time, &sx, &sz);
  hsto_line_solve(o, HST_KIND_POISSON, 0.0, HST_RHS1, dst, sx, sz);
  cplx *p = o->f[dst];
  double pmean = 0.0;
  for (int iy = 0; iy < o->ny; iy++) {
    size_t c = FIDX(o, 0, 0, iy);
    p[c] = R[c];
    pmean += creal(p[c]) * o->dyl[iy];
  }
  pmean = pmean / o->d.ly;
  for (int iy = 0; iy < o->ny; iy++) {
    size_t c = FIDX(o, 0, 0, iy);
    p[c] = creal(p[c]) - pmean;
  }
}

/* The column energies of outstats (src/hst_stats.f90:224-238):
 * E[ix][iz + nz] = sum_y w |u|^2 + |v|^2 + |w|^2, w = dyl (ix = 0) or 2 dyl. */
void hsto_column_energy(const hsto *o, double *E) {
  const cplx *V1 = o->f[1], *V2 = o->f[2], *V3 = o->f[3];
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      double qc = 0.0;
      for (int iy = 0; iy < o->ny; iy++) {
        size_t c = FIDX(o, ix, iz, iy);
        double w = ix == 0 ? o->dyl[iy] : 2.0 * o->dyl[iy];
        qc += w * creal(V1[c] * conj(V1[c]) + V2[c] * conj(V2[c]) + V3[c] * conj(V3[c]));
      }
      E[(size_t)ix * o->NZm + (iz + o->nz)] = qc;
    }
}

void hsto_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int hsto_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ================================================================== learned closure
 *
 * The learned subgrid closure has no counterpart in the reference: these
 * functions restate this library's own definitions (the filter, the samples,
 * the network and its gradient) so that the GPU kernels have a reference to
 * be compared with.
 *
 * Pinning:
 *   filter weights     closed forms: unit sum, the 4:1 trapezoid [1/8 1/4 1/4 1/4 1/8],
 *                      exactness on linear data, the top-hat transfer function
 *   to/from physical   the round trip and single-mode closed forms (tests/)
 *   network gradient   central finite differences of the FP64 loss
 *   bf16 rounding      torch's CPU bfloat16 conversion (round to nearest even)
 */

/* Top-hat filter in y of width ly/nyc centred on coarse node k, applied to the
 * piecewise-linear interpolant of the fine rows (uniform grids, same ly).
 * Fine row j (may be -2, -1 or >= nyf: the shear-periodic images) gets
 * weight (1/(b-a)) * integral over [a, b] of the hat function centred on j,
 * [a, b] = k nyf/nyc -+ nyf/(2 nyc) in fine-row units.  Up to 8 taps from j0[k]. */
static double hat_cdf(double x) {
  if (x <= -1.0) return 0.0;
  if (x <= 0.0) return 0.5 * (x + 1.0) * (x + 1.0);
  if (x <= 1.0) return 1.0 - 0.5 * (1.0 - x) * (1.0 - x);
  return 1.0;
}

int hsto_filter_weights(int nyf, int nyc, int *j0, double *w) {
  if (nyc < 2 || nyf < nyc) return 1;
  const double r = 0.5 * (double)nyf / (double)nyc;
  for (int k = 0; k < nyc; k++) {
    const double c = (double)k * (double)nyf / (double)nyc, a = c - r, b = c + r;
    const int lo = (int)floor(a), hi = (int)ceil(b);
    if (hi - lo + 1 > 8) return 2;
    j0[k] = lo;
    for (int t = 0; t < 8; t++) {
      const int j = lo + t;
      w[k * 8 + t] = (j <= hi) ? (hat_cdf(b - j) - hat_cdf(a - j)) / (b - a) : 0.0;
    }
  }
  return 0;
}

/* Filter of a fine spectral field (reference layout, ghost rows filled) to the
 * coarse grid: the Fourier columns |ix| <= nx_c, |iz| <= nz_c are kept (a
 * sharp spectral box in x and z; alfa0, beta0 equal) and each coarse row is
 * the weighted sum of fine rows above.  Coarse ghost rows are left to the
 * caller. */
void hsto_filter_field(const hsto *fo, const hsto *co, const double *fine, double *coarse) {
  const cplx *F = (const cplx *)fine;
  cplx *C = (cplx *)coarse;
  int *j0 = (int *)malloc(sizeof(int) * co->ny);
  double *w = (double *)malloc(sizeof(double) * 8 * co->ny);
  hsto_filter_weights(fo->ny, co->ny, j0, w);
  for (int ix = 0; ix <= co->nx; ix++)
    for (int iz = -co->nz; iz <= co->nz; iz++)
      for (int k = 0; k < co->ny; k++) {
        cplx s = 0.0;
        for (int t = 0; t < 8; t++) {
          if (w[k * 8 + t] == 0.0) continue;
          s += w[k * 8 + t] * F[FIDX(fo, ix, iz, j0[k] + t)];
        }
        C[FIDX(co, ix, iz, k)] = s;
      }
  free(j0);
  free(w);
}

/* A spectral field to physical values at the interior rows:
 * phys[iy][zp][x], x < 2 nxd, zp < nzd.  op 0: the field; 1: d/dx (times
 * i alfa0 ix); 2: d/dz (times i beta0 iz).  The transforms of row_nonlinear. */
void hsto_to_physical(const hsto *o, const double *field, int op, double *phys) {
  const cplx *V = (const cplx *)field;
  const int nzd = o->nzd, nx = o->nx, nz = o->nz, N = 2 * o->nxd;
#pragma omp parallel
  {
    cplx *VVx = (cplx *)malloc(sizeof(cplx) * (size_t)nzd * (nx + 1));
    const int L = nzd > N ? nzd : N;
    cplx *line = (cplx *)malloc(sizeof(cplx) * 2 * L), *w2 = line + L;
#pragma omp for
    for (int iy = 0; iy < o->ny; iy++) {
      for (int ix = 0; ix <= nx; ix++) {
        for (int k = 0; k < nzd; k++) {
          int iz = k <= nz ? k : (k >= nzd - nz ? k - nzd : 1 << 30);
          cplx v = 0.0;
          if (iz != 1 << 30) {
            v = V[FIDX(o, ix, iz, iy)];
            if (op == 1) v *= I * o->d.alfa0 * ix;
            else if (op == 2) v *= I * o->d.beta0 * iz;
          }
          line[k] = v;
        }
        fft_run(&o->pz, line, w2, +1);
        for (int k = 0; k < nzd; k++) VVx[(size_t)k * (nx + 1) + ix] = line[k];
      }
      for (int k = 0; k < nzd; k++) {
        for (int j = 0; j < N; j++) line[j] = 0.0;
        line[0] = creal(VVx[(size_t)k * (nx + 1)]);
        for (int ix = 1; ix <= nx; ix++) {
          line[ix] = VVx[(size_t)k * (nx + 1) + ix];
          line[N - ix] = conj(line[ix]);
        }
        fft_run(&o->px, line, w2, +1);
        double *out = phys + ((size_t)iy * nzd + k) * N;
        for (int j = 0; j < N; j++) out[j] = creal(line[j]);
      }
    }
    free(VVx);
    free(line);
  }
}

/* Physical values at the interior rows back to a spectral field (reference
 * layout): times factor, real forward x transform, z transform, truncation
 * to the kept modes.  Ghost rows are left to the caller. */
void hsto_from_physical(const hsto *o, const double *phys, double *field) {
  cplx *V = (cplx *)field;
  const int nzd = o->nzd, nx = o->nx, nz = o->nz, N = 2 * o->nxd;
#pragma omp parallel
  {
    cplx *HX = (cplx *)malloc(sizeof(cplx) * (size_t)nzd * (nx + 1));
    const int L = nzd > N ? nzd : N;
    cplx *line = (cplx *)malloc(sizeof(cplx) * 2 * L), *w2 = line + L;
#pragma omp for
    for (int iy = 0; iy < o->ny; iy++) {
      for (int k = 0; k < nzd; k++) {
        const double *in = phys + ((size_t)iy * nzd + k) * N;
        for (int j = 0; j < N; j++) line[j] = in[j] * o->factor;
        fft_run(&o->px, line, w2, -1);
        for (int ix = 0; ix <= nx; ix++) HX[(size_t)k * (nx + 1) + ix] = line[ix];
      }
      for (int ix = 0; ix <= nx; ix++) {
        for (int k = 0; k < nzd; k++) line[k] = HX[(size_t)k * (nx + 1) + ix];
        fft_run(&o->pz, line, w2, -1);
        for (int iz = -nz; iz <= nz; iz++) V[FIDX(o, ix, iz, iy)] = line[iz >= 0 ? iz : nzd + iz];
      }
    }
    free(HX);
    free(line);
  }
}

/* ghost rows of an arbitrary reference-layout field (wrap_rows) */
void hsto_fill_ghosts_of(const hsto *o, double *field, double sx, double sz) {
  cplx *f = (cplx *)field;
  const int ny = o->ny;
  for (int ix = 0; ix <= o->nx; ix++)
    for (int iz = -o->nz; iz <= o->nz; iz++) {
      double a = -(o->d.alfa0 * ix * sx + o->d.beta0 * iz * sz);
      cplx ph = cos(a) + I * sin(a);
      cplx *c = f + FIDX(o, ix, iz, -2);
      c[ny + 2] = c[0 + 2] * ph;
      c[ny + 1 + 2] = c[1 + 2] * ph;
      c[-1 + 2] = c[ny - 1 + 2] * conj(ph);
      c[-2 + 2] = c[ny - 2 + 2] * conj(ph);
    }
}

/* ------------------------------------------------------------------ the network
 *
 * 9 inputs (the filtered velocity gradients du_i/dx_j, i-major), two hidden
 * layers of 64 tanh units, 6 outputs (tau_uu vv ww uv vw uw).  Parameters,
 * one flat array: W1[64][9] b1[64] W2[64][64] b2[64] W3[6][64] b3[6].
 * Inputs and targets are standardised: xh = (x - mx) * rx, yh = (y - my) * ry
 * with norm = [mx(9) sx(9) my(6) sy(6)], rx = 1/sx, ry = 1/sy, in FP32.
 * Loss: mean over samples and the 6 outputs of (out - yh)^2.
 *
 * bf = 1 rounds to bfloat16 (nearest even) where the GPU kernels feed the
 * matrix cores: the standardised inputs, the weights, the hidden activations
 * and the back-propagated errors; sums, biases and tanh stay in FP32/FP64. */
#define MLP_NIN 9
#define MLP_NH 64
#define MLP_NOUT 6
#define MLP_NP (64 * 9 + 64 + 64 * 64 + 64 + 6 * 64 + 6)
#define OW1 0
#define OB1 (64 * 9)
#define OW2 (OB1 + 64)
#define OB2 (OW2 + 64 * 64)
#define OW3 (OB2 + 64)
#define OB3 (OW3 + 6 * 64)

int hsto_mlp_nparam(void) { return MLP_NP; }

float hsto_bf16_round(float x) {
  union { float f; uint32_t u; } v;
  v.f = x;
  if ((v.u & 0x7f800000u) == 0x7f800000u) return x; /* inf, nan: unchanged */
  v.u += 0x7fffu + ((v.u >> 16) & 1u);
  v.u &= 0xffff0000u;
  return v.f;
}

static inline double bfr(double x, int bf) { return bf ? (double)hsto_bf16_round((float)x) : x; }
static inline double f32r(double x, int bf) { return bf ? (double)(float)x : x; }

/* Forward and (when grad) backward pass over n samples; returns the loss
 * (sum over samples and outputs divided by n*6).  out: n*6 standardised
 * outputs when not NULL. */
static double mlp_pass(const double *par, const double *norm, int64_t n, const float *X, const float *Y, int bf,
                       double *grad, double *out) {
  double W1[64 * 9], W2[64 * 64], W3[6 * 64];
  for (int i = 0; i < 64 * 9; i++) W1[i] = bfr(par[OW1 + i], bf);
  for (int i = 0; i < 64 * 64; i++) W2[i] = bfr(par[OW2 + i], bf);
  for (int i = 0; i < 6 * 64; i++) W3[i] = bfr(par[OW3 + i], bf);
  const double *b1 = par + OB1, *b2 = par + OB2, *b3 = par + OB3;
  float mx[9], rx[9], my[6], ry[6];
  for (int k = 0; k < 9; k++) { mx[k] = (float)norm[k]; rx[k] = 1.0f / (float)norm[9 + k]; }
  for (int k = 0; k < 6; k++) { my[k] = (float)norm[18 + k]; ry[k] = 1.0f / (float)norm[24 + k]; }
  if (grad) memset(grad, 0, sizeof(double) * MLP_NP);
  double loss = 0.0;
  const double scale = 2.0 / ((double)n * 6.0);
#pragma omp parallel reduction(+ : loss)
  {
    double *g = grad ? (double *)calloc(MLP_NP, sizeof(double)) : NULL;
#pragma omp for schedule(static)
    for (int64_t s = 0; s < n; s++) {
      double x[9], h1[64], h2[64], h1b[64], h2b[64], o[6];
      for (int k = 0; k < 9; k++) {
        float xh = (X[s * 9 + k] - mx[k]) * rx[k];
        x[k] = bf ? (double)hsto_bf16_round(xh) : (double)xh;
      }
      for (int j = 0; j < 64; j++) {
        double z = 0.0;
        for (int k = 0; k < 9; k++) z += W1[j * 9 + k] * x[k];
        z = f32r(f32r(z, bf) + b1[j], bf);
        h1[j] = tanh(z);
        h1b[j] = bfr(h1[j], bf);
      }
      for (int j = 0; j < 64; j++) {
        double z = 0.0;
        for (int k = 0; k < 64; k++) z += W2[j * 64 + k] * h1b[k];
        z = f32r(f32r(z, bf) + b2[j], bf);
        h2[j] = tanh(z);
        h2b[j] = bfr(h2[j], bf);
      }
      double dy[6];
      for (int j = 0; j < 6; j++) {
        double z = 0.0;
        for (int k = 0; k < 64; k++) z += W3[j * 64 + k] * h2b[k];
        o[j] = f32r(f32r(z, bf) + b3[j], bf);
        if (out) out[s * 6 + j] = o[j];
        if (Y) {
          float yh = (Y[s * 6 + j] - my[j]) * ry[j];
          double e = o[j] - (double)yh;
          loss += e * e;
          dy[j] = f32r(scale * e, bf);
        }
      }
      if (!g || !Y) continue;
      double dyb[6], d2[64], d2b[64], d1[64], d1b[64];
      for (int j = 0; j < 6; j++) dyb[j] = bfr(dy[j], bf);
      for (int k = 0; k < 64; k++) {
        double a = 0.0;
        for (int j = 0; j < 6; j++) a += W3[j * 64 + k] * dyb[j];
        d2[k] = f32r(f32r(a, bf) * (1.0 - h2[k] * h2[k]), bf);
        d2b[k] = bfr(d2[k], bf);
      }
      for (int k = 0; k < 64; k++) {
        double a = 0.0;
        for (int j = 0; j < 64; j++) a += W2[j * 64 + k] * d2b[j];
        d1[k] = f32r(f32r(a, bf) * (1.0 - h1[k] * h1[k]), bf);
        d1b[k] = bfr(d1[k], bf);
      }
      for (int j = 0; j < 6; j++) {
        g[OB3 + j] += dy[j];
        for (int k = 0; k < 64; k++) g[OW3 + j * 64 + k] += dyb[j] * h2b[k];
      }
      for (int j = 0; j < 64; j++) {
        g[OB2 + j] += d2[j];
        for (int k = 0; k < 64; k++) g[OW2 + j * 64 + k] += d2b[j] * h1b[k];
        g[OB1 + j] += d1[j];
        for (int k = 0; k < 9; k++) g[OW1 + j * 9 + k] += d1b[j] * x[k];
      }
    }
    if (g) {
#pragma omp critical
      for (int i = 0; i < MLP_NP; i++) grad[i] += g[i];
      free(g);
    }
  }
  return loss / ((double)n * 6.0);
}

double hsto_mlp_grad(const double *par, const double *norm, int64_t n, const float *X, const float *Y, int bf,
                     double *grad) {
  return mlp_pass(par, norm, n, X, Y, bf, grad, NULL);
}

void hsto_mlp_forward(const double *par, const double *norm, int64_t n, const float *X, int bf, double *out) {
  mlp_pass(par, norm, n, X, NULL, bf, NULL, out);
}
