// Host-side set-up of libhstgpu: derived deck values, the y grid, the compact
// weights, the seeded initial field and the column-independent factorisation
// of the D0 matrix.  No HIP calls here: these entry points run without a GPU.
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "hst_internal.h"

namespace hst {

static const double PI = 3.141592653589793;
static const double TWOPI = 6.283185307179586;

static bool fft_fit(int n) {
  int j = n;
  while (j % 2 == 0) j /= 2;
  return j == 1 || j == 3;
}

// x**n for a runtime integer exponent, by binary powering (what gfortran
// evaluates for a real raised to an integer variable).
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

static double ymap(double ystretch, double xi) {
  if (ystretch <= 1.0e-10) return xi;
  if (xi <= 0.5) return 0.5 * std::tanh(2.0 * ystretch * xi) / std::tanh(ystretch);
  return 0.5 * (2.0 + std::tanh(2.0 * ystretch * (xi - 1.0)) / std::tanh(ystretch));
}

// 5x5 solve with partial pivoting; M[row][col].
static void solve5(double M[5][5], double b[5], double x[5]) {
  for (int i = 0; i < 5; i++) {
    int p = i;
    double best = std::fabs(M[i][i]);
    for (int r = i + 1; r < 5; r++)
      if (std::fabs(M[r][i]) > best) { best = std::fabs(M[r][i]); p = r; }
    if (p != i) {
      for (int c = 0; c < 5; c++) std::swap(M[i][c], M[p][c]);
      std::swap(b[i], b[p]);
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

std::string check_deck(const hst_deck &d) {
  if (d.nx < 1 || d.nz < 1) return "nx and nz must be positive";
  if (d.ny < 8) return "ny must be at least 8";
  // the checks of read_input (src/hst_input.f90:67-82)
  if (d.sl_amplitude != 0.0 && (d.s2_amplitude != 0.0 || d.s2_mean != 0.0))
    return "the Stokes layer and the spanwise shear S2 cannot both be active";
  if (d.sl_amplitude != 0.0 && !(d.sl_period > 0.0)) return "sl_period must be positive";
  if (d.sl_amplitude != 0.0 && !(d.sl_delta > 0.0)) return "sl_delta must be positive";
  if ((d.damp_kpar > 0.0) != (d.damp_kperp > 0.0))
    return "damp_kpar and damp_kperp must both be positive (damping on) or both zero (off)";
  if (d.damp_kpar > 0.0 && (d.damp_norm < 1.0 || d.damp_edge < 0.0 || d.damp_edge >= 1.0 || d.damp_rate < 0.0))
    return "damping needs damp_norm >= 1, 0 <= damp_edge < 1, damp_rate >= 0";
  if (!(d.re > 0.0)) return "re must be positive";
  if (!(d.alfa0 > 0.0) || !(d.beta0 > 0.0) || !(d.ly > 0.0)) return "alfa0, beta0 and ly must be positive";
  if (d.initial_condition == 1) {
    if (d.tg_mode < 1) return "tg_mode must be positive";
    if (d.tg_mode > d.nx || d.tg_mode > d.nz) return "tg_mode must not exceed nx or nz";
  }
  return "";
}

// Sizes and grid tables: read_input's derived block, allocate_fields'
// grid, setup_derivatives.  der is der(0:ny-1, 0:3, -2:2), iy fastest.
void grid_setup(const hst_deck &d, hst_grid_info &g, std::vector<double> &y, std::vector<double> &dyl,
                std::vector<double> &der) {
  g.nxd = 3 * (d.nx + 1) / 2;
  while (!fft_fit(g.nxd)) g.nxd++;
  g.nzd = 3 * d.nz;
  while (!fft_fit(g.nzd)) g.nzd++;
  g.nxm = d.nx + 1;
  g.nzm = 2 * d.nz + 1;
  g.nyg = d.ny + 4;
  g.ncol = g.nxm * g.nzm;
  g.lx = 2.0 * PI / d.alfa0;
  g.lz = 2.0 * PI / d.beta0;
  g.dx = PI / (d.alfa0 * g.nxd);
  g.dz = 2.0 * PI / (d.beta0 * g.nzd);
  g.dy = d.ly / d.ny;
  g.factor = 1.0 / (2.0 * g.nxd * g.nzd);
  g.ni = 1.0 / d.re;
  const int ny = d.ny;
  y.assign(ny + 4, 0.0);
  dyl.assign(ny, 0.0);
  der.assign(20 * (size_t)ny, 0.0);
  for (int iy = -2; iy <= ny + 1; iy++) {
    int m = ((iy % ny) + ny) % ny;
    y[iy + 2] = d.ly * ymap(d.ystretch, (double)m / ny) + d.ly * (double)((iy - m) / ny);
  }
  for (int iy = 0; iy < ny; iy++) dyl[iy] = 0.5 * (y[iy + 3] - y[iy + 1]);
  auto DER = [&](int iy, int n, int j) -> double & { return der[(size_t)iy + (size_t)ny * (n + 4 * (j + 2))]; };
  for (int iy = 0; iy < ny; iy++) {
    double dyj[5], M[5][5], t[5], x[5];
    for (int j = 0; j < 5; j++) dyj[j] = y[iy + j] - y[iy + 2];
    // D4: exact on y^4 (= 24) and zero on y^0..y^3
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = powi(dyj[j], 4 - i);
    for (int i = 0; i < 5; i++) t[i] = 0.0;
    t[0] = 24.0;
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(iy, 3, j - 2) = x[j];
    // D0: sum d0 f'''' = sum d4 f for y^8 .. y^4
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = (double)((5 - i) * (6 - i) * (7 - i) * (8 - i)) * powi(dyj[j], 4 - i);
    for (int i = 0; i < 5; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += DER(iy, 3, j - 2) * powi(dyj[j], 8 - i);
      t[i] = s;
    }
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(iy, 0, j - 2) = x[j];
    // D2, D1 consistent with that D0
    double Mp[5][5];
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) Mp[i][j] = powi(dyj[j], 4 - i);
    std::memcpy(M, Mp, sizeof(M));
    for (int i = 0; i < 5; i++) t[i] = 0.0;
    for (int i = 0; i <= 2; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += DER(iy, 0, j - 2) * (double)(4 - i) * (double)(3 - i) * powi(dyj[j], 2 - i);
      t[i] = s;
    }
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(iy, 2, j - 2) = x[j];
    std::memcpy(M, Mp, sizeof(M));
    for (int i = 0; i < 5; i++) t[i] = 0.0;
    for (int i = 0; i <= 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += DER(iy, 0, j - 2) * (double)(4 - i) * powi(dyj[j], 3 - i);
      t[i] = s;
    }
    solve5(M, t, x);
    for (int j = 0; j < 5; j++) DER(iy, 1, j - 2) = x[j];
  }
}

// The elimination of the D0 matrix is the same for every column (its rows
// do not depend on the wavenumber), so its pivots, multipliers, border
// columns and the real half of the record are computed once here, in the
// order of the per-column sweep (bordering with one slab).
void d0_factor(const std::vector<double> &der, int ny, D0Factor &f) {
  auto d0 = [&](int iy, int j) { return der[(size_t)iy + (size_t)ny * (0 + 4 * (j + 2))]; };
  const int mi = ny - 2;
  f.rows.assign((size_t)mi * D0Factor::NCOL, 0.0);
  double u11 = 0, u12 = 0, p1 = 0, q1 = 0, u21 = 0, u22 = 0, p2 = 0, q2 = 0;
  double w1 = 0, w2 = 0, v1 = 0, v2 = 0, sp0 = 0, sq0 = 0, sp1 = 0, sq1 = 0;
  for (int i = 0; i < mi; i++) {
    double a[5];
    for (int j = -2; j <= 2; j++) a[j + 2] = d0(i, j);
    double p = 0.0, q = 0.0, lA = 0.0, lB = 0.0;
    if (i == 0) { p = a[0]; q = a[1]; }
    if (i == 1) q = a[0];
    if (i >= 2) {
      double l = a[0];
      lA = l;
      a[1] = a[1] - l * u21;
      a[2] = a[2] - l * u22;
      p = p - l * p2;
      q = q - l * q2;
    }
    if (i >= 1) {
      double l = a[1];
      lB = l;
      a[2] = a[2] - l * u11;
      a[3] = a[3] - l * u12;
      p = p - l * p1;
      q = q - l * q1;
    }
    double rp = 1.0 / a[2];
    double t1 = a[3] * rp, t2 = a[4] * rp;
    p = p * rp;
    q = q * rp;
    double w = -(u11 * w1 + u22 * w2);
    if (i == 0) w = 1.0;
    double v = -(u11 * v1 + u22 * v2);
    if (i == 1) v = 1.0;
    sp0 = sp0 + w * p; sq0 = sq0 + w * q;
    sp1 = sp1 + v * p; sq1 = sq1 + v * q;
    double *r = &f.rows[(size_t)i * D0Factor::NCOL];
    r[D0Factor::L2] = lA; r[D0Factor::L1] = lB; r[D0Factor::RP] = rp;
    r[D0Factor::U1] = t1; r[D0Factor::U2] = t2; r[D0Factor::B1] = p; r[D0Factor::B2] = q;
    r[D0Factor::W] = w; r[D0Factor::V] = v;
    w2 = w1; w1 = w; v2 = v1; v1 = v;
    u21 = u11; u22 = u12; p2 = p1; q2 = q1;
    u11 = t1; u12 = t2; p1 = p; q1 = q;
  }
  double *c = f.rec;
  c[3] = sp0; c[4] = sq0; c[5] = w2 * u22 + w1 * u11; c[6] = w1 * u12;
  c[9] = sp1; c[10] = sq1; c[11] = v2 * u22 + v1 * u11; c[12] = v1 * u12;
  c[15] = p2 - u21 * p1; c[16] = q2 - u21 * q1; c[17] = u22 - u21 * u11; c[18] = -u21 * u12;
  c[21] = p1; c[22] = q1; c[23] = u11; c[24] = u12;
  c[0] = u21;  // the multiplier of row mi-2's value: x(mi-2) = x2 - u21 x1 - ...
  const int ga = ny - 2;
  f.c1m4 = d0(ga, -2); f.c1m3 = d0(ga, -1); f.d11 = d0(ga, 0); f.d12 = d0(ga, 1); f.c10 = d0(ga, 2);
  f.c2m3 = d0(ga + 1, -2); f.d21 = d0(ga + 1, -1); f.d22 = d0(ga + 1, 0); f.c20 = d0(ga + 1, 1); f.c21 = d0(ga + 1, 2);
}

// uniform_from_key: a deviate in [0, 1) keyed on the seed and the mode
// indices, an fmix32-shaped avalanche with every intermediate masked to 32 bits.
static double uniform_from_key(int seed, int component, int iy, int iz, int ix) {
  const int64_t MASK32 = 4294967295LL, MIX_A = 2146121005LL, MIX_B = 2032289749LL;
  int64_t h = (int64_t)seed;
  h += 2654435761LL * (int64_t)component;
  h += 2246822519LL * (int64_t)iy;
  h += 3266489917LL * (int64_t)iz;
  h += 668265263LL * (int64_t)ix;
  h &= MASK32;
  h ^= (h >> 16);
  h = (h * MIX_A) & MASK32;
  h ^= (h >> 13);
  h = (h * MIX_B) & MASK32;
  h ^= (h >> 16);
  return (double)h * (1.0 / 4294967296.0);
}

// The seeded divergence-free field u = i k x A, or the Taylor-Green vortex,
// into the reference layout V(iy, iz, ix, c) with the ghost rows.
void initial_field(const hst_deck &d, const std::vector<double> &y, std::complex<double> *V) {
  typedef std::complex<double> cx;
  const cx I(0.0, 1.0);
  const int nx = d.nx, nz = d.nz, ny = d.ny, NZm = 2 * nz + 1, NYg = ny + 4;
  const size_t nf = (size_t)(nx + 1) * NZm * NYg;
  std::memset((void *)V, 0, sizeof(cx) * 3 * nf);
  auto at = [&](int c, int ix, int iz) { return V + (size_t)c * nf + ((size_t)ix * NZm + (iz + nz)) * NYg; };
  if (d.initial_condition == 1) {
    const double ky = 2.0 * PI * d.tg_mode / d.ly;
    const int ix = d.tg_mode;
    for (int iz = -d.tg_mode; iz <= d.tg_mode; iz += 2 * d.tg_mode) {
      cx *u = at(0, ix, iz), *v = at(1, ix, iz);
      for (int r = 0; r < NYg; r++) {
        u[r] = -I * d.tg_amplitude * std::cos(ky * y[r]) / 4.0;
        v[r] = -d.tg_amplitude * std::sin(ky * y[r]) / 4.0;
      }
    }
    return;
  }
  const int mmax = ny / 3;
  std::vector<cx> ph((size_t)(2 * mmax + 1) * NYg);
  for (int m = -mmax; m <= mmax; m++) {
    const double ky = 2.0 * PI * m / d.ly;
    for (int r = 0; r < NYg; r++) ph[(size_t)(m + mmax) * NYg + r] = cx(std::cos(ky * y[r]), std::sin(ky * y[r]));
  }
#pragma omp parallel for schedule(dynamic)
  for (int ix = 0; ix <= nx; ix++) {
    const double kx = d.alfa0 * ix;
    for (int iz = -nz; iz <= nz; iz++) {
      const double kz = d.beta0 * iz;
      if (ix == 0 && iz == 0) continue;  // no mean profiles: the mean flow is S*y only
      for (int m = -mmax; m <= mmax; m++) {
        const double ky = 2.0 * PI * m / d.ly;
        const double k = std::sqrt(kx * kx + ky * ky + kz * kz);
        if (k == 0.0) continue;
        const double env = powi(k / d.kpeak, 2) * std::exp(1.0 - powi(k / d.kpeak, 2));
        if (env < 1.0e-8) continue;
        // the potential of mode (ix, iz, m); on ix = 0 the mode (-iz, -m) is the conjugate of (iz, m)
        int jz = iz, jm = m;
        bool mirror = false;
        if (ix == 0 && (iz < 0 || (iz == 0 && m < 0))) { jz = -iz; jm = -m; mirror = true; }
        cx a[3];
        for (int c = 1; c <= 3; c++) {
          double r = uniform_from_key(d.seed, c, jm, jz, ix);
          double phase = TWOPI * uniform_from_key(d.seed, c + 3, jm, jz, ix);
          a[c - 1] = r * cx(std::cos(phase), std::sin(phase));
          if (mirror) a[c - 1] = std::conj(a[c - 1]);
        }
        for (int c = 0; c < 3; c++) a[c] = d.amplitude * env * a[c];
        const cx u0 = I * (ky * a[2] - kz * a[1]);
        const cx u1 = I * (kz * a[0] - kx * a[2]);
        const cx u2 = I * (kx * a[1] - ky * a[0]);
        const cx *p = &ph[(size_t)(m + mmax) * NYg];
        cx *v0 = at(0, ix, iz), *v1 = at(1, ix, iz), *v2 = at(2, ix, iz);
        for (int r = 0; r < NYg; r++) {
          v0[r] += u0 * p[r];
          v1[r] += u1 * p[r];
          v2[r] += u2 * p[r];
        }
      }
    }
  }
}

// sum of w |u|^2 over the modes of the seeded field (see hst_initial_energy)
double initial_energy(const hst_deck &d) {
  typedef std::complex<double> cx;
  const cx I(0.0, 1.0);
  const int nx = d.nx, nz = d.nz, mmax = d.ny / 3;
  std::vector<double> part(nx + 1, 0.0);  // per x mode, summed in order: the result does not depend on threads
#pragma omp parallel for schedule(dynamic)
  for (int ix = 0; ix <= nx; ix++) {
    const double kx = d.alfa0 * ix, wgt = ix == 0 ? 1.0 : 2.0;
    double q2 = 0.0;
    for (int iz = -nz; iz <= nz; iz++) {
      const double kz = d.beta0 * iz;
      if (ix == 0 && iz == 0) continue;
      for (int m = -mmax; m <= mmax; m++) {
        const double ky = 2.0 * PI * m / d.ly;
        const double k = std::sqrt(kx * kx + ky * ky + kz * kz);
        if (k == 0.0) continue;
        const double env = powi(k / d.kpeak, 2) * std::exp(1.0 - powi(k / d.kpeak, 2));
        if (env < 1.0e-8) continue;
        int jz = iz, jm = m;
        bool mirror = false;
        if (ix == 0 && (iz < 0 || (iz == 0 && m < 0))) { jz = -iz; jm = -m; mirror = true; }
        cx a[3];
        for (int c = 1; c <= 3; c++) {
          double r = uniform_from_key(d.seed, c, jm, jz, ix);
          double phase = TWOPI * uniform_from_key(d.seed, c + 3, jm, jz, ix);
          a[c - 1] = r * cx(std::cos(phase), std::sin(phase));
          if (mirror) a[c - 1] = std::conj(a[c - 1]);
        }
        for (int c = 0; c < 3; c++) a[c] = d.amplitude * env * a[c];
        const cx u0 = I * (ky * a[2] - kz * a[1]);
        const cx u1 = I * (kz * a[0] - kx * a[2]);
        const cx u2 = I * (kx * a[1] - ky * a[0]);
        q2 += wgt * (std::norm(u0) + std::norm(u1) + std::norm(u2));
      }
    }
    part[ix] = q2;
  }
  double q2 = 0.0;
  for (double v : part) q2 += v;
  return q2;
}

Dims make_dims(const hst_deck &d, const hst_grid_info &g) {
  Dims D;
  D.nx = d.nx; D.ny = d.ny; D.nz = d.nz;
  D.NXm = g.nxm; D.NZm = g.nzm; D.NYg = g.nyg; D.ncol = g.ncol;
  D.ncs = (g.ncol + 255) / 256 * 256;
  D.nxd = g.nxd; D.nzd = g.nzd;
  D.NXp = (D.NXm + 15) / 16 * 16;
  D.alfa0 = d.alfa0; D.beta0 = d.beta0; D.ni = g.ni; D.S = d.S;
  D.factor = g.factor; D.dx = g.dx; D.dz = g.dz;
  return D;
}

D0Const d0_const(const D0Factor &f) {
  D0Const c;
  std::memcpy(c.rec, f.rec, sizeof(c.rec));
  c.c1m4 = f.c1m4; c.c1m3 = f.c1m3; c.d11 = f.d11; c.d12 = f.d12; c.c10 = f.c10;
  c.c2m3 = f.c2m3; c.d21 = f.d21; c.d22 = f.d22; c.c20 = f.c20; c.c21 = f.c21;
  return c;
}

}  // namespace hst

extern "C" int hst_initial_energy(const hst_deck *deck, double *q2) {
  if (!deck || !q2) return 1;
  std::string err = hst::check_deck(*deck);
  if (!err.empty()) {
    hst::set_error(err);
    return 1;
  }
  *q2 = hst::initial_energy(*deck);
  return 0;
}

extern "C" int hst_grid_setup(const hst_deck *deck, hst_grid_info *info, double *y, double *dyl, double *der) {
  if (!deck) return 1;
  std::string err = hst::check_deck(*deck);
  if (!err.empty()) {
    hst::set_error(err);
    return 1;
  }
  hst_grid_info g;
  std::vector<double> vy, vdyl, vder;
  hst::grid_setup(*deck, g, vy, vdyl, vder);
  if (info) *info = g;
  if (y) std::memcpy(y, vy.data(), sizeof(double) * vy.size());
  if (dyl) std::memcpy(dyl, vdyl.data(), sizeof(double) * vdyl.size());
  if (der) std::memcpy(der, vder.data(), sizeof(double) * vder.size());
  return 0;
}

extern "C" int hst_initial_field(const hst_deck *deck, double *V_host) {
  if (!deck || !V_host) return 1;
  std::string err = hst::check_deck(*deck);
  if (!err.empty()) {
    hst::set_error(err);
    return 1;
  }
  hst_grid_info g;
  std::vector<double> y, dyl, der;
  hst::grid_setup(*deck, g, y, dyl, der);
  hst::initial_field(*deck, y, reinterpret_cast<std::complex<double> *>(V_host));
  return 0;
}
