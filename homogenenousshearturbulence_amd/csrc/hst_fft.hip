// The spectral <-> physical round trip of the RK3 substep on gfx950:
//
//   k_zift   z inverse transform of u, v, w (3/2 padding fused into the load),
//            written transposed to x-fastest lines (Hv)
//   k_xpass  per (row, z point): x inverse real transform of u, v, w, the six
//            products (and the CFL number), x forward real transforms of the
//            products truncated to the nx+1 kept modes (Hp)
//   k_zfft   z forward transform of the six products, truncated to the 2nz+1
//            kept modes, written to the spectral layout (P)
//
// Every transform is an in-LDS Stockham FFT of radices 8, 4, 2 and 3 with the
// butterflies held in registers between the two barriers of a stage.  Real
// transforms of length 2M run as complex transforms of length M.  Field
// layout on the device: [row][ix][iz] with iz fastest (a Fourier column is
// one (ix, iz) pair; col = ix*(2nz+1) + iz + nz), rows -2..ny+1.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "hst_cx.h"
#include "hst_fft_dev.h"
#include "hst_internal.h"

namespace hst {

template <typename R>
static int fill_table(Cx<R> **dst, int n, double scale_den) {
  std::vector<Cx<R>> h(n);
  for (int k = 0; k < n; k++) {
    const double a = -2.0 * M_PI * (double)k / scale_den;
    h[k] = Cx<R>{(R)std::cos(a), (R)std::sin(a)};
  }
  if (hipMalloc(dst, sizeof(Cx<R>) * n) != hipSuccess) return 1;
  return hipMemcpy(*dst, h.data(), sizeof(Cx<R>) * n, hipMemcpyHostToDevice) != hipSuccess;
}

bool fft_size_supported(int n) {
  switch (n) {
    case 6: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
    case 384: case 512: case 768: case 1024: case 1536:
      return true;
    default:
      return false;
  }
}

int make_fft_tables(const Dims &D, int mixed, void **out) {
  TableSet *t = new TableSet();
  t->mixed = mixed;
  int rc = 0;
  if (!mixed) {
    rc |= fill_table(&t->d.twz, D.nzd, D.nzd);
    rc |= fill_table(&t->d.twx, D.nxd, D.nxd);
    rc |= fill_table(&t->d.twh, D.nxd, 2.0 * D.nxd);
  } else {
    rc |= fill_table(&t->f.twz, D.nzd, D.nzd);
    rc |= fill_table(&t->f.twx, D.nxd, D.nxd);
    rc |= fill_table(&t->f.twh, D.nxd, 2.0 * D.nxd);
  }
  *out = t;
  return rc;
}

void free_fft_tables(void *p) {
  TableSet *t = (TableSet *)p;
  if (!t) return;
  if (!t->mixed) { (void)hipFree(t->d.twz); (void)hipFree(t->d.twx); (void)hipFree(t->d.twh); }
  else { (void)hipFree(t->f.twz); (void)hipFree(t->f.twx); (void)hipFree(t->f.twh); }
  delete t;
}

// ------------------------------------------------------------------ launch shapes

// Lines per block of the z kernels, measured at nzd = 768: FP32 four (more
// blocks per CU than eight, the XCD order keeps the tiles that share a line in
// one L2); FP64 two in both directions (four: 1.5x slower).  Both copy the
// twiddles to LDS (FP64 12 KB a block): read through the cache inside the
// stages, each stage waited for its loads (FP64 k_zift 1.75 -> 1.58-1.70 ms,
// k_zfft 2.43 -> 2.29-2.33 ms with the copy).
#ifndef HST_ZG64_FWD
#define HST_ZG64_FWD 2
#endif
#ifndef HST_ZG64_INV
#define HST_ZG64_INV 2
#endif
#ifndef HST_ZTW_LDS64
#define HST_ZTW_LDS64 1
#endif
#ifndef HST_ZTW_LDS32
#define HST_ZTW_LDS32 1
#endif
template <typename R>
constexpr bool ztw_lds() { return sizeof(R) == 4 ? HST_ZTW_LDS32 : HST_ZTW_LDS64; }

template <typename R, int N, bool FWD>
struct ZCfg {
  static constexpr int T = fft_threads(N);
  static constexpr int LDS_MAX = 100 * 1024;
  static constexpr int G0 = sizeof(R) == 8 ? (FWD ? HST_ZG64_FWD : HST_ZG64_INV) : 4;
  // LDS line stride: lanes of the transposed accesses walk the lines (stride LS)
  // and the points (stride 1); LS = 16 dwords mod 64 spreads them over all banks
  static constexpr int DW = (int)sizeof(Cx<R>) / 4, PER = 64 / DW;
  static constexpr int LS = N + (((16 / DW) - N % PER) % PER + PER) % PER;
  static constexpr int g_lds(int g) { return (g > 1 && g * LS * (int)sizeof(Cx<R>) > LDS_MAX) ? g_lds(g / 2) : g; }
  static constexpr int g_thr(int g) { return (g > 1 && g * T > 1024) ? g_thr(g / 2) : g; }
  static constexpr int G = g_thr(g_lds(G0));
  static constexpr int NT = G * T;
};

template <typename R, int M>
struct XCfg {
  static constexpr int T = fft_threads(M);
  static constexpr int LDS_MAX = 100 * 1024;
  static constexpr int per_line = 4 * (M + 1) * (int)sizeof(Cx<R>);
  static constexpr int g_fit(int g) { return (g > 1 && (g * per_line > LDS_MAX || g * T > 512)) ? g_fit(g / 2) : g; }
  static constexpr int G = g_fit(4);
  static constexpr int NT = G * T;
};

// ------------------------------------------------------------------ transform buffers

// ------------------------------------------------------------------ XCD-aware block order

// Workgroups are dealt to the 8 XCDs round-robin by their linear id, and each
// XCD has its own L2.  The z kernels read (k_zfft) or write (k_zift) segments
// of G x points that share 128-byte lines with the next x tile, so the linear
// id is renumbered to give every XCD a contiguous range of tiles: neighbours
// then meet in the same L2 instead of fetching the line twice.  Returns
// (tile, row, component) for a grid (ntile, nrow, nc).
__device__ __forceinline__ void xcd_tile(int &tile, int &row, int &comp) {
  const int nt = gridDim.x, nr = gridDim.y, l = xcd_linear();
  tile = l % nt;
  row = (l / nt) % nr;
  comp = l / (nt * nr);
}

// dst[idx(k)] = f(k) for k < N by the NT threads of a block (blockDim.x ==
// NT): every load of a thread is issued before its first store.  A plain
// strided loop stores each value as soon as it is loaded, so it waits for
// one load after another.
template <int N, int NT, typename T, class F, class IDX>
__device__ __forceinline__ void lds_fill(T *dst, F f, IDX idx) {
  constexpr int PER = (N + NT - 1) / NT;
  T v[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int k = (int)threadIdx.x + u * NT;
    if (N % NT == 0 || k < N) v[u] = f(k);
  }
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int k = (int)threadIdx.x + u * NT;
    if (N % NT == 0 || k < N) dst[idx(k)] = v[u];
  }
}
template <int N, int NT, typename T, class F>
__device__ __forceinline__ void lds_fill(T *dst, F f) {
  lds_fill<N, NT>(dst, f, [](int k) { return k; });
}

// ------------------------------------------------------------------ k_zift

// z inverse transform of the three velocity components for the rows of a
// chunk.  grid (x tiles of G modes, rows, component); output Hv[c][row][zp][ix].
// CMP: V1 holds eta and V3 dv/dy (the time step's form between its
// substeps); grid (npt + nvt, rows): the first npt = ceil(NXm / (G/2)) blocks
// of a row recover u and w of G/2 modes as they load them (recover_uw) and
// transform the G lines, the other nvt = ceil(NXm / G) transform v of G modes.
template <typename R, int NZD, bool CMP>
__global__ void __launch_bounds__((ZCfg<R, NZD, false>::NT))
    k_zift(const double2 *__restrict__ V1, const double2 *__restrict__ V2, const double2 *__restrict__ V3,
           Cx<R> *__restrict__ Hv, Dims D, int row0, int nrows, const Cx<R> *__restrict__ tw) {
  constexpr int G = ZCfg<R, NZD, false>::G, T = ZCfg<R, NZD, false>::T, LS = ZCfg<R, NZD, false>::LS;
  static_assert(!CMP || G % 2 == 0, "the u, w blocks split a tile in two");
  constexpr int HG = G / 2, NT = ZCfg<R, NZD, false>::NT, PER = (NZD + NT - 1) / NT;
  int tile, rowc, c;
  xcd_tile(tile, rowc, c);
  const int npt = (D.NXm + HG - 1) / HG;
  const bool pair = CMP && tile < npt;
  if (CMP && !pair) {
    tile -= npt;
    c = 1;
  }
  __shared__ Cx<R> lds[G * LS];
  __shared__ Cx<R> twl[ztw_lds<R>() ? NZD : 1];  // the twiddles: lanes of a stage read scattered entries
  const Cx<R> *twp = ztw_lds<R>() ? twl : tw;
  const int row = row0 + rowc, nz = D.nz;
  const size_t rb = (size_t)row * D.ncol;
  if (!pair) {
    const int cc = CMP ? 1 : c, ix0 = tile * G;
    const double2 *src = (cc == 0 ? V1 : (cc == 1 ? V2 : V3)) + rb + (size_t)ix0 * D.NZm;
    if constexpr (ztw_lds<R>()) lds_fill<NZD, NT>(twl, [&](int k) { return tw[k]; });
    lds_fill<G * NZD, NT>(
        lds,
        [&](int e) {
          const int j = e / NZD, k = e - j * NZD;
          const bool in = ix0 + j < D.NXm && (k <= nz || k >= NZD - nz);
          const int iz = in ? (k <= nz ? k : k - NZD) : 0, jc = ix0 + j < D.NXm ? j : 0;
          const Cx<R> v = cx_from<R>(src[(size_t)jc * D.NZm + iz + nz]);
          return in ? v : Cx<R>{(R)0, (R)0};
        },
        [&](int e) { const int j = e / NZD; return j * LS + (e - j * NZD); });
    __syncthreads();
    const int g = threadIdx.x / T, t = threadIdx.x % T;
    fft_line<R, NZD, T, +1>(lds + g * LS, t, twp);
    Cx<R> *dst = Hv + ((size_t)cc * nrows + rowc) * NZD * D.NXp;
    for (int e = threadIdx.x; e < G * NZD; e += blockDim.x) {
      const int j = e % G, zp = e / G;
      if (ix0 + j < D.NXm) dst[hoff<R>(NZD, D.NXp, zp, ix0 + j)] = lds[j * LS + zp];
    }
  } else {
    const int ix0 = tile * HG;
    // every load of the block's lines first, then the recovery; the loads
    // are unconditional (a lane with no mode reads the column (ix, 0)), as a
    // load under a branch is waited for where the branch ends
    Cx<double> xv[HG][PER], ev[HG][PER];
#pragma unroll
    for (int j = 0; j < HG; j++) {
      const size_t c0 = rb + (size_t)min(ix0 + j, D.NXm - 1) * D.NZm + nz;  // the column (ix, 0)
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int k = threadIdx.x + q * NT;
        const bool in = ix0 + j < D.NXm && k < NZD && (k <= nz || k >= NZD - nz);
        const int iz = in ? (k <= nz ? k : k - NZD) : 0;
        const Cx<double> x = ld(V3 + c0 + iz), e = ld(V1 + c0 + iz);
        xv[j][q] = in ? x : Cx<double>{0.0, 0.0};
        ev[j][q] = in ? e : Cx<double>{0.0, 0.0};
      }
    }
    // the twiddles, while the lines' loads are in flight
    if constexpr (ztw_lds<R>()) lds_fill<NZD, NT>(twl, [&](int k) { return tw[k]; });
#pragma unroll
    for (int j = 0; j < HG; j++) {
      const int ix = ix0 + j;
      const double al = D.alfa0 * ix;
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int k = threadIdx.x + q * NT;
        if (k < NZD) {
          const int iz = k <= nz ? k : k - NZD;
          const double be = D.beta0 * iz;
          Cx<double> ud, wd;
          recover_uw(al, be, al * al + be * be, ix == 0 && iz == 0, xv[j][q], ev[j][q], ud, wd);
          lds[j * LS + k] = Cx<R>{(R)ud.x, (R)ud.y};
          lds[(HG + j) * LS + k] = Cx<R>{(R)wd.x, (R)wd.y};
        }
      }
    }
    __syncthreads();
    const int g = threadIdx.x / T, t = threadIdx.x % T;
    fft_line<R, NZD, T, +1>(lds + g * LS, t, twp);
    for (int h = 0; h < 2; h++) {
      Cx<R> *dst = Hv + ((size_t)(2 * h) * nrows + rowc) * NZD * D.NXp;
      for (int e = threadIdx.x; e < HG * NZD; e += blockDim.x) {
        const int j = e % HG, zp = e / HG;
        if (ix0 + j < D.NXm) dst[hoff<R>(NZD, D.NXp, zp, ix0 + j)] = lds[(h * HG + j) * LS + zp];
      }
    }
  }
}

// ------------------------------------------------------------------ k_zfft

template <typename R, int NZD>
__global__ void __launch_bounds__((ZCfg<R, NZD, true>::NT))
    k_zfft(const Cx<R> *__restrict__ Hp, Cx<R> *__restrict__ P0, size_t pstride, Dims D, int row0, int nrows,
           const Cx<R> *__restrict__ tw) {
  constexpr int G = ZCfg<R, NZD, true>::G, T = ZCfg<R, NZD, true>::T, LS = ZCfg<R, NZD, true>::LS;
  constexpr int NT = ZCfg<R, NZD, true>::NT;
  __shared__ Cx<R> lds[G * LS];
  __shared__ Cx<R> twl[ztw_lds<R>() ? NZD : 1];
  int tile, rowc, p;
  xcd_tile(tile, rowc, p);
  const int row = row0 + rowc, ix0 = tile * G;
  const Cx<R> *src = Hp + ((size_t)p * nrows + rowc) * NZD * D.NXp;
  // the lines and the twiddles: every load in flight before the first store
  constexpr int PER = (G * NZD + NT - 1) / NT, PT = ztw_lds<R>() ? (NZD + NT - 1) / NT : 0;
  Cx<R> v[PER], tv[PT > 0 ? PT : 1];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = threadIdx.x + u * NT, j = e % G, zp = e / G;
    if ((G * NZD) % NT == 0 || e < G * NZD)
      v[u] = (ix0 + j < D.NXm) ? src[hoff<R>(NZD, D.NXp, zp, ix0 + j)] : Cx<R>{(R)0, (R)0};
  }
#pragma unroll
  for (int u = 0; u < PT; u++) {
    const int k = threadIdx.x + u * NT;
    if (NZD % NT == 0 || k < NZD) tv[u] = tw[k];
  }
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = threadIdx.x + u * NT;
    if ((G * NZD) % NT == 0 || e < G * NZD) lds[(e % G) * LS + e / G] = v[u];
  }
#pragma unroll
  for (int u = 0; u < PT; u++) {
    const int k = threadIdx.x + u * NT;
    if (NZD % NT == 0 || k < NZD) twl[k] = tv[u];
  }
  const Cx<R> *twp = ztw_lds<R>() ? twl : tw;
  __syncthreads();
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  const int nz = D.nz;
  if constexpr (sizeof(R) == 4) {
    // FP32: the last stage stores the kept modes of line g straight to P (the
    // modes of one x mode are consecutive there), measured 4 % faster at
    // nzd = 768; FP64 measured slower that way and goes through LDS
    Cx<R> *dst = P0 + (size_t)p * pstride + (size_t)row * D.ncol + (size_t)(ix0 + g) * D.NZm + nz;
    const bool okg = ix0 + g < D.NXm;
    fft_line_emit<R, NZD, T, -1>(lds + g * LS, t, twp, [&](int k, const Cx<R> &val) {
      if (okg && k <= nz) dst[k] = val;
      else if (okg && k >= NZD - nz) dst[k - NZD] = val;
    });
  } else {
    fft_line<R, NZD, T, -1>(lds + g * LS, t, twp);
    Cx<R> *dst = P0 + (size_t)p * pstride + (size_t)row * D.ncol + (size_t)ix0 * D.NZm;
    for (int j = 0; j < G; j++) {
      if (ix0 + j >= D.NXm) break;
      for (int izm = threadIdx.x; izm < D.NZm; izm += blockDim.x) {
        const int iz = izm - nz;
        dst[(size_t)j * D.NZm + izm] = lds[j * LS + (iz >= 0 ? iz : NZD + iz)];
      }
    }
  }
}

// ------------------------------------------------------------------ products

// The five combinations of the products a substep needs (MODE 3 of the x
// passes): vv - uu, vv - ww, uv, vw, uw, each times factor, on the even (x)
// and odd (y) points of a line.  The right-hand sides depend on uu, vv, ww
// only through these two differences (and through uv, vw, uw), so five
// spectra replace the six of build_products.
template <typename R>
__device__ __forceinline__ Cx<R> combined_product(int p, const Cx<R> &u, const Cx<R> &v, const Cx<R> &w, R factor) {
  switch (p) {
    case 0: return Cx<R>{(v.x * v.x - u.x * u.x) * factor, (v.y * v.y - u.y * u.y) * factor};
    case 1: return Cx<R>{(v.x * v.x - w.x * w.x) * factor, (v.y * v.y - w.y * w.y) * factor};
    case 2: return Cx<R>{u.x * v.x * factor, u.y * v.y * factor};
    case 3: return Cx<R>{v.x * w.x * factor, v.y * w.y * factor};
    default: return Cx<R>{u.x * w.x * factor, u.y * w.y * factor};
  }
}

// ------------------------------------------------------------------ k_xpass

// One (row, z point) line per group of T threads, G lines per block.
// MODE 0: six products; 1: the round trip (u, v, w times factor); 2: CFL only.
template <typename R, int M, int MODE>
__global__ void __launch_bounds__((XCfg<R, M>::NT))
    k_xpass(const Cx<R> *__restrict__ Hv, Cx<R> *__restrict__ Hp, Dims D, int row0, int nrows,
            const Cx<R> *__restrict__ twx, const Cx<R> *__restrict__ twh, const double *__restrict__ dyl,
            double *__restrict__ cfl_part, int want_cfl) {
  constexpr int G = XCfg<R, M>::G, T = XCfg<R, M>::T, LS = M + 1;
  __shared__ Cx<R> vel[3 * G * LS];
  __shared__ Cx<R> work[G * LS];
  __shared__ double red[XCfg<R, M>::NT / 64 + 1];
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  const int rowc = blockIdx.y, row = row0 + rowc;
  const int zp = blockIdx.x * G + g;  // nzd is a multiple of G (checked by the launcher)
  const int nx = D.nx, NZD = D.nzd;
  Cx<R> *wl = work + g * LS;
  // X -> Z of the half-length complex inverse transform, per component
  for (int c = 0; c < 3; c++) {
    const Cx<R> *src = Hv + ((size_t)c * nrows + rowc) * NZD * D.NXp;
    for (int k = t; k <= nx; k += T) wl[k] = src[hoff<R>(NZD, D.NXp, zp, k)];
    __syncthreads();
    Cx<R> *vl = vel + (c * G + g) * LS;
    for (int k = t; k < M; k += T) {
      Cx<R> a = (k <= nx) ? wl[k] : Cx<R>{(R)0, (R)0};
      if (k == 0) a.y = (R)0;  // c2r: the imaginary part of the kx = 0 bin is ignored
      const int j = M - k;
      Cx<R> b = (k > 0 && j <= nx) ? conj(wl[j]) : Cx<R>{(R)0, (R)0};
      Cx<R> e = conj(twh[k]);  // exp(+i pi k / M)
      vl[k] = (a + b) + mul_i(e * (a - b));
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 3; c++) fft_line<R, M, T, +1>(vel + (c * G + g) * LS, t, twx);
  // vel[c][m] = x[2m] + i x[2m+1]
  const Cx<R> *u = vel + (0 * G + g) * LS, *v = vel + (1 * G + g) * LS, *w = vel + (2 * G + g) * LS;
  if (want_cfl) {
    double cm = 0.0;
    const int iy = row - 2;
    if (iy >= 0 && iy < D.ny) {
      const double idx = 1.0 / D.dx, idy = 1.0 / dyl[iy], idz = 1.0 / D.dz;
      for (int m = t; m < M; m += T) {
        double a = fabs((double)u[m].x) * idx + fabs((double)v[m].x) * idy + fabs((double)w[m].x) * idz;
        double b = fabs((double)u[m].y) * idx + fabs((double)v[m].y) * idy + fabs((double)w[m].y) * idz;
        cm = fmax(cm, fmax(a, b));
      }
    }
    for (int o = 32; o > 0; o >>= 1) cm = fmax(cm, __shfl_xor(cm, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cm;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0.0;
      for (int i = 0; i < (int)((blockDim.x + 63) >> 6); i++) r = fmax(r, red[i]);
      cfl_part[blockIdx.y * gridDim.x + blockIdx.x] = r;
    }
  }
  if constexpr (MODE == 2) return;
  const R factor = (R)D.factor;
  constexpr int NP = (MODE == 0 ? 6 : (MODE == 3 ? 5 : 3));
  for (int p = 0; p < NP; p++) {
    for (int m = t; m < M; m += T) {
      Cx<R> r;
      if constexpr (MODE == 1) {
        const Cx<R> a = (p == 0 ? u[m] : (p == 1 ? v[m] : w[m]));
        r = Cx<R>{a.x * factor, a.y * factor};
      } else if constexpr (MODE == 3) {
        r = combined_product<R>(p, u[m], v[m], w[m], factor);
      } else {
        const Cx<R> a = (p == 0 || p == 3 || p == 5) ? u[m] : (p == 1 || p == 4 ? v[m] : w[m]);
        const Cx<R> b = (p == 0) ? u[m] : (p == 1 || p == 3) ? v[m] : w[m];
        // uu vv ww uv vw uw, each (a b) factor on the even and the odd points
        r = Cx<R>{a.x * b.x * factor, a.y * b.y * factor};
      }
      wl[m] = r;
    }
    __syncthreads();
    fft_line<R, M, T, -1>(wl, t, twx);
    Cx<R> *dst = Hp + ((size_t)p * nrows + rowc) * NZD * D.NXp;
    for (int k = t; k <= nx; k += T) {
      const Cx<R> Z = wl[k], Zc = conj(wl[(M - k) % M]);
      const Cx<R> e = twh[k];  // exp(-i pi k / M)
      const Cx<R> s = (Z + Zc) * (R)0.5, d = (Z - Zc) * (R)0.5;
      dst[hoff<R>(NZD, D.NXp, zp, k)] = s - mul_i(e * d);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ k_xline

// The x pass for power-of-two lengths M >= 64: T = M/8 threads per line,
// E = 8 elements per thread; a block is one wavefront holding 64/T lines, or
// at M = 1024 two wavefronts holding one line.  A
// thread holds the elements t + T e (e < E) of a line in registers; the
// Stockham stages exchange through LDS only between stages, and the last
// stage of a transform leaves the elements t + T e in the registers of
// thread t again, which is where the first stage of the next transform
// reads them: the products are formed in registers between the inverse and
// the forward transforms.  LDS lines are padded by one element per eight.
template <int M>
struct XL {
  static constexpr int T = M / 8;
  static constexpr int E = M / T;
  // lines per block: at least four neighbouring z points (a tile's four
  // points are one 128-byte run of the product buffer), a wavefront at least
  static constexpr int LPB = 64 / T > 4 ? 64 / T : 4;
  static constexpr int NTH = LPB * T;  // threads per block
  static constexpr int PL = M + M / 8;
  // kept x modes a line can have: nx + 1 <= 2 nxd / 3 (the 3/2 rule)
  static constexpr int NXS = 2 * M / 3 + 2;
};

__device__ __forceinline__ int padx(int p) { return p + (p >> 3); }

// The exchange between two Stockham stages of a line: a line of up to 64
// threads lives in one wavefront and in its own part of LDS, so the
// wavefront need only wait for its own LDS traffic (the LDS serves a
// wavefront's requests in order; the fences keep the compiler from moving
// them across); the 1024-point lines span two wavefronts and take the
// block's barrier.
template <int M>
__device__ __forceinline__ void xl_sync() {
  if constexpr (XL<M>::T <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// The x pass's twiddles in LDS, stage by stage: a stage of stride S and
// radix RAD (not the last) multiplies output k2 of butterfly n1 by
// w^(n1 k2 S), w = exp(-2 pi i / M); its table holds them [k2 - 1][n1], so
// that the lanes of a wavefront, whose n1 are consecutive (or equal), read
// neighbouring words (read as tw[n1 k2 S], a stage's lanes hit one bank).
template <int M, int ST = 0, int S = 1>
__host__ __device__ constexpr int xtw_off(int st) {
  if constexpr (ST + 1 >= n_stages(M)) {
    return 0;
  } else {
    constexpr int RAD = radix_at(M, ST), Ls = M / (S * RAD);
    return st <= ST ? 0 : (RAD - 1) * Ls + xtw_off<M, ST + 1, S * RAD>(st);
  }
}
template <int M>
__host__ __device__ constexpr int xtw_size() { return xtw_off<M>(n_stages(M)); }
// the exponent n1 k2 S of entry i of the staged table
template <int M, int ST = 0, int S = 1>
__device__ __forceinline__ int xtw_exponent(int i) {
  if constexpr (ST + 1 >= n_stages(M)) {
    return 0;
  } else {
    constexpr int RAD = radix_at(M, ST), Ls = M / (S * RAD), n = (RAD - 1) * Ls;
    if (i < n) return (i % Ls) * (1 + i / Ls) * S;
    return xtw_exponent<M, ST + 1, S * RAD>(i - n);
  }
}

template <typename R, int M, int SIGN, int ST, int S>
__device__ __forceinline__ void xl_fft(Cx<R> (&v)[XL<M>::E], Cx<R> *L, int t, const Cx<R> *__restrict__ tw) {
  constexpr int T = XL<M>::T, E = XL<M>::E;
  if constexpr (ST < n_stages(M)) {
    constexpr int RAD = radix_at(M, ST);
    constexpr int Ls = M / (S * RAD);
    constexpr int NB = E / RAD;
    constexpr bool FIRST = (S == 1), LAST = (S * RAD == M);
    Cx<R> c[NB][RAD];
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if constexpr (FIRST) {
#pragma unroll
        for (int n2 = 0; n2 < RAD; n2++) c[j][n2] = v[j + NB * n2];
      } else {
        const int b = t + T * j, q = b % S, n1 = b / S;
#pragma unroll
        for (int n2 = 0; n2 < RAD; n2++) c[j][n2] = L[padx(q + S * (n1 + Ls * n2))];
      }
    }
    xl_sync<M>();  // every read of L done before it is written again
#pragma unroll
    for (int j = 0; j < NB; j++) {
      dft<R, SIGN, RAD>(c[j]);
      if constexpr (!LAST) {
        const int b = t + T * j, n1 = b / S;
        constexpr int OFF = xtw_off<M>(ST);
#pragma unroll
        for (int k2 = 1; k2 < RAD; k2++) {
          Cx<R> w = tw[OFF + (k2 - 1) * Ls + n1];
          if (SIGN > 0) w.y = -w.y;
          c[j][k2] = c[j][k2] * w;
        }
      }
    }
    if constexpr (LAST) {
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int k2 = 0; k2 < RAD; k2++) v[j + NB * k2] = c[j][k2];
    } else {
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const int b = t + T * j, q = b % S, n1 = b / S;
#pragma unroll
        for (int k2 = 0; k2 < RAD; k2++) L[padx(q + S * k2 + S * RAD * n1)] = c[j][k2];
      }
      xl_sync<M>();
      xl_fft<R, M, SIGN, ST + 1, S * RAD>(v, L, t, tw);
    }
  }
}

// Z(k) = (X(k) + conj X(M-k)) + i exp(i pi k/M) (X(k) - conj X(M-k)) at k = t + T e:
// the half-length complex input of the length-2M inverse real transform.
// Every load of the line is issued before the first is used: with the loads
// and the arithmetic interleaved per element, each element's conditional
// loads were waited for before the next element's went out.
template <typename R, int M>
__device__ __forceinline__ void xl_c2r_load(Cx<R> (&dst)[XL<M>::E], const Cx<R> *__restrict__ src, int NZD, int NXp,
                                            int zp, bool ok, int t, int nx, const Cx<R> *__restrict__ twh) {
  constexpr int T = XL<M>::T, E = XL<M>::E;
  Cx<R> a[E], b[E], ep[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int k = t + T * e;
    a[e] = Cx<R>{(R)0, (R)0};
    b[e] = Cx<R>{(R)0, (R)0};
    if (ok && k <= nx) a[e] = src[hoff<R>(NZD, NXp, zp, k)];
    if (ok && k > 0 && M - k <= nx) b[e] = src[hoff<R>(NZD, NXp, zp, M - k)];
    ep[e] = twh[k];
  }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int k = t + T * e;
    Cx<R> av = a[e];
    if (k == 0) av.y = (R)0;  // c2r: the imaginary part of the kx = 0 bin is ignored
    const Cx<R> bv = conj(b[e]);
    dst[e] = (av + bv) + mul_i(conj(ep[e]) * (av - bv));  // conj(twh[k]) = exp(+i pi k / M)
  }
}

// occupancy floor (wavefronts per SIMD): the FP64 lines fit in 256 VGPRs (two
// wavefronts per SIMD) with a few bytes of spill instead of taking one; the
// FP32 lines are bound by the vector ALU and run 4 % faster at four
// wavefronts with 36 bytes of spill than at three without
#ifndef XLINE_F32_WAVES
#define XLINE_F32_WAVES 4
#endif
template <typename R, int M>
constexpr int xline_min_waves() { return sizeof(R) == 8 ? 2 : XLINE_F32_WAVES; }

template <typename R, int M, int MODE>
__global__ void __launch_bounds__(XL<M>::NTH, (xline_min_waves<R, M>()))
    k_xline(const Cx<R> *__restrict__ Hv, Cx<R> *__restrict__ Hp, Dims D, int nrows, const Cx<R> *__restrict__ twx,
            const Cx<R> *__restrict__ twh, const double *__restrict__ dyl, double *__restrict__ cfl_part,
            int want_cfl, int row0) {
  constexpr int T = XL<M>::T, E = XL<M>::E, LPB = XL<M>::LPB, PL = XL<M>::PL;
  constexpr int NTH = XL<M>::NTH, HT = htile<R>();
  __shared__ Cx<R> lds[LPB * PL];
  __shared__ Cx<R> twl[xtw_size<M>() > 0 ? xtw_size<M>() : 1];  // the staged twiddles
  lds_fill<xtw_size<M>(), NTH>(twl, [&](int k) { return twx[xtw_exponent<M>(k)]; });
  // exp(-i pi k / M) of the R2C post-processing, read by every product: from
  // LDS (read from global memory inside the product loop, each element waited
  // for its own load)
  __shared__ Cx<R> twhl[M];
  lds_fill<M, NTH>(twhl, [&](int k) { return twh[k]; });
  __syncthreads();
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  const int NZD = D.nzd, nx = D.nx;
  // tiled Hv: neighbouring z points share cache lines, so neighbouring
  // blocks stay on one XCD (one L2)
  const int line0 = xcd_linear() * LPB, line = line0 + g;
  const bool ok = line < NZD * nrows;
  const int rowc = ok ? line / NZD : 0, zp = ok ? line - rowc * NZD : 0;
  // where the block's lines start in a product plane of the tiled buffer
  // (-1: past the last line), for the stores of every product
  __shared__ long long lbase[LPB];
  if (t == 0) lbase[g] = ok ? (long long)rowc * NZD * D.NXp + (long long)zp * HT : -1;
  Cx<R> *L = lds + g * PL;
  Cx<R> u[E], v[E], w[E];
  // C2R pre-processing of u, v, w straight from the x modes in HBM
  const size_t plane = (size_t)NZD * D.NXp;
  xl_c2r_load<R, M>(u, Hv + ((size_t)0 * nrows + rowc) * plane, NZD, D.NXp, zp, ok, t, nx, twh);
  xl_c2r_load<R, M>(v, Hv + ((size_t)1 * nrows + rowc) * plane, NZD, D.NXp, zp, ok, t, nx, twh);
  xl_c2r_load<R, M>(w, Hv + ((size_t)2 * nrows + rowc) * plane, NZD, D.NXp, zp, ok, t, nx, twh);
  xl_fft<R, M, +1, 0, 1>(u, L, t, twl);
  xl_fft<R, M, +1, 0, 1>(v, L, t, twl);
  xl_fft<R, M, +1, 0, 1>(w, L, t, twl);
  // u[e] = x(2k) + i x(2k+1) at k = t + T e
  if (want_cfl) {
    double cm = 0.0;
    const int iy = row0 + rowc - 2;
    if (ok && iy >= 0 && iy < D.ny) {
      const double idx = 1.0 / D.dx, idy = 1.0 / dyl[iy], idz = 1.0 / D.dz;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const double a = fabs((double)u[e].x) * idx + fabs((double)v[e].x) * idy + fabs((double)w[e].x) * idz;
        const double b = fabs((double)u[e].y) * idx + fabs((double)v[e].y) * idy + fabs((double)w[e].y) * idz;
        cm = fmax(cm, fmax(a, b));
      }
    }
    for (int o = 32; o > 0; o >>= 1) cm = fmax(cm, __shfl_xor(cm, o));
    if constexpr (XL<M>::NTH > 64) {  // the block's wavefronts
      __shared__ double red[XL<M>::NTH / 64];
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cm;
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int i = 1; i < XL<M>::NTH / 64; i++) cm = fmax(cm, red[i]);
        cfl_part[blockIdx.x] = cm;
      }
    } else {
      if (threadIdx.x == 0) cfl_part[blockIdx.x] = cm;
    }
  }
  if constexpr (MODE == 2) return;
  // the R2C post-processing's halves, (Z(k) +- conj Z(M - k)) / 2, folded
  // into the factor (halving is exact above the denormal range): four
  // multiplications a mode fewer, the same values up to how the compiler
  // contracts the multiply-adds
  const R factor = (R)(0.5 * D.factor);
  constexpr int NP = (MODE == 0 ? 6 : (MODE == 3 ? 5 : 3));
  // unrolled: with p known at compile time no register array is selected at
  // run time (a run-time select of arrays would put them in scratch memory)
#pragma unroll
  for (int p = 0; p < NP; p++) {
    Cx<R> z[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      if constexpr (MODE == 1) {
        const Cx<R> a = (p == 0 ? u[e] : (p == 1 ? v[e] : w[e]));
        z[e] = Cx<R>{a.x * factor, a.y * factor};
      } else if constexpr (MODE == 3) {
        z[e] = combined_product<R>(p, u[e], v[e], w[e], factor);
      } else {
        const Cx<R> a = (p == 0 || p == 3 || p == 5) ? u[e] : (p == 1 || p == 4 ? v[e] : w[e]);
        const Cx<R> b = (p == 0) ? u[e] : ((p == 1 || p == 3) ? v[e] : w[e]);
        z[e] = Cx<R>{a.x * b.x * factor, a.y * b.y * factor};
      }
    }
    xl_fft<R, M, -1, 0, 1>(z, L, t, twl);
    // R2C post-processing needs Z(k) and Z(M - k)
#pragma unroll
    for (int e = 0; e < E; e++) L[padx(t + T * e)] = z[e];
    xl_sync<M>();
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int k = t + T * e;
      if (k <= nx) {
        const Cx<R> Z = z[e], Zc = conj(L[padx((M - k) & (M - 1))]);
        const Cx<R> ep = twhl[k];  // exp(-i pi k / M)
        z[e] = (Z + Zc) - mul_i(ep * (Z - Zc));
      }
    }
    // the kept modes go to the line's own LDS, [ix], once every lane of the
    // line has read its Z(M - k) (a separate staging array measured 6 %
    // slower in FP32: 33.5 against 22.5 KB of LDS per block)
    xl_sync<M>();
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int k = t + T * e;
      if (k <= nx) L[k] = z[e];
    }
    __syncthreads();
    // the block's lines to the tiles: consecutive threads take the lines of
    // one tile's HT modes, so each tile's run of LPB z points is one store
    const int nt = (nx + HT) / HT;
    Cx<R> *const hp = Hp + (size_t)p * nrows * plane;
    for (int i = threadIdx.x; i < nt * LPB * HT; i += NTH) {
      const int tile = i / (LPB * HT), r = i - tile * (LPB * HT), gl = r / HT;
      const int ix = tile * HT + (r - gl * HT);
      const long long b = lbase[gl];
      if (ix <= nx && b >= 0) hp[b + (long long)tile * NZD * HT + (r - gl * HT)] = lds[gl * PL + ix];
    }
    // every tile read before the next product's transform writes the lines
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launchers

template <typename R, int NZD>
static int run_zift(const Dims &D, const double2 *const V[3], void *Hv, int row0, int nrows, const Tables<R> &tb,
                    int compact, hipStream_t s) {
  typedef ZCfg<R, NZD, false> C;
  if (compact) {
    dim3 grid((D.NXm + C::G / 2 - 1) / (C::G / 2) + (D.NXm + C::G - 1) / C::G, nrows, 1);
    hipLaunchKernelGGL((k_zift<R, NZD, true>), grid, dim3(C::NT), 0, s, V[0], V[1], V[2], (Cx<R> *)Hv, D, row0,
                       nrows, tb.twz);
  } else {
    dim3 grid((D.NXm + C::G - 1) / C::G, nrows, 3);
    hipLaunchKernelGGL((k_zift<R, NZD, false>), grid, dim3(C::NT), 0, s, V[0], V[1], V[2], (Cx<R> *)Hv, D, row0,
                       nrows, tb.twz);
  }
  return hipGetLastError();
}

template <typename R, int NZD>
static int run_zfft(const Dims &D, int np, const void *Hp, void *P0, size_t pstride, int row0, int nrows,
                    const Tables<R> &tb, hipStream_t s) {
  typedef ZCfg<R, NZD, true> C;
  dim3 grid((D.NXm + C::G - 1) / C::G, nrows, np);
  hipLaunchKernelGGL((k_zfft<R, NZD>), grid, dim3(C::NT), 0, s, (const Cx<R> *)Hp, (Cx<R> *)P0, pstride, D, row0,
                     nrows, tb.twz);
  return hipGetLastError();
}

template <typename R, int M>
static int run_xpass(const Dims &D, int mode, const void *Hv, void *Hp, int row0, int nrows, const Tables<R> &tb,
                     const double *dyl, double *cfl_part, int want_cfl, int *n_part, hipStream_t s) {
  if constexpr ((M & (M - 1)) == 0 && M >= 64) {
    typedef XL<M> C;
    if (D.nx + 1 > C::NXS) return hipErrorInvalidValue;
    const int lines = D.nzd * nrows;
    dim3 grid((lines + C::LPB - 1) / C::LPB);
    if (n_part) *n_part = grid.x;
#define LX(MD, WC)                                                                                                 \
  hipLaunchKernelGGL((k_xline<R, M, MD>), grid, dim3(C::NTH), 0, s, (const Cx<R> *)Hv, (Cx<R> *)Hp, D, nrows, tb.twx, \
                     tb.twh, dyl, cfl_part, WC, row0)
    if (mode == 0) LX(0, want_cfl);
    else if (mode == 1) LX(1, want_cfl);
    else if (mode == 3) LX(3, want_cfl);
    else LX(2, 1);
#undef LX
    return hipGetLastError();
  } else {
    typedef XCfg<R, M> C;
    if (D.nzd % C::G) return hipErrorInvalidValue;
    dim3 grid(D.nzd / C::G, nrows);
    if (n_part) *n_part = grid.x * grid.y;
    if (mode == 0)
      hipLaunchKernelGGL((k_xpass<R, M, 0>), grid, dim3(C::NT), 0, s, (const Cx<R> *)Hv, (Cx<R> *)Hp, D, row0, nrows,
                         tb.twx, tb.twh, dyl, cfl_part, want_cfl);
    else if (mode == 1)
      hipLaunchKernelGGL((k_xpass<R, M, 1>), grid, dim3(C::NT), 0, s, (const Cx<R> *)Hv, (Cx<R> *)Hp, D, row0, nrows,
                         tb.twx, tb.twh, dyl, cfl_part, want_cfl);
    else if (mode == 3)
      hipLaunchKernelGGL((k_xpass<R, M, 3>), grid, dim3(C::NT), 0, s, (const Cx<R> *)Hv, (Cx<R> *)Hp, D, row0, nrows,
                         tb.twx, tb.twh, dyl, cfl_part, want_cfl);
    else
      hipLaunchKernelGGL((k_xpass<R, M, 2>), grid, dim3(C::NT), 0, s, (const Cx<R> *)Hv, (Cx<R> *)Hp, D, row0, nrows,
                         tb.twx, tb.twh, dyl, cfl_part, 1);
    return hipGetLastError();
  }
}

#define HST_FOR_SIZES(X) X(6) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536)

template <typename R>
static int zift_dispatch(const Dims &D, const double2 *const V[3], void *Hv, int row0, int nrows, const Tables<R> &tb,
                         int compact, hipStream_t s) {
  switch (D.nzd) {
#define CASE(n) case n: return run_zift<R, n>(D, V, Hv, row0, nrows, tb, compact, s);
    HST_FOR_SIZES(CASE)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <typename R>
static int zfft_dispatch(const Dims &D, int np, const void *Hp, void *P0, size_t pstride, int row0, int nrows,
                         const Tables<R> &tb, hipStream_t s) {
  switch (D.nzd) {
#define CASE(n) case n: return run_zfft<R, n>(D, np, Hp, P0, pstride, row0, nrows, tb, s);
    HST_FOR_SIZES(CASE)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <typename R>
static int xpass_dispatch(const Dims &D, int mode, const void *Hv, void *Hp, int row0, int nrows, const Tables<R> &tb,
                          const double *dyl, double *cfl_part, int want_cfl, int *n_part, hipStream_t s) {
  switch (D.nxd) {
#define CASE(n) case n: return run_xpass<R, n>(D, mode, Hv, Hp, row0, nrows, tb, dyl, cfl_part, want_cfl, n_part, s);
    HST_FOR_SIZES(CASE)
#undef CASE
  }
  return hipErrorInvalidValue;
}

int launch_zift(const Dims &D, int mixed, const double2 *const V[3], void *Hv, int row0, int nrows,
                const void *tables, hipStream_t s, int compact) {
  const TableSet *t = (const TableSet *)tables;
  return mixed ? zift_dispatch<float>(D, V, Hv, row0, nrows, t->f, compact, s)
               : zift_dispatch<double>(D, V, Hv, row0, nrows, t->d, compact, s);
}

int launch_xpass(const Dims &D, int mixed, int mode, const void *Hv, void *Hp, int row0, int nrows,
                 const void *tables, const double *dyl, double *cfl_part, int want_cfl, int *n_part, hipStream_t s) {
  const TableSet *t = (const TableSet *)tables;
  return mixed ? xpass_dispatch<float>(D, mode, Hv, Hp, row0, nrows, t->f, dyl, cfl_part, want_cfl, n_part, s)
               : xpass_dispatch<double>(D, mode, Hv, Hp, row0, nrows, t->d, dyl, cfl_part, want_cfl, n_part, s);
}

int launch_zfft(const Dims &D, int mixed, int np, const void *Hp, void *P0, size_t pstride, int row0, int nrows,
                const void *tables, hipStream_t s) {
  const TableSet *t = (const TableSet *)tables;
  return mixed ? zfft_dispatch<float>(D, np, Hp, P0, pstride, row0, nrows, t->f, s)
               : zfft_dispatch<double>(D, np, Hp, P0, pstride, row0, nrows, t->d, s);
}

// Algorithmic bytes (each array read or written once; interior of the chunk).
double bytes_zift(const Dims &D, int mixed, int nrows) {
  const double cs = mixed ? 8.0 : 16.0;
  return 3.0 * nrows * ((double)D.ncol * 16.0 + (double)D.NXm * D.nzd * cs);
}
double bytes_xpass(const Dims &D, int mixed, int mode, int nrows) {
  const double cs = mixed ? 8.0 : 16.0, h = (double)D.NXm * D.nzd * nrows * cs;
  return mode == 0 ? 9.0 * h : (mode == 3 ? 8.0 * h : (mode == 1 ? 6.0 * h : 3.0 * h));
}
double bytes_zfft(const Dims &D, int mixed, int np, int nrows) {
  const double cs = mixed ? 8.0 : 16.0;
  return np * (double)nrows * ((double)D.NXm * D.nzd * cs + (double)D.ncol * cs);
}

}  // namespace hst
