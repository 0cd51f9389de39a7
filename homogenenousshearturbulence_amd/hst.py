"""Python binding of libhstgpu.so (the C ABI of include/hst_gpu.h).

The calls keep the names and argument meaning of the reference's Fortran
modules (hst_equations, hst_transforms, hst_linsolve, hst_derivatives):
timestep, transform/products (nonlinear), buildrhs_prepare, buildrhs,
shear_shift, line_solve, linsolve, fill_ghosts, outstats.  Errors raise
HstError with the library's message (the reference stops on every error).

There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .deck import Deck, first_timestep, hst_deck, hst_grid_info, new_timestep

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhstgpu.so")

V1, V2, V3, RHS1, RHS2, OLD1, OLD2, SCRATCH = 1, 2, 3, 4, 5, 6, 7, 8
KIND_D2V, KIND_ETA, KIND_POISSON, KIND_D0, KIND_DY = 1, 2, 3, 4, 5
FP64, MIXED = 0, 1

# RK_rai (src/hst_params.f90:108-111): one row per substep
RK_RAI = ((120.0 / 32.0, 2.0, 0.0), (120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0), (120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0))

EXPORTS = (
    "hst_grid_setup", "hst_initial_field", "hst_initial_energy", "hst_gpu_init", "hst_gpu_free", "hst_gpu_sync", "hst_gpu_last_error",
    "hst_gpu_grid", "hst_gpu_upload_field", "hst_gpu_download_field", "hst_gpu_zero_field", "hst_gpu_fill_ghosts",
    "hst_gpu_nonlinear", "hst_gpu_compute_cfl", "hst_gpu_product_to_field", "hst_gpu_roundtrip",
    "hst_gpu_buildrhs_prepare", "hst_gpu_buildrhs", "hst_gpu_shear_shift", "hst_gpu_line_solve", "hst_gpu_linsolve",
    "hst_gpu_timestep", "hst_gpu_outstats", "hst_gpu_set_chunk_rows", "hst_gpu_memory", "hst_gpu_profile",
    "hst_gpu_profile_count", "hst_gpu_profile_entry", "hst_gpu_pressure", "hst_gpu_spectrum",
    "hst_gpu_stokes_apply", "hst_gpu_damping_mask", "hst_gpu_outstats_extra", "hst_gpu_project_divergence",
    "hst_gpu_bandwidth_probe", "hst_gpu_line_solve_noalloc", "hst_gpu_fft_z", "hst_gpu_fft_x",
    "hst_gpu_products", "hst_gpu_set_products",
)


class HstError(RuntimeError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HstError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    vp, d, i, dp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    sig = {
        "hst_grid_setup": [vp, vp, vp, vp, vp],
        "hst_initial_field": [vp, vp],
        "hst_initial_energy": [vp, dp],
        "hst_gpu_init": [vp, i],
        "hst_gpu_free": [],
        "hst_gpu_sync": [],
        "hst_gpu_last_error": [ctypes.c_char_p, i],
        "hst_gpu_grid": [vp, vp, vp, vp],
        "hst_gpu_upload_field": [i, vp],
        "hst_gpu_download_field": [i, vp],
        "hst_gpu_zero_field": [i],
        "hst_gpu_fill_ghosts": [i, d, d],
        "hst_gpu_nonlinear": [dp],
        "hst_gpu_compute_cfl": [dp],
        "hst_gpu_product_to_field": [i, i],
        "hst_gpu_roundtrip": [],
        "hst_gpu_buildrhs_prepare": [dp, d, d],
        "hst_gpu_buildrhs": [dp, i],
        "hst_gpu_shear_shift": [d, d],
        "hst_gpu_line_solve": [i, d, i, i, d, d],
        "hst_gpu_linsolve": [d, d],
        "hst_gpu_timestep": [d, dp, dp],
        "hst_gpu_outstats": [d, dp],
        "hst_gpu_set_chunk_rows": [i],
        "hst_gpu_memory": [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)],
        "hst_gpu_profile": [i],
        "hst_gpu_profile_count": [],
        "hst_gpu_profile_entry": [i, ctypes.c_char_p, i, ctypes.POINTER(ctypes.c_int64), dp, dp],
        "hst_gpu_pressure": [d, i],
        "hst_gpu_spectrum": [vp, ctypes.POINTER(ctypes.c_int32)],
        "hst_gpu_stokes_apply": [d],
        "hst_gpu_project_divergence": [d],
        "hst_gpu_damping_mask": [],
        "hst_gpu_outstats_extra": [dp],
        "hst_gpu_bandwidth_probe": [d, i, dp],
        "hst_gpu_line_solve_noalloc": [i, d, i, i, vp, vp, d, d],
        "hst_gpu_fft_z": [vp, ctypes.c_int64, i],
        "hst_gpu_fft_x": [vp, ctypes.c_int64, i],
        "hst_gpu_products": [i, ctypes.c_int64, vp, vp],
        "hst_gpu_set_products": [i, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().hst_gpu_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise HstError(f"{what}: {last_error()} (status {rc})")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def grid_setup(deck: Deck):
    """Derived sizes, y, dyl, der(iy, n, j) (host-only, no GPU)."""
    c = deck.to_c()
    info = hst_grid_info()
    y = np.zeros(deck.ny + 4)
    dyl = np.zeros(deck.ny)
    der = np.zeros(20 * deck.ny)
    _check(lib().hst_grid_setup(ctypes.byref(c), ctypes.byref(info), y.ctypes.data, dyl.ctypes.data,
                                der.ctypes.data), "hst_grid_setup")
    return info, y, dyl, der.reshape((5, 4, deck.ny)).transpose(2, 1, 0)


def initial_field(deck: Deck) -> np.ndarray:
    """generate_initial_field on the host: complex128 (3, nx+1, 2nz+1, ny+4)."""
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    c = deck.to_c()
    _check(lib().hst_initial_field(ctypes.byref(c), V.ctypes.data), "hst_initial_field")
    return V


def initial_energy(deck: Deck) -> float:
    """<u_i u_i> of the deck's seeded random field (uniform y grid)."""
    c = deck.to_c()
    q = ctypes.c_double(0.0)
    _check(lib().hst_initial_energy(ctypes.byref(c), ctypes.byref(q)), "hst_initial_energy")
    return q.value


class Solver:
    """The device state of one deck: V, rhs, oldrhs and the transform and
    solver buffers on the GPU, and the clock on the host (FP64)."""

    def __init__(self, deck: Deck, precision: int = FP64, chunk_rows: int = 0):
        self.deck = deck
        self.precision = precision
        self._c = deck.to_c()
        _check(lib().hst_gpu_set_chunk_rows(chunk_rows), "hst_gpu_set_chunk_rows")
        _check(lib().hst_gpu_init(ctypes.byref(self._c), precision), "hst_gpu_init")
        info = hst_grid_info()
        self.y = np.zeros(deck.ny + 4)
        self.dyl = np.zeros(deck.ny)
        der = np.zeros(20 * deck.ny)
        _check(lib().hst_gpu_grid(ctypes.byref(info), self.y.ctypes.data, self.dyl.ctypes.data, der.ctypes.data),
               "hst_gpu_grid")
        self.info = info
        self.der = der.reshape((5, 4, deck.ny)).transpose(2, 1, 0)
        self.shape = (deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4)
        self.time = deck.time
        self.cfl = 0.0

    def close(self) -> None:
        lib().hst_gpu_free()

    # ------------------------------------------------------------ fields
    def upload(self, fid: int, F: np.ndarray) -> None:
        F = np.ascontiguousarray(F, dtype=np.complex128)
        assert F.shape == self.shape, (F.shape, self.shape)
        _check(lib().hst_gpu_upload_field(fid, F.ctypes.data), "hst_gpu_upload_field")

    def download(self, fid: int) -> np.ndarray:
        F = np.empty(self.shape, dtype=np.complex128)
        _check(lib().hst_gpu_download_field(fid, F.ctypes.data), "hst_gpu_download_field")
        return F

    def field(self, fid: int) -> np.ndarray:
        return self.download(fid)

    @property
    def V(self) -> np.ndarray:
        return np.stack([self.download(c) for c in (V1, V2, V3)])

    def set_V(self, V: np.ndarray) -> None:
        for c in range(3):
            self.upload(c + 1, V[c])

    def zero(self, fid: int) -> None:
        _check(lib().hst_gpu_zero_field(fid), "hst_gpu_zero_field")

    def reset_old(self) -> None:
        self.zero(OLD1)
        self.zero(OLD2)

    def initial_field(self) -> None:
        self.set_V(initial_field(self.deck))

    # ------------------------------------------------------------ clock (src/hst_derivatives.f90:85-128)
    def s2_of(self, t: float) -> float:
        d = self.deck
        if (d.s2_amplitude == 0.0 and d.s2_mean == 0.0) or t < d.s2_start:
            return 0.0
        if d.s2_period > 0.0:
            return d.s2_mean + d.s2_amplitude * np.sin(6.283185307179586 / d.s2_period * (t - d.s2_start))
        return d.s2_mean + d.s2_amplitude

    def s2_integral(self, t1: float, t2: float) -> float:
        d = self.deck
        if (d.s2_amplitude == 0.0 and d.s2_mean == 0.0) or t2 <= d.s2_start:
            return 0.0
        a, b = max(t1, d.s2_start), t2
        if d.s2_period > 0.0:
            om = 6.283185307179586 / d.s2_period
            return d.s2_mean * (b - a) + d.s2_amplitude / om * (np.cos(om * (a - d.s2_start)) - np.cos(om * (b - d.s2_start)))
        return (d.s2_mean + d.s2_amplitude) * (b - a)

    def gamma_y(self, t: float) -> float:
        return self.s2_integral(self.deck.s2_start, t)

    def shear_shifts(self, t: float):
        d = self.deck
        return float(np.fmod(d.S * t * d.ly, d.lx)) % d.lx, float(np.fmod(self.gamma_y(t) * d.ly, d.lz)) % d.lz

    # ------------------------------------------------------------ the substep's pieces
    def fill_ghosts(self, fid: int, sx: float, sz: float) -> None:
        _check(lib().hst_gpu_fill_ghosts(fid, sx, sz), "hst_gpu_fill_ghosts")

    def fill_all_ghosts(self, t: float | None = None) -> None:
        sx, sz = self.shear_shifts(self.time if t is None else t)
        for c in (V1, V2, V3):
            self.fill_ghosts(c, sx, sz)

    def nonlinear(self, want_cfl: bool = False) -> float:
        c = ctypes.c_double(0.0)
        _check(lib().hst_gpu_nonlinear(ctypes.byref(c) if want_cfl else None), "hst_gpu_nonlinear")
        return c.value

    def compute_cfl(self) -> float:
        c = ctypes.c_double(0.0)
        _check(lib().hst_gpu_compute_cfl(ctypes.byref(c)), "hst_gpu_compute_cfl")
        return c.value

    def roundtrip(self) -> None:
        _check(lib().hst_gpu_roundtrip(), "hst_gpu_roundtrip")

    def product(self, p: int) -> np.ndarray:
        """Product p (1..6: uu vv ww uv vw uw) of the last nonlinear pass."""
        _check(lib().hst_gpu_product_to_field(p, SCRATCH), "hst_gpu_product_to_field")
        return self.download(SCRATCH)

    def buildrhs_prepare(self, ode, deltat: float, time: float) -> None:
        o = np.asarray(ode, dtype=np.float64)
        _check(lib().hst_gpu_buildrhs_prepare(_dp(o), deltat, time), "hst_gpu_buildrhs_prepare")

    def buildrhs(self, ode, g: int) -> None:
        o = np.asarray(ode, dtype=np.float64)
        _check(lib().hst_gpu_buildrhs(_dp(o), g), "hst_gpu_buildrhs")

    def shear_shift(self, dt_sub: float, time: float) -> None:
        _check(lib().hst_gpu_shear_shift(dt_sub, time), "hst_gpu_shear_shift")

    def line_solve(self, kind: int, lam: float, src: int, dst: int, sx: float, sz: float) -> None:
        _check(lib().hst_gpu_line_solve(kind, lam, src, dst, sx, sz), "hst_gpu_line_solve")

    def line_solve_noalloc(self, kind: int, lam: float, col0: int, ncols: int, src_ptr: int, dst_ptr: int,
                           sx: float, sz: float) -> None:
        """line_solve of the columns col0 .. col0+ncols-1 between caller-owned
        device arrays complex128 [ny+4][ncols] (device addresses), with the
        library's fixed workspace: no device allocation.  Enqueued on the
        library's stream; sync() waits for it."""
        _check(lib().hst_gpu_line_solve_noalloc(kind, lam, col0, ncols, src_ptr, dst_ptr, sx, sz),
               "hst_gpu_line_solve_noalloc")

    def linsolve(self, lam: float, time: float) -> None:
        _check(lib().hst_gpu_linsolve(lam, time), "hst_gpu_linsolve")

    def timestep(self, deltat: float, want_cfl: bool = False) -> None:
        t = ctypes.c_double(self.time)
        c = ctypes.c_double(0.0)
        _check(lib().hst_gpu_timestep(deltat, ctypes.byref(t), ctypes.byref(c) if want_cfl else None),
               "hst_gpu_timestep")
        self.time = t.value
        if want_cfl:
            self.cfl = c.value

    def outstats(self) -> np.ndarray:
        out = np.zeros(10)
        _check(lib().hst_gpu_outstats(self.time, _dp(out)), "hst_gpu_outstats")
        return out

    def pressure(self, dst: int = 5) -> np.ndarray:
        """The fluctuation pressure of the current velocity at self.time
        (compute_pressure, src/hst_pressure.f90:35-112), solved into field dst
        (default HST_RHS2) and returned."""
        _check(lib().hst_gpu_pressure(self.time, dst), "hst_gpu_pressure")
        return self.download(dst)

    def spectrum(self):
        """(E[ix, iz + nz], n): half the column energies averaged over the
        outstats calls since the previous call (spectrum_write,
        src/hst_stats.f90:110-134); restarts the average."""
        E = np.zeros((self.deck.nx + 1, 2 * self.deck.nz + 1))
        n = ctypes.c_int32(0)
        _check(lib().hst_gpu_spectrum(E.ctypes.data, ctypes.byref(n)), "hst_gpu_spectrum")
        return E, n.value

    def stokes_apply(self) -> None:
        """stokes_apply at self.time (src/hst_stokes.f90:103-118), ghosts included."""
        _check(lib().hst_gpu_stokes_apply(self.time), "hst_gpu_stokes_apply")

    def project_divergence(self) -> None:
        """project_restart at self.time (src/hst_io.f90:213-250), ghosts filled before and after."""
        _check(lib().hst_gpu_project_divergence(self.time), "hst_gpu_project_divergence")

    def damping_mask(self) -> None:
        _check(lib().hst_gpu_damping_mask(), "hst_gpu_damping_mask")

    def outstats_extra(self) -> np.ndarray:
        """The Stokes-layer region line and the damping line of the last
        outstats: energy out/in, dissipation out/in, direction (deg), damped
        columns, half their energy, half their lost power."""
        out = np.zeros(8)
        _check(lib().hst_gpu_outstats_extra(_dp(out)), "hst_gpu_outstats_extra")
        return out

    def sync(self) -> None:
        _check(lib().hst_gpu_sync(), "hst_gpu_sync")

    def memory(self):
        b, ch = ctypes.c_int64(0), ctypes.c_int(0)
        lib().hst_gpu_memory(ctypes.byref(b), ctypes.byref(ch))
        return b.value, ch.value

    # ------------------------------------------------------------ timing
    def profile(self, enable: bool) -> None:
        _check(lib().hst_gpu_profile(1 if enable else 0), "hst_gpu_profile")

    def profile_table(self):
        rows = []
        n = lib().hst_gpu_profile_count()
        name = ctypes.create_string_buffer(64)
        for k in range(n):
            la, ms, by = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0)
            _check(lib().hst_gpu_profile_entry(k, name, 64, ctypes.byref(la), ctypes.byref(ms), ctypes.byref(by)),
                   "hst_gpu_profile_entry")
            rows.append({"kernel": name.value.decode(), "launches": la.value, "ms": ms.value, "bytes": by.value})
        return rows


def bandwidth_probe(gib: float = 4.0, reps: int = 5) -> dict:
    """The device's practical HBM bandwidth (hst_gpu_bandwidth_probe): best
    read-only, write-only and copy GB/s of plain streaming kernels."""
    out = np.zeros(6)
    _check(lib().hst_gpu_bandwidth_probe(gib, reps, _dp(out)), "hst_gpu_bandwidth_probe")
    return {"read_GBps": round(out[0], 1), "write_GBps": round(out[1], 1), "copy_GBps": round(out[2], 1),
            "grids": [int(out[3]), int(out[4]), int(out[5])], "GiB": gib}


def run_deck(deck: Deck, precision: int = FP64, nsteps: int | None = None, solver: Solver | None = None) -> Solver:
    """The driver's start-up and time loop (src/hst.f90:63-125) on the GPU:
    fresh field, the layer profile, ghosts, the first step from the CFL
    number of the initial field, then per step timestep() and new_timestep()
    (src/hst.f90:163-169) until t_max or nstep (or nsteps steps when given).
    The CFL number is computed only where it sets the step (cflmax > 0).
    s.dts lists the step lengths taken."""
    s = solver or Solver(deck, precision)
    s.time = deck.time
    s.initial_field()
    s.stokes_apply()  # the layer profile imposed on the fresh field (src/hst.f90:76-77)
    s.fill_all_ghosts()
    adaptive = deck.cflmax > 0.0
    dt = first_timestep(deck.deltat)
    if adaptive:
        s.cfl = s.compute_cfl()
    dt = new_timestep(dt, s.cfl, deck.cflmax, deck.dt_fixed)
    istep = 0
    s.dts = []
    while (s.time < deck.t_max - 0.5 * dt and istep < deck.nstep) if nsteps is None else istep < nsteps:
        s.timestep(dt, want_cfl=adaptive)
        s.dts.append(dt)
        istep += 1
        dt = new_timestep(dt, s.cfl, deck.cflmax, deck.dt_fixed)
    s.nsteps = istep
    s.deltat = dt
    return s
