"""ctypes binding of the CPU oracle (oracle/libhst_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the reference the GPU library is tested
against.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from homogenenousshearturbulence_amd.deck import Deck, first_timestep, hst_grid_info, new_timestep

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhst_oracle.so")

RK_RAI = ((120.0 / 32.0, 2.0, 0.0), (120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0), (120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0))


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, dp, ip, d, i = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.hsto_create.restype = vp
        L.hsto_create.argtypes = [vp]
        L.hsto_free.argtypes = [vp]
        L.hsto_info.argtypes = [vp, vp, vp, vp, vp]
        L.hsto_field.restype = vp
        L.hsto_field.argtypes = [vp, i]
        L.hsto_products.restype = vp
        L.hsto_products.argtypes = [vp]
        L.hsto_uniform_from_key.restype = d
        L.hsto_uniform_from_key.argtypes = [i, i, i, i, i]
        L.hsto_initial_field.argtypes = [vp]
        L.hsto_fill_ghosts.argtypes = [vp, i, d, d]
        L.hsto_fill_all_ghosts.argtypes = [vp, d]
        L.hsto_shear_shifts.argtypes = [vp, d, dp, dp]
        L.hsto_gamma_y.restype = d
        L.hsto_gamma_y.argtypes = [vp, d]
        L.hsto_s2_of.restype = d
        L.hsto_s2_of.argtypes = [vp, d]
        L.hsto_line_solve.argtypes = [vp, i, d, i, i, d, d]
        L.hsto_nonlinear.restype = d
        L.hsto_nonlinear.argtypes = [vp, i, i]
        L.hsto_buildrhs_prepare.argtypes = [vp, dp, d, d]
        L.hsto_buildrhs.argtypes = [vp, dp, i]
        L.hsto_shear_shift.argtypes = [vp, d, d]
        L.hsto_linsolve.argtypes = [vp, d, d]
        L.hsto_timestep.argtypes = [vp, d, dp, dp]
        L.hsto_outstats_at.argtypes = [vp, d, dp]
        L.hsto_set_threads.argtypes = [i]
        L.hsto_max_threads.restype = i
        L.hsto_pressure.argtypes = [vp, d, i]
        L.hsto_stokes_apply.argtypes = [vp, d]
        L.hsto_project_divergence.argtypes = [vp, d]
        L.hsto_damping_mask.argtypes = [vp]
        L.hsto_outstats_extra.argtypes = [vp, vp]
        L.hsto_gdamp.restype = vp
        L.hsto_gdamp.argtypes = [vp]
        L.hsto_column_energy.argtypes = [vp, vp]
        # learned closure
        L.hsto_filter_weights.restype = i
        L.hsto_filter_weights.argtypes = [i, i, vp, vp]
        L.hsto_filter_field.argtypes = [vp, vp, vp, vp]
        L.hsto_to_physical.argtypes = [vp, vp, i, vp]
        L.hsto_from_physical.argtypes = [vp, vp, vp]
        L.hsto_fill_ghosts_of.argtypes = [vp, vp, d, d]
        L.hsto_mlp_nparam.restype = i
        L.hsto_bf16_round.restype = ctypes.c_float
        L.hsto_bf16_round.argtypes = [ctypes.c_float]
        L.hsto_mlp_grad.restype = d
        L.hsto_mlp_grad.argtypes = [vp, vp, ctypes.c_int64, vp, vp, i, vp]
        L.hsto_mlp_forward.argtypes = [vp, vp, ctypes.c_int64, vp, i, vp]
        _lib = L
    return _lib


class Oracle:
    """One FP64 CPU state of the reference path (fields in the reference layout)."""

    def __init__(self, deck: Deck):
        self.deck = deck
        self._c = deck.to_c()
        self.h = lib().hsto_create(ctypes.byref(self._c))
        info = hst_grid_info()
        ny = deck.ny
        self.y = np.zeros(ny + 4)
        self.dyl = np.zeros(ny)
        self.der_f = np.zeros(20 * ny)
        lib().hsto_info(self.h, ctypes.byref(info), self.y.ctypes.data, self.dyl.ctypes.data, self.der_f.ctypes.data)
        self.info = info
        # der(iy, n, j) -> der[iy, n, j+2]
        self.der = self.der_f.reshape((5, 4, ny)).transpose(2, 1, 0)
        self.shape = (deck.nx + 1, 2 * deck.nz + 1, ny + 4)
        self.time = deck.time
        self.cfl = 0.0

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().hsto_free(h)
            self.h = None

    def field(self, fid: int) -> np.ndarray:
        """Live view of field fid (1..8) as complex128 (nx+1, 2nz+1, ny+4)."""
        n = int(np.prod(self.shape))
        p = lib().hsto_field(self.h, fid)
        buf = (ctypes.c_double * (2 * n)).from_address(p)
        return np.frombuffer(buf, dtype=np.complex128).reshape(self.shape)

    def products(self) -> np.ndarray:
        n = int(np.prod(self.shape))
        p = lib().hsto_products(self.h)
        buf = (ctypes.c_double * (12 * n)).from_address(p)
        return np.frombuffer(buf, dtype=np.complex128).reshape((6,) + self.shape)

    @property
    def V(self) -> np.ndarray:
        return np.stack([self.field(c) for c in (1, 2, 3)])

    def set_V(self, V: np.ndarray) -> None:
        for c in range(3):
            self.field(c + 1)[...] = V[c]

    def initial_field(self) -> None:
        lib().hsto_initial_field(self.h)

    def fill_ghosts(self, fid: int, sx: float, sz: float) -> None:
        lib().hsto_fill_ghosts(self.h, fid, sx, sz)

    def fill_all_ghosts(self, t: float | None = None) -> None:
        lib().hsto_fill_all_ghosts(self.h, self.time if t is None else t)

    def shear_shifts(self, t: float):
        sx, sz = ctypes.c_double(), ctypes.c_double()
        lib().hsto_shear_shifts(self.h, t, ctypes.byref(sx), ctypes.byref(sz))
        return sx.value, sz.value

    def gamma_y(self, t: float) -> float:
        return lib().hsto_gamma_y(self.h, t)

    def s2_of(self, t: float) -> float:
        return lib().hsto_s2_of(self.h, t)

    def line_solve(self, kind: int, lam: float, src: int, dst: int, sx: float, sz: float) -> None:
        lib().hsto_line_solve(self.h, kind, lam, src, dst, sx, sz)

    def nonlinear(self, mode: int = 0, want_cfl: bool = False) -> float:
        return lib().hsto_nonlinear(self.h, mode, int(want_cfl))

    def timestep(self, deltat: float) -> None:
        t = ctypes.c_double(self.time)
        c = ctypes.c_double(0.0)
        lib().hsto_timestep(self.h, deltat, ctypes.byref(t), ctypes.byref(c))
        self.time = t.value
        self.cfl = c.value

    def pressure(self, dst: int = 5) -> np.ndarray:
        """compute_pressure at the current time into field dst (default HST_RHS2)."""
        lib().hsto_pressure(self.h, self.time, dst)
        return self.field(dst).copy()

    def stokes_apply(self) -> None:
        lib().hsto_stokes_apply(self.h, self.time)

    def project_divergence(self) -> None:
        lib().hsto_project_divergence(self.h, self.time)

    def damping_mask(self) -> None:
        lib().hsto_damping_mask(self.h)

    def outstats_extra(self) -> np.ndarray:
        out = np.zeros(8)
        lib().hsto_outstats_extra(self.h, out.ctypes.data)
        return out

    def gdamp(self) -> np.ndarray:
        """gdamp[ix, iz + nz] as last built."""
        n = (self.deck.nx + 1) * (2 * self.deck.nz + 1)
        buf = (ctypes.c_double * n).from_address(lib().hsto_gdamp(self.h))
        return np.frombuffer(buf, dtype=np.float64).reshape((self.deck.nx + 1, 2 * self.deck.nz + 1)).copy()

    def column_energy(self) -> np.ndarray:
        """E[ix, iz + nz] of the current velocity (one outstats sample)."""
        E = np.zeros((self.deck.nx + 1, 2 * self.deck.nz + 1))
        lib().hsto_column_energy(self.h, E.ctypes.data)
        return E

    def outstats(self) -> np.ndarray:
        out = np.zeros(10)
        lib().hsto_outstats_at(self.h, self.time, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out


def remap_field(Vf: np.ndarray, sizes: dict, deck: Deck, y: np.ndarray) -> np.ndarray:
    """restart_read's remap of a file field (src/hst_io.f90:155-167): the
    Fourier modes present in both copied, every target row from the file row
    nearest to it on the file's uniform grid (the end point wrapped); ghost
    rows zero.  Vf: the file's field as cplio.read_field returns it."""
    nx_s, nz_s, ny_s = sizes["nx"], sizes["nz"], sizes["ny"]
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    dy_s = deck.ly / ny_s
    nxc, nzc = min(deck.nx, nx_s), min(deck.nz, nz_s)
    for iy in range(deck.ny):
        src = int(np.floor(np.mod(y[iy + 2], deck.ly) / dy_s + 0.5))  # nint of a non-negative value
        if src == ny_s:
            src = 0
        V[:, :nxc + 1, deck.nz - nzc:deck.nz + nzc + 1, iy + 2] = Vf[:, :nxc + 1, nz_s - nzc:nz_s + nzc + 1, src + 2]
    return V


def uniform_from_key(seed, component, iy, iz, ix) -> float:
    return lib().hsto_uniform_from_key(seed, component, iy, iz, ix)


def set_threads(n: int) -> None:
    lib().hsto_set_threads(n)


def max_threads() -> int:
    return lib().hsto_max_threads()


def compute_cfl(o: Oracle) -> float:
    """transform_to_physical + compute_cfl (src/hst_transforms.f90:163-179)
    of the current velocity (the products it leaves are not read)."""
    return o.nonlinear(mode=1, want_cfl=True)


def run_deck(deck: Deck, nsteps: int | None = None) -> Oracle:
    """The driver's start-up and time loop (src/hst.f90:63-125): fresh field
    (no restart file), the layer profile, ghosts at the start time, the first
    step from the CFL number of the initial field (src/hst.f90:84-87), then per
    step timestep() and new_timestep() (:163-169) until t_max or nstep, or
    nsteps steps when given.  o.dts lists the step lengths taken."""
    o = Oracle(deck)
    o.time = deck.time
    o.initial_field()
    o.stokes_apply()  # the layer profile imposed on the fresh field (src/hst.f90:76-77)
    o.fill_all_ghosts()
    dt = first_timestep(deck.deltat)
    o.cfl = compute_cfl(o)
    dt = new_timestep(dt, o.cfl, deck.cflmax, deck.dt_fixed)
    istep = 0
    o.dts = []
    while (o.time < deck.t_max - 0.5 * dt and istep < deck.nstep) if nsteps is None else istep < nsteps:
        o.timestep(dt)
        o.dts.append(dt)
        istep += 1
        dt = new_timestep(dt, o.cfl, deck.cflmax, deck.dt_fixed)
    o.nsteps = istep
    o.deltat = dt
    return o


def _reset_old(self) -> None:
    self.field(6)[...] = 0.0
    self.field(7)[...] = 0.0


Oracle.reset_old = _reset_old


# ---------------------------------------------------------------- learned closure


def filter_weights(ny_fine: int, ny_coarse: int):
    j0 = np.zeros(ny_coarse, dtype=np.int32)
    w = np.zeros((ny_coarse, 8))
    rc = lib().hsto_filter_weights(ny_fine, ny_coarse, j0.ctypes.data, w.ctypes.data)
    if rc:
        raise ValueError(f"filter weights: status {rc}")
    return j0, w


def filter_field(fine: "Oracle", coarse: "Oracle", f: np.ndarray) -> np.ndarray:
    """The coarse filtered field (interior rows; ghosts zero) of a fine field."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    out = np.zeros(coarse.shape, dtype=np.complex128)
    lib().hsto_filter_field(fine.h, coarse.h, f.ctypes.data, out.ctypes.data)
    return out


def to_physical(o: "Oracle", f: np.ndarray, op: int = 0) -> np.ndarray:
    """Physical values [ny][nzd][2 nxd]; op 1: d/dx, 2: d/dz."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    out = np.zeros((o.deck.ny, o.info.nzd, 2 * o.info.nxd))
    lib().hsto_to_physical(o.h, f.ctypes.data, op, out.ctypes.data)
    return out


def from_physical(o: "Oracle", phys: np.ndarray) -> np.ndarray:
    phys = np.ascontiguousarray(phys, dtype=np.float64)
    out = np.zeros(o.shape, dtype=np.complex128)
    lib().hsto_from_physical(o.h, phys.ctypes.data, out.ctypes.data)
    return out


def fill_ghosts_of(o: "Oracle", f: np.ndarray, sx: float, sz: float) -> None:
    assert f.flags.c_contiguous and f.dtype == np.complex128
    lib().hsto_fill_ghosts_of(o.h, f.ctypes.data, sx, sz)


def dy(o: "Oracle", f: np.ndarray, sx: float, sz: float) -> np.ndarray:
    """d/dy of a field with ghost rows filled: the D0 solve of D1 f (KIND_DY)."""
    o.field(8)[...] = f
    o.line_solve(5, 0.0, 8, 7, sx, sz)
    return o.field(7).copy()


def gradients(o: "Oracle", U, sx: float, sz: float):
    """The 9 physical gradient fields du_i/dx_j (i-major) of U (3 fields, ghosts valid)."""
    g = []
    for i in range(3):
        g.append(to_physical(o, U[i], 1))
        g.append(to_physical(o, dy(o, U[i], sx, sz), 0))
        g.append(to_physical(o, U[i], 2))
    return g


def samples(o: "Oracle", U, Pbar, sx: float, sz: float):
    """(x[n][9], y[n][6]) at every coarse grid point, index ((iy nzd + zp) 2 nxd + x)."""
    g = gradients(o, U, sx, sz)
    u = [to_physical(o, U[i]) for i in range(3)]
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2))
    tau = [to_physical(o, Pbar[p]) - u[a] * u[b] for p, (a, b) in enumerate(pairs)]
    x = np.stack([a.ravel() for a in g], axis=1)
    y = np.stack([a.ravel() for a in tau], axis=1)
    return x, y


def mlp_nparam() -> int:
    return lib().hsto_mlp_nparam()


def bf16_round(x: float) -> float:
    return lib().hsto_bf16_round(x)


def mlp_grad(params, norm, x, y, bf16: bool):
    """(loss, gradient) of the mean squared standardised error; bf16 rounds
    where the GPU kernels feed the matrix cores."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    nm = np.ascontiguousarray(norm, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(p.size)
    loss = lib().hsto_mlp_grad(p.ctypes.data, nm.ctypes.data, x.shape[0], x.ctypes.data, y.ctypes.data, int(bf16),
                               g.ctypes.data)
    return loss, g


def mlp_forward(params, norm, x, bf16: bool) -> np.ndarray:
    """Standardised outputs [n][6]."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    nm = np.ascontiguousarray(norm, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros((x.shape[0], 6))
    lib().hsto_mlp_forward(p.ctypes.data, nm.ctypes.data, x.shape[0], x.ctypes.data, int(bf16), out.ctypes.data)
    return out


def closure_stress(o: "Oracle", U, params, norm, sx: float, sz: float, bf16: bool = True):
    """The modelled stress spectra (6 fields, ghosts at sx, sz) of U: gradients,
    the network at every grid point, the forward transform.  `o` is a helper
    state of the coarse deck (its fields 7 and 8 are used as work space)."""
    g = gradients(o, U, sx, sz)
    x = np.stack([a.ravel() for a in g], axis=1).astype(np.float32)
    out = mlp_forward(params, norm, x, bf16)
    sy = np.float32(norm[24:30]).astype(np.float32)
    my = np.float32(norm[18:24]).astype(np.float32)
    tau = (out.astype(np.float32) * sy + my).astype(np.float64)
    shape = g[0].shape
    T = []
    for p in range(6):
        f = from_physical(o, tau[:, p].reshape(shape))
        fill_ghosts_of(o, f, sx, sz)
        T.append(f)
    return np.stack(T)


def step_with_closure(o: "Oracle", helper: "Oracle", deltat: float, params, norm) -> None:
    """hsto_timestep with the modelled stresses added to the products after
    every product pass (the closure's hook in the substep)."""
    L = lib()
    t = o.time
    for ode in RK_RAI:
        o.nonlinear(0)
        sx, sz = o.shear_shifts(t)
        o.products()[...] += closure_stress(helper, o.V, params, norm, sx, sz)
        c = (ctypes.c_double * 3)(*ode)
        L.hsto_buildrhs_prepare(o.h, c, deltat, t)
        L.hsto_buildrhs(o.h, c, 1)
        L.hsto_buildrhs(o.h, c, 2)
        dts = 2.0 / ode[0] * deltat
        L.hsto_shear_shift(o.h, dts, t)
        t = t + dts
        L.hsto_linsolve(o.h, ode[0] / deltat, t)
    o.time = t
