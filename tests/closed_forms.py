"""The reference's closed-form tests, restated over a solver object.

A solver has: deck, y (rows -2..ny+1), time, set_V(V), V (property),
fill_all_ghosts(), reset_old(), timestep(dt), gamma_y(t), s2_of(t).
Both the CPU oracle and the GPU library are driven through these, so the
same known answers pin both.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PI = 3.141592653589793


def kelvin(make_solver, deck, gk=0.0):
    """test_kelvin (tests/test_kelvin.f90:49-136): one Kelvin mode (ix, iz, ky) =
    (1, 1, 1) of the linearised equations against its closed form.  gk: the
    column's damping rate (the closed form gains exp(-gk t), :102-111); with
    the hard mask (deck.damp_hard and gk > 0) the mode must vanish and the
    errors are the absolute max |v| and |eta|."""
    deck = dataclasses.replace(deck, linear=True, time=0.0)
    s = make_solver(deck)
    ix, iz, mky = 1, 1, 1
    v0, eta0 = 1.0 + 0.0j, 0.3 - 0.2j
    kx, kz, ky0 = deck.alfa0 * ix, deck.beta0 * iz, 2.0 * PI * mky / deck.ly
    kh2 = kx * kx + kz * kz
    kh = math.sqrt(kh2)
    k02 = kh2 + ky0 * ky0
    u0 = (1j * kx * (1j * ky0 * v0) - 1j * kz * eta0) / kh2
    w0 = (1j * kz * (1j * ky0 * v0) + 1j * kx * eta0) / kh2
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    ph = np.exp(1j * ky0 * s.y)
    V[0, ix, iz + deck.nz] = u0 * ph
    V[1, ix, iz + deck.nz] = v0 * ph
    V[2, ix, iz + deck.nz] = w0 * ph
    s.set_V(V)
    dt = deck.deltat if deck.deltat > 0.0 else 1.0e-3
    n = int(round(deck.t_max / dt))
    for _ in range(n):
        s.timestep(dt)
    V = s.V
    time = s.time
    kyt = ky0 - deck.S * kx * time - kz * s.gamma_y(time)
    k2t = kh2 + kyt * kyt
    k2int = 0.0
    for iq in range(20001):
        tq = time * iq / 20000.0
        c = kh2 + (ky0 - deck.S * kx * tq - kz * s.gamma_y(tq)) ** 2
        if iq in (0, 20000):
            c = 0.5 * c
        k2int += c
    k2int = k2int * time / 20000.0
    visc = math.exp(-deck.ni * k2int)
    vex = v0 * k02 / k2t * visc
    check_eta = deck.s2_period == 0.0 and deck.s2_start <= 0.0
    s2c = s.s2_of(0.0)
    c = deck.S * kx + s2c * kz
    if abs(c) > 0.0:
        etaex = (eta0 + (s2c * 1j * kx - deck.S * 1j * kz) * v0 * k02 / (kh * c)
                 * (math.atan(ky0 / kh) - math.atan(kyt / kh))) * visc
    else:
        etaex = (eta0 + (s2c * 1j * kx - deck.S * 1j * kz) * v0 * time) * visc
    yy = s.y[2:deck.ny + 2]
    pht = np.exp(1j * kyt * yy)
    vnum = V[1, ix, iz + deck.nz, 2:deck.ny + 2]
    etanum = 1j * kz * V[0, ix, iz + deck.nz, 2:deck.ny + 2] - 1j * kx * V[2, ix, iz + deck.nz, 2:deck.ny + 2]
    if deck.damp_hard and gk > 0.0:
        return float(np.abs(vnum).max()), float(np.abs(etanum).max()), n
    vex = vex * math.exp(-gk * time)
    etaex = etaex * math.exp(-gk * time)
    err_v = float(np.abs(vnum - vex * pht).max()) / abs(vex)
    err_eta = float(np.abs(etanum - etaex * pht).max()) / abs(etaex)
    return err_v, (err_eta if check_eta else None), n


def taylor_green(make_solver, deck, n=20):
    """test_taylorgreen (tests/test_taylorgreen.f90:154-242): 2-D Taylor-Green
    vortices in the x-y and z-y planes decay as exp(-2 ni k^2 t) with the
    nonlinear terms on.  Returns the worst relative error."""
    deck = dataclasses.replace(deck, S=0.0, linear=False, time=0.0)
    mx, mz, my = 3, 1, 2
    assert abs(mx * deck.alfa0 - 2 * PI * my / deck.ly) < 1e-12
    dt = deck.deltat if deck.deltat > 0.0 else 1.0e-3
    worst = 0.0
    for orient in (1, 2):
        s = make_solver(deck)
        k = 2.0 * PI
        ky = k
        s.reset_old()
        V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
        sy, cy = np.sin(ky * s.y), np.cos(ky * s.y)
        if orient == 1:
            V[0, mx, deck.nz] = 0.5 * sy
            V[1, mx, deck.nz] = 0.5j * cy
        else:
            V[2, 0, mz + deck.nz] = 0.5 * sy
            V[2, 0, -mz + deck.nz] = 0.5 * sy
            V[1, 0, mz + deck.nz] = 0.5j * cy
            V[1, 0, -mz + deck.nz] = -0.5j * cy
        s.set_V(V)
        s.fill_all_ghosts()
        for _ in range(n):
            s.timestep(dt)
        decay = math.exp(-deck.ni * 2.0 * k * k * s.time)
        ex = np.zeros_like(V)
        yy = s.y
        if orient == 1:
            ex[0, mx, deck.nz] = 0.5 * np.sin(ky * yy) * decay
            ex[1, mx, deck.nz] = 0.5j * np.cos(ky * yy) * decay
        else:
            ex[2, 0, mz + deck.nz] = 0.5 * np.sin(ky * yy) * decay
            ex[2, 0, -mz + deck.nz] = 0.5 * np.sin(ky * yy) * decay
            ex[1, 0, mz + deck.nz] = 0.5j * np.cos(ky * yy) * decay
            ex[1, 0, -mz + deck.nz] = -0.5j * np.cos(ky * yy) * decay
        Vn = s.V
        err = float(np.abs(Vn[..., 2:deck.ny + 2] - ex[..., 2:deck.ny + 2]).max())
        worst = max(worst, err / (0.5 * decay))
    return worst


def forcing(make_solver, deck):
    """test_forcing (tests/test_forcing.f90:321-378): eta forcing of
    u = sin(ky), w = sin(kx) after one inviscid step; v stays zero."""
    deck = dataclasses.replace(deck, S=0.0, linear=False, re=1.0e12, time=0.0)
    mx, my = 3, 2
    k = mx * deck.alfa0
    ky = k
    s = make_solver(deck)
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    V[0, 0, deck.nz] = np.sin(ky * s.y)
    V[2, mx, deck.nz] = -0.5j
    s.set_V(V)
    s.fill_all_ghosts()
    dt = 1.0e-4
    ialfa = 1j * mx * deck.alfa0
    eta0 = 0.0 * V[0, mx, deck.nz, 2] - ialfa * V[2, mx, deck.nz, 2]
    s.timestep(dt)
    Vn = s.V
    yy = s.y[2:deck.ny + 2]
    eta1 = -ialfa * Vn[2, mx, deck.nz, 2:deck.ny + 2]
    expected = eta0 + 1j * k * k * dt / 2.0 * np.sin(ky * yy)
    err = float(np.abs(eta1 - expected).max())
    ref = float(np.abs(expected - eta0).max())
    errv = float(np.abs(Vn[1, mx, deck.nz, 2:deck.ny + 2]).max())
    return err / ref, errv


def _tg_plane_field(deck, y, mx=3, ky=2.0 * PI):
    """u = cos(kx) sin(ky), v = -sin(kx) cos(ky) as stored modes (ix = mx)."""
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    V[0, mx, deck.nz] = 0.5 * np.sin(ky * y)
    V[1, mx, deck.nz] = 0.5j * np.cos(ky * y)
    return V


def pressure_taylor_green(make_solver, deck):
    """test_pressure (tests/test_pressure.f90:32-100): the x-y Taylor-Green
    vortex (S = 0) has p = -(cos 2kx + cos 2ky)/4: p(2 mx, 0) = -1/8 and
    p(0, 0) = -cos(2 ky y)/4 (zero-mean gauge).  Returns (max error of those
    two modes, max magnitude of every other mode), interior rows."""
    deck = dataclasses.replace(deck, S=0.0, linear=False, time=0.0)
    mx, my = 3, 2
    assert abs(mx * deck.alfa0 - 2 * PI * my / deck.ly) < 1e-12
    s = make_solver(deck)
    s.time = 0.0
    ky = mx * deck.alfa0
    s.set_V(_tg_plane_field(deck, s.y, mx, ky))
    s.fill_all_ghosts()
    p = s.pressure()[..., 2:deck.ny + 2]
    yy = s.y[2:deck.ny + 2]
    nz = deck.nz
    err = max(float(np.abs(p[2 * mx, nz] + 0.125).max()),
              float(np.abs(p[0, nz] + 0.25 * np.cos(2.0 * ky * yy)).max()))
    rest = p.copy()
    rest[2 * mx, nz] = 0.0
    rest[0, nz] = 0.0
    return err, float(np.abs(rest).max())


def pressure_shear(make_solver, deck, s2=0.5):
    """The mean-shear source of compute_pressure (src/hst_pressure.f90:1-9,
    the -2 (S i alfa + S2 i beta) D0 v term) on two single modes of v,
    v = A cos(ky y) at (ix, iz) = (mx, 0) and (mx, 1), u = w = 0: no product
    reaches these two modes, so each solves (d_yy - k^2) p = -2 i (S alfa +
    S2 beta) v exactly: p = 2 i (S alfa + S2 beta) A cos(ky y) / (k^2 + ky^2).
    Returns the max error over both modes relative to the larger amplitude."""
    deck = dataclasses.replace(deck, S=1.0, linear=False, time=0.0, s2_mean=s2, s2_amplitude=0.0,
                               s2_period=0.0, s2_start=0.0)
    s = make_solver(deck)
    s.time = 0.0
    mx, my = 2, 1
    ky = 2 * PI * my / deck.ly
    A = (0.3 - 0.4j, 0.25 + 0.1j)
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    for A_, iz in zip(A, (0, 1)):
        V[1, mx, iz + deck.nz] = A_ * np.cos(ky * s.y)
    s.set_V(V)
    s.fill_all_ghosts()
    p = s.pressure()[..., 2:deck.ny + 2]
    yy = s.y[2:deck.ny + 2]
    s2now = s.s2_of(0.0)
    err, big = 0.0, 0.0
    for A_, iz in zip(A, (0, 1)):
        al, be = deck.alfa0 * mx, deck.beta0 * iz
        B = 2j * (deck.S * al + s2now * be) * A_ / (al * al + be * be + ky * ky)
        err = max(err, float(np.abs(p[mx, iz + deck.nz] - B * np.cos(ky * yy)).max()))
        big = max(big, abs(B))
    return err / big



def stokes_layer(make_solver, deck):
    """test_stokes (tests/test_stokes.f90:32-77): no fluctuations; the layer is
    prescribed for two periods, then driven by its body force and viscosity
    alone for the rest of t_max; the (0,0) mode of w must follow the analytic
    layer W(y, t) = A exp(-s) cos(omega (t - t0) - s), s = sqrt((eta/delta)^2 +
    0.01), eta = y - ly/2 (src/hst_stokes.f90:6-8), and everything else must
    stay zero.  Returns (max error / max |W|, max of everything else)."""
    deck = dataclasses.replace(deck, time=0.0)
    s = make_solver(deck)
    s.time = 0.0
    s.set_V(np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128))
    s.stokes_apply()
    s.fill_all_ghosts()
    dt = deck.deltat if deck.deltat > 0.0 else 5.0e-3
    n = int(round(deck.t_max / dt))
    for _ in range(n):
        s.timestep(dt)
    V = s.V[..., 2:deck.ny + 2]
    yy = s.y[2:deck.ny + 2]
    t = s.time
    sv = np.sqrt(((yy - 0.5 * deck.ly) / deck.sl_delta) ** 2 + 0.1 ** 2)
    ramp = 1.0 if t >= deck.sl_start else 0.0
    wex = ramp * deck.sl_amplitude * np.exp(-sv) * np.cos(2 * PI / deck.sl_period * (t - deck.sl_start) - sv)
    nz = deck.nz
    err = float(np.abs(V[2, 0, nz] - wex).max())
    other = max(float(np.abs(V[2, 0, nz].imag).max()), float(np.abs(V[0, 0, nz]).max()),
                float(np.abs(V[1, 0, nz]).max()))
    V = V.copy()
    V[2, 0, nz] = 0.0
    other = max(other, float(np.abs(V).max()))
    return err / float(np.abs(wex).max()), other


def cfl_single_mode(deck, y, au=0.7, aw=0.4, peak_row=None, ghost_amp=50.0):
    """A field whose CFL number is known in closed form (compute_cfl,
    src/hst_transforms.f90:163-179): u, v, w all in the Fourier mode
    (ix, iz) = (1, 0) with real amplitudes, so that each is 2 a cos(alfa0 x)
    and all three peak together at x = 0, a physical point.  u and w have the
    same amplitude on every row; v has amplitude 1 on row peak_row only (an
    interior row, by default the row of smallest dyl), a much larger one on the
    ghost rows (which the CFL number must not see) and 0.25 elsewhere.  The
    CFL number is then 2 au/dx + 2 aw/dz + max over interior rows of
    2 av(iy)/dyl(iy), with dx = lx/(2 nxd), dz = lz/nzd.
    Returns (V, expected)."""
    ny = deck.ny
    nxd, nzd = deck.nxd, deck.nzd
    dyl = np.array([0.5 * (y[iy + 3] - y[iy + 1]) for iy in range(ny)])
    if peak_row is None:
        peak_row = int(np.argmin(dyl))
    av = np.full(ny + 4, 0.25)
    av[peak_row + 2] = 1.0
    av[[0, 1, ny + 2, ny + 3]] = ghost_amp
    V = np.zeros((3, deck.nx + 1, 2 * deck.nz + 1, ny + 4), dtype=np.complex128)
    V[0, 1, deck.nz, :] = au
    V[1, 1, deck.nz, :] = av
    V[2, 1, deck.nz, :] = aw
    dx, dz = deck.lx / (2 * nxd), deck.lz / nzd
    expected = 2 * au / dx + 2 * aw / dz + float(np.max(2 * av[2:ny + 2] / dyl))
    return V, expected
