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


def kelvin(make_solver, deck):
    """test_kelvin (tests/test_kelvin.f90:49-136): one Kelvin mode (ix, iz, ky) =
    (1, 1, 1) of the linearised equations against its closed form."""
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
