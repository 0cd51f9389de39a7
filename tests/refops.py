"""numpy helpers of the tests: stencil operators through the ghost rows, the
dense cyclic matrix of a line, closed forms.  Fields are complex128 arrays
(nx+1, 2nz+1, ny+4) in the reference layout (row iy at index iy + 2)."""
from __future__ import annotations

import numpy as np

KIND_D2V, KIND_ETA, KIND_POISSON, KIND_D0, KIND_DY = 1, 2, 3, 4, 5


def k2_table(deck):
    ix = np.arange(deck.nx + 1)[:, None]
    iz = np.arange(-deck.nz, deck.nz + 1)[None, :]
    a = deck.alfa0 * ix
    b = deck.beta0 * iz
    return a * a + b * b


def coef(kind, lam, ni, kk, der, iy, j):
    """coef of src/hst_linsolve.f90:236-252 (kk may be an array)."""
    d0, d1, d2, d4 = der[iy, 0, j + 2], der[iy, 1, j + 2], der[iy, 2, j + 2], der[iy, 3, j + 2]
    if kind == KIND_D2V:
        return lam * (d2 - kk * d0) - ni * (d4 - 2.0 * kk * d2 + kk * kk * d0)
    if kind == KIND_ETA:
        return lam * d0 - ni * (d2 - kk * d0)
    if kind == KIND_POISSON:
        return d2 - kk * d0
    return d0 + 0.0 * kk


def apply_operator(kind, lam, ni, k2, der, F):
    """The operator of a system applied to F through its ghost rows, interior rows."""
    ny = der.shape[0]
    out = np.zeros_like(F)
    for iy in range(ny):
        acc = 0.0
        for j in range(-2, 3):
            c = coef(kind, lam, ni, k2, der, iy, j)
            acc = acc + c * F[:, :, iy + j + 2]
        out[:, :, iy + 2] = acc
    return out


def stencil(n, der, F):
    """sum_j der(iy, n, j) F(iy + j) on the interior rows."""
    ny = der.shape[0]
    out = np.zeros_like(F)
    for iy in range(ny):
        out[:, :, iy + 2] = sum(der[iy, n, j + 2] * F[:, :, iy + j + 2] for j in range(-2, 3))
    return out


def wrap_phase(deck, sx, sz):
    ix = np.arange(deck.nx + 1)[:, None]
    iz = np.arange(-deck.nz, deck.nz + 1)[None, :]
    return np.exp(1j * -(deck.alfa0 * ix * sx + deck.beta0 * iz * sz))


def fill_ghosts(deck, F, sx, sz):
    """wrap_rows (src/hst_mpi.f90:340-355) in numpy."""
    ny = deck.ny
    ph = wrap_phase(deck, sx, sz)
    F[:, :, ny + 2] = F[:, :, 2] * ph
    F[:, :, ny + 3] = F[:, :, 3] * ph
    F[:, :, 1] = F[:, :, ny + 1] * np.conj(ph)
    F[:, :, 0] = F[:, :, ny] * np.conj(ph)


def dense_line_matrix(kind, lam, ni, kk, der, ph):
    """The cyclic pentadiagonal matrix of one line (src/hst_linsolve.f90:5-10):
    a(i, i+j mod ny) = coef * (ph if i+j >= ny, conj(ph) if i+j < 0)."""
    ny = der.shape[0]
    A = np.zeros((ny, ny), dtype=np.complex128)
    if kind == KIND_DY:
        kind = KIND_D0
    for i in range(ny):
        for j in range(-2, 3):
            c = coef(kind, lam, ni, kk, der, i, j)
            col = i + j
            w = 1.0
            if col >= ny:
                w = ph
            elif col < 0:
                w = np.conj(ph)
            A[i, col % ny] += c * w
    return A


def random_field(deck, seed_re=3, seed_im=4, uniform=None):
    """The random field of test_linsolve (tests/test_linsolve.f90:97-107), (0,0) zero."""
    F = np.zeros((deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4), dtype=np.complex128)
    for ix in range(deck.nx + 1):
        for iz in range(-deck.nz, deck.nz + 1):
            if ix == 0 and iz == 0:
                continue
            for iy in range(deck.ny):
                F[ix, iz + deck.nz, iy + 2] = complex(uniform(seed_re, 1, iy, iz, ix) - 0.5,
                                                      uniform(seed_im, 1, iy, iz, ix) - 0.5)
    return F


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def physical_velocity(deck, V, nxd, nzd):
    """u, v, w at the 2nxd x nzd points of every row (-2..ny+1) by numpy's own
    FFTs, independent of the oracle's and the library's transforms: the z
    inverse (+i, unnormalised) with mode iz at z index iz mod nzd, then the x
    complex-to-real transform of length 2nxd (unnormalised), which, as every
    c2r does, takes only the real part of the kx = 0 bin.  (3, ny+4, nzd, 2nxd)."""
    nx, nz, ny = deck.nx, deck.nz, deck.ny
    out = np.empty((3, ny + 4, nzd, 2 * nxd))
    for c in range(3):
        Z = np.zeros((nx + 1, nzd, ny + 4), dtype=np.complex128)
        for iz in range(-nz, nz + 1):
            Z[:, iz % nzd, :] = V[c][:, iz + nz, :]
        Z = np.fft.ifft(Z, axis=1) * nzd
        X = np.zeros((nxd + 1, nzd, ny + 4), dtype=np.complex128)
        X[:nx + 1] = Z
        phys = np.fft.irfft(X, n=2 * nxd, axis=0) * (2 * nxd)
        out[c] = phys.transpose(2, 1, 0)
    return out


def cfl_bruteforce(deck, V, y, nxd, nzd):
    """max over the interior rows 0..ny-1 and every physical point of
    |u|/dx + |v|/dyl + |w|/dz: dx = lx/(2nxd) and dz = lz/nzd the spacings of
    the physical points, dyl(iy) = (y(iy+1) - y(iy-1))/2 from the grid's own
    rows (ghosts included: the periodic extension)."""
    ny = deck.ny
    phys = physical_velocity(deck, V, nxd, nzd)
    dx, dz = deck.lx / (2 * nxd), deck.lz / nzd
    best = 0.0
    for iy in range(ny):
        dyl = 0.5 * (y[iy + 3] - y[iy + 1])
        u, v, w = phys[0, iy + 2], phys[1, iy + 2], phys[2, iy + 2]
        best = max(best, float((np.abs(u) / dx + np.abs(v) / dyl + np.abs(w) / dz).max()))
    return best
