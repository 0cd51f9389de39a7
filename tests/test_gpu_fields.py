"""The per-field entry points behind the Fortran host's hst_fft and
hst_transforms (FFT, IFT, RFT, HFT, build_products, and the product spectra
handed to buildrhs), on the reference's buffer shapes, against numpy's FFTs
(an independent FP64 implementation) and the oracle's right-hand sides."""
import ctypes
import dataclasses

import numpy as np
import pytest

from conftest import deck_path
from homogenenousshearturbulence_amd import hst
from homogenenousshearturbulence_amd.deck import read_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    d = read_input(deck_path("small"))
    s = hst.Solver(d, hst.FP64)
    yield s
    s.close()


def test_z_transforms_match_numpy(solver):
    d, L = solver.deck, hst.lib()
    n, nlines = d.nzd, (d.nx + 1) * (d.ny + 4)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((nlines, n)) + 1j * rng.standard_normal((nlines, n))
    for sign, want in ((-1, np.fft.fft(a, axis=1)), (+1, np.fft.ifft(a, axis=1) * n)):
        b = a.copy()
        hst._check(L.hst_gpu_fft_z(b.ctypes.data, nlines, sign), "fft_z")
        assert np.abs(b - want).max() <= 1e-13 * np.abs(want).max(), sign


def test_x_real_transforms_match_numpy(solver):
    d, L = solver.deck, hst.lib()
    m = d.nxd
    nlines = d.nzd * (d.ny + 4)
    rng = np.random.default_rng(4)
    # c2r (RFT): modes 0..m, imaginary parts of 0 and m ignored
    X = rng.standard_normal((nlines, m + 1)) + 1j * rng.standard_normal((nlines, m + 1))
    b = X.copy()
    hst._check(L.hst_gpu_fft_x(b.ctypes.data, nlines, 1), "fft_x c2r")
    Xh = X.copy()
    Xh[:, 0] = Xh[:, 0].real
    Xh[:, m] = Xh[:, m].real
    want = np.fft.irfft(Xh, n=2 * m, axis=1) * (2 * m)
    got = b.view(np.float64).reshape(nlines, 2 * m + 2)
    assert np.abs(got[:, :2 * m] - want).max() <= 1e-13 * np.abs(want).max()
    assert not got[:, 2 * m:].any()
    # r2c (HFT): 2m real points (the two spare reals of a line are not read)
    x = rng.standard_normal((nlines, 2 * m + 2))
    x[:, 2 * m:] = 7.0
    b = np.ascontiguousarray(x).view(np.complex128).copy()
    hst._check(L.hst_gpu_fft_x(b.ctypes.data, nlines, -1), "fft_x r2c")
    want = np.fft.rfft(x[:, :2 * m], axis=1)
    assert np.abs(b - want).max() <= 1e-13 * np.abs(want).max()


def test_products_of_a_group(solver):
    d, L = solver.deck, hst.lib()
    lw = 2 * (d.nxd + 1)
    nlines = d.nzd * (d.ny + 4)
    rng = np.random.default_rng(5)
    vel = rng.standard_normal((3, nlines, lw))
    f = 1.0 / (2.0 * d.nxd * d.nzd)
    for g, pairs in ((1, ((0, 0), (1, 1), (2, 2))), (2, ((0, 1), (1, 2), (0, 2)))):
        prod = rng.standard_normal((3, nlines, lw))
        before = prod.copy()
        hst._check(L.hst_gpu_products(g, nlines, vel.ctypes.data, prod.ctypes.data), "products")
        for p, (a, b) in enumerate(pairs):
            want = vel[a, :, :lw - 2] * vel[b, :, :lw - 2] * f
            assert np.array_equal(prod[p, :, :lw - 2], want), (g, p)
        assert np.array_equal(prod[:, :, lw - 2:], before[:, :, lw - 2:])


@pytest.mark.parametrize("precision", [hst.FP64, hst.MIXED])
def test_buildrhs_from_given_spectra(oracle_built, precision):
    """hst_gpu_set_products with the oracle's product spectra in the VVdz
    shape, then buildrhs(ODE, g): the right-hand sides of the oracle's
    buildrhs (FP64 1e-12; mixed: FP32 spectra and stencils, 1e-5)."""
    O = oracle_built
    d = dataclasses.replace(read_input(deck_path("small")), time=0.21)
    o = O.Oracle(d)
    o.time = d.time
    o.initial_field()
    o.fill_all_ghosts()
    o.nonlinear(0)
    ode = np.array(hst.RK_RAI[1])
    c = (ctypes.c_double * 3)(*ode)
    s = hst.Solver(d, precision)
    try:
        s.set_V(o.V)
        s.reset_old()
        s.buildrhs_prepare(ode, 0.002, d.time)
        O.lib().hsto_buildrhs_prepare(o.h, c, 0.002, d.time)
        P = o.products()
        for g in (1, 2):
            # VVdz(nzd, nx+1, ny+4, 3) in Fortran order = [3][ny+4][nx+1][nzd]
            vvdz = np.zeros((3, d.ny + 4, d.nx + 1, d.nzd), dtype=np.complex128)
            for m in range(3):
                for iz in range(-d.nz, d.nz + 1):
                    vvdz[m, :, :, iz % d.nzd] = P[3 * (g - 1) + m][:, iz + d.nz, :].T
            hst._check(hst.lib().hst_gpu_set_products(g, vvdz.ctypes.data), "set_products")
            s.buildrhs(ode, g)
            O.lib().hsto_buildrhs(o.h, c, g)
        tol = 1e-12 if precision == hst.FP64 else 1e-5
        for f in (hst.RHS1, hst.RHS2):
            got, want = s.download(f)[..., 2:-2], o.field(f)[..., 2:-2]
            assert np.abs(got - want).max() <= tol * np.abs(want).max(), f
    finally:
        s.close()


@pytest.mark.parametrize("precision", [hst.FP64, hst.MIXED])
def test_fused_step_conserves_energy(oracle_built, precision):
    """The fused time step (which transforms vv - uu, vv - ww, uv, vw, uw
    instead of the six products) conserves energy like the six-product
    oracle: S = 0, nu = 1e-12, the rate (dq2/dt)/q2 of a step of 1e-4 after a
    first step; FP64 equal to the oracle's to 1e-8, mixed below 1e-2."""
    O = oracle_built
    d = dataclasses.replace(read_input(deck_path("pressure")), S=0.0, re=1e12)
    dt = 1e-4

    def q(V):
        w = np.full(d.nx + 1, 2.0)
        w[0] = 1.0
        return float((w[:, None, None] * (np.abs(V[:, :, :, 2:d.ny + 2]) ** 2).sum(axis=0)).sum() / d.ny)

    s = hst.Solver(d, precision)
    try:
        s.time = 0.0
        s.initial_field()
        s.fill_all_ghosts(0.0)
        s.reset_old()
        s.timestep(dt)
        q0 = q(s.V)
        s.timestep(dt)
        rate = (q(s.V) - q0) / dt / q0
    finally:
        s.close()
    o = O.Oracle(d)
    o.time = 0.0
    o.initial_field()
    o.fill_all_ghosts(0.0)
    o.timestep(dt)
    q0 = q(o.V)
    o.timestep(dt)
    want = (q(o.V) - q0) / dt / q0
    assert abs(rate) <= 1e-2
    if precision == hst.FP64:
        assert abs(rate - want) <= 1e-8, (rate, want)
