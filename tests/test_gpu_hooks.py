"""The physics hooks of the substep on the GPU through the C ABI: the Stokes
layer (src/hst_stokes.f90) and the large-scale damping (src/hst_damping.f90),
against the reference's golden field, its closed-form tests and the oracle.

Tolerances: FP64 pieces 1e-12 relative; whole runs 1e-10 (the reference's
regression bar); the mixed form 1e-5; the closed forms at the reference's bar.
"""
import ctypes
import dataclasses

import numpy as np
import pytest

from conftest import GOLDEN, deck_path
from homogenenousshearturbulence_amd import cplio, hst
from homogenenousshearturbulence_amd.deck import read_input
import closed_forms

pytestmark = pytest.mark.gpu

RK = hst.RK_RAI


@pytest.fixture(scope="module")
def O(oracle_built):
    return oracle_built


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _factory(precision=hst.FP64):
    made = []

    def make(deck):
        for s in made:
            s.close()
        made.clear()
        s = hst.Solver(deck, precision)
        made.append(s)
        return s

    return make, made


@pytest.mark.parametrize("precision,tol", [(hst.FP64, 1e-10), (hst.MIXED, 1e-5)])
def test_golden_stokes_layer(precision, tol):
    """tests/regression.sh with small_stokes: 50 steps of the layer with its
    body force on the stretched grid vs the reference's stored field."""
    d = read_input(deck_path("small_stokes"))
    g = hst.run_deck(d, precision)
    try:
        t, Vr, _ = cplio.read_field(f"{GOLDEN}/reference/small_stokes.fld")
        assert g.nsteps == 50 and g.time == t
        err = cplio.compare_fields(Vr, g.V)
        print("small_stokes", precision, err)
        assert err <= tol
    finally:
        g.close()


def test_stokes_layer_closed_form():
    make, made = _factory()
    try:
        err, other = closed_forms.stokes_layer(make, read_input(deck_path("stokes")))
    finally:
        for s in made:
            s.close()
    print("Stokes layer", err, other)
    assert err <= 1e-4 and other <= 1e-12


DAMPED = {"kelvin_damped": 2.0, "kelvin_hard": 1.0, "kelvin_outside": 0.0}


@pytest.mark.parametrize("name", list(DAMPED))
def test_kelvin_damping(name):
    make, made = _factory()
    try:
        err_v, err_eta, _ = closed_forms.kelvin(make, read_input(deck_path(name)), DAMPED[name])
    finally:
        for s in made:
            s.close()
    print(name, err_v, err_eta)
    if name == "kelvin_hard":
        assert err_v <= 1e-12 and err_eta <= 1e-12
    else:
        assert err_v <= 1e-3 and err_eta <= 1e-3


def _damped_small(**kw):
    """The small deck with damping on the large scales (S2 oscillating so
    that the region turns every substep with damp_follow)."""
    d = read_input(deck_path("small"))
    base = dict(damp_kpar=4.5, damp_kperp=13.0, damp_rate=3.0, damp_edge=0.3, damp_mean=True,
                s2_amplitude=0.6, s2_period=0.5)
    base.update(kw)
    return dataclasses.replace(d, **base)


@pytest.mark.parametrize("hard", [False, True])
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("precision,tol", [(hst.FP64, 1e-12), (hst.MIXED, 1e-5)])
def test_damped_steps_match_oracle(O, hard, exact, precision, tol):
    """Three steps of the seeded field with soft or hard damping (the region
    following the turning shear direction) vs the oracle; the statistics and
    the damping line of outstats too."""
    d = _damped_small(damp_hard=hard, exact_shift=exact)
    o = O.run_deck(d, nsteps=3)
    g = hst.run_deck(d, precision, nsteps=3)
    try:
        assert g.time == o.time
        assert rel(g.V, o.V) <= tol
        so, sg = o.outstats(), g.outstats()
        assert np.abs(sg[:8] - so[:8]).max() <= max(tol, 1e-12) * np.abs(so[:8]).max() * 10
        eo, eg = o.outstats_extra(), g.outstats_extra()
        print("damping line", eg, eo)
        assert eg[5] == eo[5] and eg[5] > 0
        assert abs(eg[4] - eo[4]) <= 1e-12
        assert np.abs(eg[6:] - eo[6:]).max() <= max(tol, 1e-12) * np.abs(eo[6:]).max() * 10
    finally:
        g.close()
    # the damping is felt: the undamped run differs
    o2 = O.run_deck(dataclasses.replace(d, damp_kpar=0.0, damp_kperp=0.0), nsteps=3)
    assert rel(o.V, o2.V) > 1e3 * tol


def test_damping_pieces_match_oracle(O):
    """buildrhs_prepare with the soft damping term and the stand-alone hard
    mask against the oracle's pieces."""
    for hard in (False, True):
        d = _damped_small(damp_hard=hard)
        t = 0.13
        o = O.Oracle(d)
        o.initial_field()
        o.fill_all_ghosts(t)
        o.time = t
        rng = np.random.default_rng(3)
        for f in (6, 7):
            o.field(f)[...] = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
        old = [o.field(6).copy(), o.field(7).copy()]
        V0 = o.V.copy()
        o.nonlinear(mode=0)
        ode = RK[2]
        arr = (ctypes.c_double * 3)(*ode)
        O.lib().hsto_buildrhs_prepare(o.h, arr, 0.002, t)
        O.lib().hsto_buildrhs(o.h, arr, 1)
        O.lib().hsto_buildrhs(o.h, arr, 2)
        o.damping_mask()
        want = [o.field(f).copy() for f in (4, 5, 6, 7)]
        g = hst.Solver(d, hst.FP64)
        try:
            g.set_V(V0)
            g.upload(hst.OLD1, old[0])
            g.upload(hst.OLD2, old[1])
            g.nonlinear()
            g.buildrhs_prepare(ode, 0.002, t)
            g.buildrhs(ode, 1)
            g.buildrhs(ode, 2)
            g.damping_mask()
            got = [g.download(f) for f in (4, 5, 6, 7)]
        finally:
            g.close()
        for a, b in zip(got, want):
            assert rel(a[..., 2:-2], b[..., 2:-2]) <= 1e-12
        if hard:  # some columns are masked
            assert np.sum(np.abs(want[0][..., 2:-2]).max(axis=-1) == 0) > 0


@pytest.mark.parametrize("precision,tol", [(hst.FP64, 1e-12), (hst.MIXED, 1e-5)])
def test_stokes_statistics_match_oracle(O, precision, tol):
    """outstats with the layer (the (0,0) column left out) and the region line
    of stokes_runtime.dat after two steps of small_stokes."""
    d = read_input(deck_path("small_stokes"))
    o = O.run_deck(d, nsteps=2)
    g = hst.run_deck(d, precision, nsteps=2)
    try:
        so, sg = o.outstats(), g.outstats()
        eo, eg = o.outstats_extra(), g.outstats_extra()
    finally:
        g.close()
    print("stokes line", eg, eo)
    assert np.abs(sg[:8] - so[:8]).max() <= 10 * tol * np.abs(so[:8]).max()
    assert np.abs(eg[:4] - eo[:4]).max() <= 10 * tol * np.abs(eo[:4]).max()
    assert eg[0] > 0 and eg[1] > 0


def test_project_divergence_matches_oracle(O):
    """project_restart on the GPU (one lane per column, the original rows kept
    in registers) against the oracle, at a time with a nonzero shear wrap."""
    d = read_input(deck_path("small_s2x"))
    t = 0.31
    o = O.Oracle(d)
    o.initial_field()
    rng = np.random.default_rng(9)
    for k in (1, 2, 3):
        o.field(k)[..., 2:-2] += 0.01 * (rng.standard_normal(o.shape)[..., 2:-2] + 1j * rng.standard_normal(o.shape)[..., 2:-2])
    V0 = o.V.copy()
    o.time = t
    o.project_divergence()
    g = hst.Solver(d, hst.FP64)
    try:
        g.set_V(V0)
        g.time = t
        g.project_divergence()
        Vg = g.V
    finally:
        g.close()
    assert rel(Vg, o.V) <= 1e-12
    assert rel(o.V, V0) > 1e-6  # the repair changed the field
