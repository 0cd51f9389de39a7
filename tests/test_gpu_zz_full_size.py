"""The one-GPU workloads at their full sizes.

* nx = nz = 128, ny = 257 (Re = 4000, dt = 0.002, the default shift; the deck
  box128_y257.in): two steps against the oracle, which finishes them in
  seconds on the host's cores, per field at the reference's 1e-10 regression
  bar (FP64) and at 1e-5 (the mixed form).
* nx = nz = 256, ny = 513 with the exact shift (Re = 8000, dt = 0.001; the
  deck box256_y513_shift.in): one step against the oracle (about 50 s
  of the host's cores), at 1e-10 (FP64) and 1e-5 (the mixed form).
* nx = nz = 512, ny = 1025 (nxd = 1024, nzd = 1536), the mixed form, Re =
  20000, dt = 5e-4 (box512_y1025.in), through a property that does not
  depend on the size: steps with the physical-space planes processed in
  chunks of 200 y rows give the same state as with every row in one chunk.
  The state is compared through outstats (energy, dissipation, the Reynolds
  stresses and the mean flows: deterministic sums over every point, so a
  checksum of the whole state), bit for bit.  The state holds about 230 GB of
  the 288 GB; where it does not fit the test is skipped.

The file sorts after the other GPU tests, so that under pytest -x a box with
less free memory cannot hide them.
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from homogenenousshearturbulence_amd import hst
from homogenenousshearturbulence_amd.deck import read_input
import refops

pytestmark = pytest.mark.gpu

DECKS = os.path.join(ROOT, "homogenenousshearturbulence_amd", "decks")


@pytest.fixture(scope="module")
def box128_oracle(oracle_built):
    d = read_input(os.path.join(DECKS, "box128_y257.in"))
    return d, oracle_built.run_deck(d, nsteps=2)


@pytest.mark.parametrize("precision,tol", [(hst.FP64, 1e-10), (hst.MIXED, 1e-5)])
def test_box128_steps_match_oracle(box128_oracle, precision, tol):
    d, o = box128_oracle
    g = hst.run_deck(d, precision, nsteps=2)
    try:
        assert g.time == o.time
        V = g.V
        errs = [refops.rel_l2(V[c][..., 2:-2], o.V[c][..., 2:-2]) for c in range(3)]
    finally:
        g.close()
    print("nx = nz = 128, ny = 257, two steps: rel L2 (u, v, w)", errs)
    assert max(errs) <= tol


@pytest.fixture(scope="module")
def box256_oracle(oracle_built):
    d = read_input(os.path.join(DECKS, "box256_y513_shift.in"))
    return d, oracle_built.run_deck(d, nsteps=1)


@pytest.mark.parametrize("precision,tol", [(hst.FP64, 1e-10), (hst.MIXED, 1e-5)])
def test_box256_shift_step_matches_oracle(box256_oracle, precision, tol):
    d, o = box256_oracle
    g = hst.run_deck(d, precision, nsteps=1)
    try:
        assert g.time == o.time
        V = g.V
        errs = [refops.rel_l2(V[c][..., 2:-2], o.V[c][..., 2:-2]) for c in range(3)]
    finally:
        g.close()
    print("nx = nz = 256, ny = 513, exact shift, one step: rel L2 (u, v, w)", errs)
    assert max(errs) <= tol


def test_box512_mixed_chunked_steps_identical():
    d = read_input(os.path.join(DECKS, "box512_y1025.in"))
    V0 = hst.initial_field(d)
    stats = []
    for chunk in (0, 200):
        try:
            s = hst.Solver(d, hst.MIXED, chunk)
        except hst.HstError as e:
            pytest.skip(f"the nx = nz = 512, ny = 1025 state does not fit on this device: {e}")
        try:
            s.set_V(V0)
            s.time = d.time
            s.fill_all_ghosts()
            for _ in range(2):
                s.timestep(d.deltat)
            stats.append(s.outstats())
        finally:
            s.close()
    print("outstats after two steps", stats[0])
    assert np.all(np.isfinite(stats[0])) and stats[0][0] > 0
    assert np.array_equal(stats[0], stats[1])
