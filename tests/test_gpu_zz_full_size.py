"""The largest one-GPU workload at its full size: nx = nz = 512, ny = 1025
(nxd = 1024, nzd = 1536), the mixed form, Re = 20000, dt = 5e-4 (the deck
box512_y1025.in), through a property that does not depend on the size: steps
with the physical-space planes processed in chunks of 200 y rows give the same
state as with every row in one chunk.  The state is compared through outstats
(energy, dissipation, the Reynolds stresses and the mean flows: deterministic
sums over every point, so a checksum of the whole state), bit for bit.

The state holds about 230 GB of the 288 GB; where it does not fit the test is
skipped.  The file sorts after the other GPU tests, so that under pytest -x a
box with less free memory cannot hide them.
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from homogenenousshearturbulence_amd import hst
from homogenenousshearturbulence_amd.deck import read_input

pytestmark = pytest.mark.gpu


def test_box512_mixed_chunked_steps_identical():
    d = read_input(os.path.join(ROOT, "homogenenousshearturbulence_amd", "decks", "box512_y1025.in"))
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
