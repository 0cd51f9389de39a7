"""The Fortran host's modules as a drop-in for the reference's: test programs
written with the call shapes of the reference's Fortran tests
(tests/test_linsolve.f90, tests/test_roundtrip.f90,
tests/test_conservation.f90) are compiled against hst_linsolve,
hst_derivatives, hst_fft, hst_transforms, hst_equations and hst_io
(tests/fortran/) and run on the GPU."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, deck_path

pytestmark = pytest.mark.gpu

FT = os.path.join(ROOT, "tests", "fortran")


def run(tmp_path, prog, deck, precision="fp64"):
    shutil.copy(deck_path(deck), tmp_path / "hst.in")
    print(f"running {prog} on {deck}.in ({precision})", flush=True)
    env = dict(os.environ, HST_PRECISION=precision)
    r = subprocess.run([os.path.join(FT, prog), "hst.in"], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=150)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


def test_line_solve_with_array_arguments(tmp_path):
    """line_solve(sys, lambda, src, dst) on sections of V and rhs: the
    operator of every kind through the ghost rows returns the field to
    1e-11; KIND_DY satisfies D0 dst = D1 src after fill_ghosts_field."""
    out = run(tmp_path, "ftest_linsolve", "small")
    assert out.count("relative error") == 5


def test_transforms_and_files_by_field(tmp_path):
    """transform_to_physical, the products buffer, products_to_spectral and
    vvdz_to_field (FFT/IFT/RFT/HFT of hst_fft on the GPU): the round trip
    returns the field to 1e-12; restart_write/restart_read return it too."""
    out = run(tmp_path, "ftest_roundtrip", "small")
    assert out.count("round trip of component") == 3


def test_conservation_of_the_products(tmp_path):
    """tests/test_conservation.f90's check through the pieces of a substep:
    uu, vv, ww alone and uv + vw + uw inject no energy (rate < 1e-2)."""
    out = run(tmp_path, "ftest_conservation", "pressure")
    assert out.count("(dq2/dt)/q2") == 8
