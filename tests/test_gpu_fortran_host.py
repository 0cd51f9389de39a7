"""The Fortran host (homogenenousshearturbulence_amd/fortran/hst) as a drop-in
for the reference's driver: tests/regression.sh runs the decks in a fresh
directory for 50 steps and compares Dati.cart.out with the stored fields at
1e-10 (compare_fields.py's max norm); the same here, on the GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, deck_path
from homogenenousshearturbulence_amd import cplio

pytestmark = pytest.mark.gpu

HST = os.path.join(ROOT, "homogenenousshearturbulence_amd", "fortran", "hst")


def run_hst(tmp_path, name, precision="fp64"):
    shutil.copy(deck_path(name), tmp_path / "hst.in")
    env = dict(os.environ, HST_PRECISION=precision)
    print(f"running the Fortran driver on {name}.in ({precision})", flush=True)
    r = subprocess.run([HST, "hst.in"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


@pytest.mark.parametrize("name", ["small", "small_s2x", "small_stokes"])
def test_regression_decks(tmp_path, name):
    run_hst(tmp_path, name)
    out = tmp_path / "Dati.cart.out"
    t, V, _ = cplio.read_field(str(out))
    tr, Vr, _ = cplio.read_field(f"{GOLDEN}/reference/{name}.fld")
    assert t == tr
    assert cplio.compare_fields(Vr, V) <= 1e-10
    # the header (text and the five raw doubles) is the reference's, byte for byte
    a, b = open(out, "rb").read(), open(f"{GOLDEN}/reference/{name}.fld", "rb").read()
    ia, ib = a.find(b"Vfield=\n"), b.find(b"Vfield=\n")
    assert a[:ia] == b[:ib]
    rt = np.loadtxt(tmp_path / "statistics" / "Runtimedata")
    if name == "small_stokes":  # t_max = 0.025: t = 0, 0.01, 0.02; the region line beside it
        assert rt.shape == (3, 13)
        sl = np.loadtxt(tmp_path / "statistics" / "stokes_runtime.dat")
        assert sl.shape == (3, 5) and np.array_equal(sl[:, 0], rt[:, 0]) and (sl[:, 1:] > 0).all()
    else:
        assert rt.shape == (11, 13)  # t = 0 and every dt_stat = 0.01 up to 0.1
        assert abs(rt[-1, 0] - tr) < 1e-12


def test_mixed_precision_driver(tmp_path):
    run_hst(tmp_path, "small", precision="mixed")
    _, V, _ = cplio.read_field(str(tmp_path / "Dati.cart.out"))
    _, Vr, _ = cplio.read_field(f"{GOLDEN}/reference/small.fld")
    err = cplio.compare_fields(Vr, V)
    assert 0 < err <= 1e-5, err


def test_restart_continues_exactly(tmp_path):
    """tests/crossbranch.sh: 25 steps, restart, 25 more = the 50-step field."""
    d = tmp_path / "a"
    d.mkdir()
    text = open(deck_path("small")).read().replace("t_max = 0.1", "t_max = 0.05")
    (d / "hst.in").write_text(text)
    r = subprocess.run([HST, "hst.in"], cwd=d, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:]
    text2 = open(deck_path("small")).read().replace("time_from_restart = .false.", "time_from_restart = .true.")
    (d / "hst.in").write_text(text2)
    r = subprocess.run([HST, "hst.in"], cwd=d, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:]
    t, V, _ = cplio.read_field(str(d / "Dati.cart.out"))
    tr, Vr, _ = cplio.read_field(f"{GOLDEN}/reference/small.fld")
    assert abs(t - tr) < 1e-12
    assert cplio.compare_fields(Vr, V) <= 1e-10


def test_snapshots_pressure_and_spectrum(tmp_path, oracle_built):
    """With dt_field = 0.05 the driver writes, beside fields/field<n>.fld, the
    pressure p_fields/pField<n>.fld (headerless, one component; src/hst.f90:
    104-116) and statistics/spectra/exz<n>.dat.  The pressure equals the
    oracle's compute_pressure of the snapshot's velocity at 1e-12; the
    spectrum's sample count is that of the outstats calls since the previous
    snapshot and its sum is ly times their mean energy."""
    from homogenenousshearturbulence_amd.deck import read_input

    text = open(deck_path("small")).read().replace("dt_field = 1e9", "dt_field = 0.05")
    (tmp_path / "hst.in").write_text(text)
    d = read_input(str(tmp_path / "hst.in"))
    r = subprocess.run([HST, "hst.in"], cwd=tmp_path, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rt = np.loadtxt(tmp_path / "statistics" / "Runtimedata")
    t_prev = -1.0
    for n, nsamp in ((1, 6), (2, 5)):
        t, V, _ = cplio.read_field(str(tmp_path / "fields" / f"field{n}.fld"))
        p = np.fromfile(tmp_path / "p_fields" / f"pField{n}.fld", dtype=np.complex128)
        p = p.reshape((d.nx + 1, 2 * d.nz + 1, d.ny + 4))
        o = oracle_built.Oracle(d)
        o.set_V(V)
        o.fill_all_ghosts(t)
        o.time = t
        want = o.pressure()
        err = np.abs(p[..., 2:-2] - want[..., 2:-2]).max() / np.abs(want[..., 2:-2]).max()
        assert err <= 1e-12, (n, err)
        assert np.array_equal(p[..., -2:], p[..., 2:4]) and np.array_equal(p[..., :2], p[..., -4:-2])
        path = tmp_path / "statistics" / "spectra" / f"exz{n}.dat"
        head = open(path).readline()
        assert head.startswith("# column energies ly/2 sum_y |u|^2 averaged over t =")
        assert f"({nsamp} samples)" in head
        E = np.loadtxt(path)
        assert E.shape == ((d.nx + 1) * (2 * d.nz + 1), 5)
        assert np.array_equal(E[: 2 * d.nz + 1, 1], np.arange(-d.nz, d.nz + 1))
        sel = (rt[:, 0] > t_prev + 1e-9) & (rt[:, 0] <= t + 1e-9)
        assert sel.sum() == nsamp
        assert abs(E[:, 4].sum() - d.ly * rt[sel, 9].mean()) <= 1e-6 * E[:, 4].sum()  # 8 printed digits
        t_prev = t


@pytest.mark.parametrize("remap", ["strict", "mode_index"])
def test_restart_remap(tmp_path, oracle_built, remap):
    """restart_read of a file of another resolution (src/hst_io.f90:63-190):
    the common Fourier modes copied, the nearest file row per target row of a
    stretched grid; with restart_remap = 'mode_index' and another alpha0 the
    divergence repair (project_restart) follows.  t_max = 0: the driver writes
    the loaded field back as Dati.cart.out, compared with the oracle's remap
    (exactly) and repair (1e-12)."""
    from oracle import oracle as O

    from homogenenousshearturbulence_amd.deck import read_input

    src = str(tmp_path / "src.fld")
    shutil.copy(f"{GOLDEN}/reference/small.fld", src)
    text = open(deck_path("small")).read()
    text = text.replace("nx = 15, ny = 32, nz = 15", "nx = 21, ny = 45, nz = 11")
    text = text.replace("ly = 2.0 /", "ly = 2.0, ystretch = 1.2 /")
    text = text.replace("t_max = 0.1", "t_max = 0.0").replace("dt_stat = 0.01", "dt_stat = 1e9")
    text = text.replace("time_from_restart = .false.", "time_from_restart = .true., restart_file = 'src.fld', "
                        f"restart_remap = '{remap}'")
    if remap == "mode_index":
        text = text.replace("alfa0 = 2.0943951023931953", "alfa0 = 1.5")
    (tmp_path / "hst.in").write_text(text)
    d = read_input(str(tmp_path / "hst.in"))
    assert (d.nx, d.ny, d.nz, d.ystretch) == (21, 45, 11, 1.2)
    r = subprocess.run([HST, "hst.in"], cwd=tmp_path, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "modes and rows remapped" in r.stdout
    t, V, sizes = cplio.read_field(str(tmp_path / "Dati.cart.out"))
    tf, Vf, sf = cplio.read_field(src)
    assert t == tf and sizes == {"nx": 21, "nz": 11, "ny": 45}
    o = O.Oracle(d)
    want = O.remap_field(Vf, sf, d, o.y)
    if remap == "strict":
        assert np.array_equal(V[..., 2:-2], want[..., 2:-2])
    else:
        o.set_V(want)
        o.time = t
        o.project_divergence()
        err = np.abs(V[..., 2:-2] - o.V[..., 2:-2]).max() / np.abs(o.V).max()
        assert err <= 1e-12, err
        assert np.abs(V[..., 2:-2] - want[..., 2:-2]).max() > 1e-6
