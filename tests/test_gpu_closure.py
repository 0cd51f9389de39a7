"""The learned closure on the GPU against the oracle (through the C ABI).

Tolerances: the filter, the transforms and the samples are FP64 (1e-12, the
samples then rounded to FP32); the network runs on the matrix cores with
bfloat16 operands, so it is compared with the oracle's bf16-rounded form at
the level that FP32 accumulation order and tanhf leave (rel L2 2e-3), and
with the FP64 form at the bf16 level (3e-2).
"""
import dataclasses

import numpy as np
import pytest

from conftest import ROOT
from homogenenousshearturbulence_amd import closure, hst
from homogenenousshearturbulence_amd.deck import read_input

pytestmark = pytest.mark.gpu


def fine_deck():
    d = read_input(f"{ROOT}/homogenenousshearturbulence_amd/decks/box32_y65.in")
    return dataclasses.replace(d, re=400.0)


def coarse_deck():
    return dataclasses.replace(fine_deck(), nx=16, nz=16, ny=17)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = np.zeros(closure.NPARAM, dtype=np.float32)
    for name, a in closure.unpack(p).items():
        a[...] = rng.standard_normal(a.shape) * (scale if name.startswith("W") else 0.1)
    return p


def synthetic(n, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((n, 9)) * np.linspace(0.5, 4.0, 9) + np.linspace(-1, 1, 9)).astype(np.float32)
    A = rng.standard_normal((9, 6)) * 0.3
    y = (np.tanh(x @ A) + 0.3 * (x[:, :6] * x[:, 3:]) + 0.05 * rng.standard_normal((n, 6))).astype(np.float32)
    return x, y


def norm_of(x, y):
    return np.concatenate([x.mean(0), x.std(0), y.mean(0), y.std(0)]).astype(np.float64)


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    return oracle


def test_network_forward_matches_oracle(O):
    c = closure.Closure(coarse_deck(), 0)
    try:
        x, y = synthetic(1000, 1)  # 1000: a partly filled last tile
        nm = norm_of(x, y)
        p = random_params(2)
        c.set_params(p)
        c.set_norm(nm)
        got = c.predict(x)
        out = O.mlp_forward(p, nm, x, bf16=True)
        want = (out.astype(np.float32) * nm[24:30].astype(np.float32) + nm[18:24].astype(np.float32))
        print("forward rel L2 vs bf16 oracle", rel(got, want))
        assert rel(got, want) < 2e-3
        w64 = O.mlp_forward(p, nm, x, bf16=False) * nm[24:30] + nm[18:24]
        assert rel(got, w64) < 3e-2
    finally:
        c.close()


def test_network_gradient_matches_oracle(O):
    c = closure.Closure(coarse_deck(), 4096)
    try:
        x, y = synthetic(3000, 3)
        c.set_dataset(x, y)
        nm = c.normalise()
        assert np.allclose(nm, norm_of(x.astype(np.float64), y.astype(np.float64)), rtol=1e-10, atol=1e-12)
        c.init_params(7)
        p = c.params()
        for first, n in ((0, 3000), (1000, 517)):
            g, loss = c.grad(first, n)
            lo, go = O.mlp_grad(p, nm, x[first:first + n], y[first:first + n], bf16=True)
            l64, g64 = O.mlp_grad(p, nm, x[first:first + n], y[first:first + n], bf16=False)
            print(first, n, "grad rel L2 vs bf16 oracle", rel(g, go), "vs FP64", rel(g, g64), "loss", loss, lo)
            assert rel(g, go) < 2e-3
            assert rel(g, g64) < 3e-2
            assert abs(loss - lo) < 1e-3 * lo
            # every parameter block carries gradient
            for name, blk in closure.unpack(g).items():
                assert np.abs(blk).max() > 0, name
    finally:
        c.close()


def test_adam_step_and_training(O):
    c = closure.Closure(coarse_deck(), 8192)
    try:
        x, y = synthetic(8192, 5)
        c.set_dataset(x, y)
        c.normalise()
        c.init_params(11)
        p0 = c.params().astype(np.float64)
        g, loss0 = c.grad(0, 1024)
        c.set_params(p0.astype(np.float32))
        # one batch holding exactly the first 1024 samples: n_train = batch
        lr = 1e-3
        c.train(1, 1024, lr, seed=3, n_train=1024)
        p1 = c.params().astype(np.float64)
        g = g.astype(np.float64)
        # first Adam step: -lr m^ / (sqrt(v^) + eps) = -lr g / (|g| + eps)
        want = p0 - lr * g / (np.abs(g) + 1e-8)
        big = np.abs(g) > 1e-4 * np.abs(g).max()
        assert np.abs(p1 - want)[big].max() < 1e-3 * lr
        assert np.abs(p1 - p0).max() <= lr * (1 + 1e-5)
        # training lowers the loss of the synthetic map; the held-out loss follows
        c.init_params(11)
        loss, val = c.train(30, 1024, 3e-3, seed=5, n_train=7168)
        print("train loss", loss[[0, 9, 19, 29]], "held out", val[[0, 9, 19, 29]])
        assert loss[-1] < 0.25 * loss[0]
        assert val[-1] < 0.3 * val[0]
        assert np.all(np.isfinite(loss)) and np.all(np.isfinite(val))
    finally:
        c.close()


@pytest.fixture(scope="module")
def filtered(O):
    """A fine FP64 state after a few steps, filtered onto the coarse grid."""
    fd, cd = fine_deck(), coarse_deck()
    s = hst.Solver(fd, hst.FP64)
    c = closure.Closure(cd, 1 << 20)
    try:
        s.time = 0.0
        s.initial_field()
        s.fill_all_ghosts()
        for _ in range(4):
            s.timestep(fd.deltat)
        s.nonlinear()
        c.filter(s.time)
        V = s.V
        P = np.stack([s.product(p + 1) for p in range(6)])
        U = c.velocity()
        Pbar = np.stack([c.field(closure.PBAR + p) for p in range(6)])
        n = c.sample(s.time)
        X, Y = c.dataset()
        yield dict(fd=fd, cd=cd, t=s.time, V=V, P=P, U=U, Pbar=Pbar, X=X, Y=Y, n=n)
    finally:
        c.close()
        s.close()


def test_filter_matches_oracle(O, filtered):
    f = filtered
    fo, co = O.Oracle(f["fd"]), O.Oracle(f["cd"])
    sx, sz = co.shear_shifts(f["t"])
    for k, src in enumerate(list(f["V"]) + list(f["P"])):
        want = O.filter_field(fo, co, src)
        O.fill_ghosts_of(co, want, sx, sz)
        got = f["U"][k] if k < 3 else f["Pbar"][k - 3]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-13, (k, err)


def test_samples_match_oracle(O, filtered):
    f = filtered
    co = O.Oracle(f["cd"])
    sx, sz = co.shear_shifts(f["t"])
    x, y = O.samples(co, f["U"], f["Pbar"], sx, sz)
    assert f["n"] == x.shape[0] == f["cd"].ny * co.info.nzd * 2 * co.info.nxd
    ex = np.abs(f["X"] - x).max(axis=0) / np.abs(x).max(axis=0)
    ey = np.abs(f["Y"] - y).max(axis=0) / np.abs(y).max(axis=0)
    print("sample errors", ex.max(), ey.max())
    assert ex.max() < 1e-6 and ey.max() < 1e-5
    # the stresses of a filtered field: tau_uu, vv, ww are positive on average
    assert (y[:, :3].mean(axis=0) > 0).all()


def test_stress_matches_oracle(O, filtered):
    f = filtered
    cd = f["cd"]
    c = closure.Closure(cd, 0)
    try:
        nm = norm_of(f["X"].astype(np.float64), f["Y"].astype(np.float64))
        p = random_params(4, scale=0.2)
        c.set_params(p)
        c.set_norm(nm)
        for i in range(3):
            c.set_field(i, f["U"][i])
        T = c.stress(f["t"])
        co = O.Oracle(cd)
        sx, sz = co.shear_shifts(f["t"])
        want = O.closure_stress(co, f["U"], p, nm, sx, sz)
        for k in range(6):
            print(k, "stress spectrum rel L2", rel(T[k], want[k]))
            assert rel(T[k], want[k]) < 2e-3
    finally:
        c.close()


def test_attached_step_matches_oracle(O, filtered):
    """A coarse FP64 step with the closure attached against the oracle's step
    with the same stresses added to its products; detached, the closure's
    effect is gone."""
    f = filtered
    cd = f["cd"]
    nm = norm_of(f["X"].astype(np.float64), f["Y"].astype(np.float64))
    p = random_params(6, scale=0.2)
    s = hst.Solver(cd, hst.FP64)
    c = closure.Closure(cd, 0)
    try:
        c.set_params(p)
        c.set_norm(nm)
        s.set_V(f["U"])
        s.reset_old()
        s.time = f["t"]
        c.attach(True)
        s.timestep(cd.deltat)
        Vc = s.V
        c.attach(False)
        s.set_V(f["U"])
        s.reset_old()
        s.time = f["t"]
        s.timestep(cd.deltat)
        Vp = s.V
    finally:
        c.close()
        s.close()
    o, helper = O.Oracle(cd), O.Oracle(cd)
    o.set_V(f["U"])
    o.reset_old()
    o.time = f["t"]
    O.step_with_closure(o, helper, cd.deltat, p, nm)
    for k in range(3):
        e = rel(Vc[k][..., 2:-2], o.V[k][..., 2:-2])
        d = rel(Vp[k][..., 2:-2], Vc[k][..., 2:-2])
        print(k, "attached vs oracle", e, "attached vs plain", d)
        assert e < 1e-6
        assert d > 100 * e


def test_steps_across_attach_and_detach(filtered):
    """Between plain time steps the velocity is held as eta and dv/dy; a step
    with the closure attached (its hook reads u, v, w) restores u, w first,
    and the plain steps after a detach go back to the compact form.  Plain,
    attached, attached, plain, plain steps in a row against the same sequence
    with the velocity read after every step (a restore each time)."""
    f = filtered
    cd = f["cd"]
    nm = norm_of(f["X"].astype(np.float64), f["Y"].astype(np.float64))
    p = random_params(8, scale=0.2)
    res = []
    for read_between in (False, True):
        s = hst.Solver(cd, hst.FP64)
        c = closure.Closure(cd, 0)
        try:
            c.set_params(p)
            c.set_norm(nm)
            s.set_V(f["U"])
            s.reset_old()
            s.time = f["t"]
            for attached in (False, True, True, False, False):
                c.attach(attached)
                s.timestep(cd.deltat)
                if read_between:
                    assert np.all(np.isfinite(s.V))
            c.attach(False)
            res.append(s.V)
        finally:
            c.close()
            s.close()
    print("attach/detach sequence, with and without reads between", rel(res[0], res[1]))
    assert rel(res[0], res[1]) < 1e-12


def test_free_and_init_detach_the_closure(O, filtered):
    """A closure attached to a solver state is detached when that state is
    freed: a state set up afterwards (here on another grid, where the
    closure's sizes would not fit its buffers) steps as the plain method,
    equal to the oracle's plain steps."""
    from conftest import deck_path

    f = filtered
    cd = f["cd"]
    nm = norm_of(f["X"].astype(np.float64), f["Y"].astype(np.float64))
    c = closure.Closure(cd, 0)
    try:
        c.set_params(random_params(6, scale=0.2))
        c.set_norm(nm)
        s = hst.Solver(cd, hst.FP64)
        c.attach(True)
        s.close()
        d2 = read_input(deck_path("small"))
        g = hst.run_deck(d2, hst.FP64, nsteps=3)
        try:
            Vg = g.V
        finally:
            g.close()
    finally:
        c.close()
    o = O.run_deck(d2, nsteps=3)
    e = rel(Vg[..., 2:-2], o.V[..., 2:-2])
    print("plain steps after the freed state vs oracle", e)
    assert e < 1e-10
