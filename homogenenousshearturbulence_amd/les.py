"""The learned closure end to end: a DNS generates the data, the network is
trained on the GPU, and a coarse run with the network is measured against the
filtered DNS beside the coarse run without it (the plain method).

    python -m homogenenousshearturbulence_amd.les [--out result.json] [--quick]

Steps (no counterpart in the reference; every GPU step is libhstgpu.so):
  1. DNS of the fine grid (box128_y257: nx = nz = 128, ny = 257) at Re = 1000
     and dt = 0.001 from the seeded field (the deck's own Re = 4000 is not
     resolved on 128 modes: its CFL number grows from 0.6 to 2.7 between
     t = 0.2 and 0.3 at constant energy and the run diverges by t = 0.35); at
     the training times the velocity and its six products are
     filtered onto the coarse grid (les32_y65: nx = nz = 32, ny = 65) and the
     (gradients, stresses) of every coarse grid point join the data set; at
     the record times the filtered velocity is kept as the yardstick.
  2. Training: Adam over the data set, the last snapshot held out.
  3. Coarse runs from the filtered DNS field at t0: without the closure and
     with it attached; the same records.
  4. The yardstick: the filtered DNS's energy, dissipation and uv stress histories
     and its x and z spectra against each coarse run.

If the run with the closure stops producing finite statistics it is stopped
and reported as diverged; the plain coarse run is the fallback.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import time as _time

import numpy as np

from . import closure, hst
from .deck import Deck, read_input

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "decks")


@dataclasses.dataclass
class Plan:
    dns: Deck
    les: Deck
    precision: int = hst.FP64
    t_train: tuple = (0.2, 0.3, 0.4, 0.5, 0.6, 0.7)
    holdout: int = 1          # last snapshots held out of the training
    t0: float = 0.7           # start of the coarse runs
    t_end: float = 1.5
    dt_record: float = 0.1
    epochs: int = 200
    batch: int = 65536
    lr: float = 1e-3
    seed: int = 1
    # weights of the six stresses (uu vv ww uv vw uw) in the loss: a target's
    # standard deviation in the standardisation is divided by sqrt(weight), so
    # its standardised errors count weight times more (the network and its
    # de-standardised output are the same model otherwise)
    target_weight: tuple = (1.0, 1.0, 1.0, 1.0, 1.0, 1.0)

    @property
    def record_times(self):
        n = int(round((self.t_end - self.t0) / self.dt_record))
        return tuple(round(self.t0 + k * self.dt_record, 10) for k in range(n + 1))


def default_plan(quick: bool = False) -> Plan:
    les = read_input(os.path.join(DECKS, "les32_y65.in"))
    dns = dataclasses.replace(read_input(os.path.join(DECKS, "box128_y257.in")), re=les.re, deltat=0.001)
    p = Plan(dns=dns, les=les)
    if quick:
        p = dataclasses.replace(p, t_train=(0.2, 0.3, 0.4), t0=0.4, t_end=0.8, dt_record=0.1, epochs=60)
    return p


def _steps_to(s, t: float, dt: float) -> None:
    n = int(round((t - s.time) / dt))
    for _ in range(n):
        s.timestep(dt)


def run_dns(plan: Plan, log=print):
    """Steps 1: the data set (in a closure context) and the filtered records."""
    s = hst.Solver(plan.dns, plan.precision)
    npts = plan.les.ny * plan.les.nzd * 2 * plan.les.nxd
    c = closure.Closure(plan.les, capacity=npts * len(plan.t_train))
    records = {}
    try:
        s.time = 0.0
        s.initial_field()
        s.fill_all_ghosts()
        events = sorted(set(plan.t_train) | set(plan.record_times))
        t_start = _time.perf_counter()
        for t in events:
            _steps_to(s, t, plan.dns.deltat)
            cfl = s.nonlinear(want_cfl=True) * plan.dns.deltat
            st = s.outstats()
            log(f"[les] DNS t = {s.time:.3f}: energy {st[0]:.5f}, dissipation {st[1]:.5f}, uv {st[2]:.5f}, "
                f"CFL {cfl:.3f}")
            if not np.all(np.isfinite(st)):
                raise hst.HstError(f"the DNS diverged by t = {s.time}")
            c.filter(s.time)
            if t in plan.t_train:
                n = c.sample(s.time)
                log(f"[les]   {n} samples")
            if t in plan.record_times:
                records[t] = c.velocity()
        log(f"[les] DNS to t = {s.time:.3f} in {_time.perf_counter() - t_start:.1f} s")
    finally:
        s.close()
    return c, records


def train(plan: Plan, c: closure.Closure, log=print):
    """Step 2."""
    npts = plan.les.ny * plan.les.nzd * 2 * plan.les.nxd
    n_train = npts * (len(plan.t_train) - plan.holdout)
    nm = c.normalise()
    if any(w != 1.0 for w in plan.target_weight):
        nm[24:30] = nm[24:30] / np.sqrt(np.asarray(plan.target_weight, dtype=np.float64))
        c.set_norm(nm)
    c.init_params(plan.seed)
    t0 = _time.perf_counter()
    loss, val = c.train(plan.epochs, plan.batch, plan.lr, seed=plan.seed, n_train=n_train)
    log(f"[les] {plan.epochs} epochs in {_time.perf_counter() - t0:.1f} s: loss {loss[0]:.4f} -> {loss[-1]:.4f}, "
        f"held out {val[0]:.4f} -> {val[-1]:.4f}")
    return loss, val


def coarse_run(plan: Plan, U0: np.ndarray, c: closure.Closure | None, log=print):
    """Step 3: one coarse run from U0 at t0; the velocity at the record times."""
    s = hst.Solver(plan.les, hst.FP64)
    out = {}
    diverged = None
    try:
        s.set_V(U0)
        s.reset_old()
        s.time = plan.t0
        if c is not None:
            c.attach(True)
        for t in plan.record_times:
            _steps_to(s, t, plan.les.deltat)
            st = s.outstats()
            if not np.all(np.isfinite(st)):
                diverged = t
                break
            out[t] = s.V
    finally:
        if c is not None:
            c.attach(False)
        s.close()
    if diverged is not None:
        log(f"[les] coarse run {'with' if c else 'without'} the closure diverged by t = {diverged}")
    return out, diverged


def statistics(plan: Plan, fields: dict) -> dict:
    """outstats of coarse fields, through a coarse state."""
    s = hst.Solver(plan.les, hst.FP64)
    try:
        res = {}
        for t, V in fields.items():
            s.set_V(V)
            s.time = t
            res[t] = s.outstats()
        return res
    finally:
        s.close()


def spectra(V: np.ndarray, deck: Deck):
    """Energy per |kx| and per |kz| summed over the other indices and the rows
    (the interior rows; ix > 0 counted twice for the implied conjugates)."""
    e = (np.abs(V[..., 2:-2]) ** 2).sum(axis=(0, 3)) / deck.ny  # [ix][iz]
    w = np.full(e.shape[0], 2.0)
    w[0] = 1.0
    e = e * w[:, None]
    ex = e.sum(axis=1)
    ez_full = e.sum(axis=0)
    nz = deck.nz
    ez = ez_full[nz:].copy()
    ez[1:] += ez_full[:nz][::-1]
    return ex, ez


def evaluate(plan: Plan, ref: dict, runs: dict) -> dict:
    """Step 4: errors of each coarse run against the filtered DNS."""
    sref = statistics(plan, ref)
    result = {"times": list(plan.record_times), "filtered_dns": {}, "runs": {}}
    names = {"energy": 0, "dissipation": 1, "uv": 2}
    result["filtered_dns"] = {k: [float(sref[t][i]) for t in plan.record_times] for k, i in names.items()}
    t_last = plan.record_times[-1]
    ex_ref, ez_ref = spectra(ref[t_last], plan.les)
    half = plan.les.nx // 2
    for name, (fields, diverged) in runs.items():
        r = {"diverged_at": diverged}
        if diverged is not None:
            result["runs"][name] = r
            continue
        st = statistics(plan, fields)
        for k, i in names.items():
            series = [float(st[t][i]) for t in plan.record_times]
            r[k] = series
            refs = result["filtered_dns"][k]
            r[k + "_error"] = float(np.mean([abs(a - b) / abs(b) for a, b in zip(series[1:], refs[1:])]))
        ex, ez = spectra(fields[t_last], plan.les)
        r["spectrum_x"] = ex.tolist()
        r["spectrum_z"] = ez.tolist()
        # mean |log| ratio over the upper half of the resolved wavenumbers
        r["spectrum_error"] = float(0.5 * (np.mean(np.abs(np.log(ex[half:] / ex_ref[half:]))) +
                                           np.mean(np.abs(np.log(ez[half:] / ez_ref[half:])))))
        result["runs"][name] = r
    result["filtered_dns"]["spectrum_x"] = ex_ref.tolist()
    result["filtered_dns"]["spectrum_z"] = ez_ref.tolist()
    return result


def run(plan: Plan, log=print) -> dict:
    c, records = run_dns(plan, log)
    try:
        loss, val = train(plan, c, log)
        U0 = records[plan.t0]
        plain = coarse_run(plan, U0, None, log)
        model = coarse_run(plan, U0, c, log)
    finally:
        c.close()
    res = evaluate(plan, records, {"plain": plain, "closure": model})
    res["train_loss"] = [float(v) for v in loss]
    res["heldout_loss"] = [float(v) for v in val]
    res["samples"] = int(plan.les.ny * plan.les.nzd * 2 * plan.les.nxd * len(plan.t_train))
    for k in ("energy_error", "dissipation_error", "uv_error", "spectrum_error"):
        a = res["runs"]["plain"].get(k)
        b = res["runs"]["closure"].get(k)
        if a is not None and b is not None:
            log(f"[les] {k}: plain {a:.4f}, closure {b:.4f}")
    return res


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=0)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    args = ap.parse_args()
    plan = default_plan(args.quick)
    if args.epochs:
        plan = dataclasses.replace(plan, epochs=args.epochs)
    plan = dataclasses.replace(plan, precision=hst.FP64 if args.precision == "fp64" else hst.MIXED)
    res = run(plan, log=lambda m: print(m, flush=True))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(res, fh)


if __name__ == "__main__":
    main()
