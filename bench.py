"""RK3 steps/s of the hst substep on one MI355X at nx = nz = 256, ny = 513
(the deck homogenenousshearturbulence_amd/decks/box256_y513_shift.in: Re 8000,
exact_shift, dt = 0.001): the mixed-precision form (FP32 transforms, products
and derivative right-hand sides; FP64 solves, Crank-Nicolson operators, mean
mode and phase) against the FP64 form, the relative L2 error of the mixed
fields after the deck's 100 steps, the HBM roofline of the dominant kernel,
and the FP64 CPU oracle timed on the host cores beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--accuracy-steps N]

Prints one JSON line.  Inputs are resident in HBM before the timed region;
the timed region is K whole RK3 steps bracketed by device synchronisation.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from homogenenousshearturbulence_amd import hst  # noqa: E402
from homogenenousshearturbulence_amd.deck import read_input  # noqa: E402

METRIC = "RK3 steps/s, nx=nz=256 ny=513: mixed vs FP64 at stated L2 error; % HBM peak"
HBM_PEAK_GBS = 8000.0
DECK = os.path.join(ROOT, "homogenenousshearturbulence_amd", "decks", "box256_y513_shift.in")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_form(deck, precision, V0, warmup, steps, acc_steps, prof_steps=2, chunk=0):
    """Times `steps` RK3 steps after `warmup`, runs on to `acc_steps` steps in
    total for the accuracy statement, then times `prof_steps` more per kernel."""
    s = hst.Solver(deck, precision, chunk)
    try:
        s.set_V(V0)
        s.zero(hst.OLD1)
        s.zero(hst.OLD2)
        s.time = 0.0
        s.fill_all_ghosts(0.0)
        dt = deck.deltat
        for _ in range(warmup):
            s.timestep(dt)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.timestep(dt)
        s.sync()
        elapsed = time.perf_counter() - t0
        done = warmup + steps
        while done < acc_steps:
            s.timestep(dt)
            done += 1
        s.sync()
        stats = s.outstats()
        V = s.V if acc_steps > 0 else None
        s.profile(True)
        for _ in range(prof_steps):
            s.timestep(dt)
        s.sync()
        table = s.profile_table()
        s.profile(False)
        mem, chunk = s.memory()
        return {"elapsed": elapsed, "V": V, "stats": stats, "table": table, "steps_done": done, "mem": mem,
                "chunk": chunk}
    finally:
        s.close()


def cpu_baseline(deck, threads):
    """The FP64 CPU oracle on the host: one full substep (transforms, products,
    right-hand sides, exact shift, solves) of the same deck, scaled to steps/s."""
    from oracle import oracle

    oracle.set_threads(threads)
    o = oracle.Oracle(deck)
    o.initial_field()
    o.fill_all_ghosts(0.0)
    import ctypes

    L = oracle.lib()
    ode = (ctypes.c_double * 3)(*oracle.RK_RAI[0])
    dt = deck.deltat
    t0 = time.perf_counter()
    o.nonlinear(mode=0)
    L.hsto_buildrhs_prepare(o.h, ode, dt, 0.0)
    L.hsto_buildrhs(o.h, ode, 1)
    L.hsto_buildrhs(o.h, ode, 2)
    dts = 2.0 / oracle.RK_RAI[0][0] * dt
    L.hsto_shear_shift(o.h, dts, 0.0)
    L.hsto_linsolve(o.h, oracle.RK_RAI[0][0] / dt, dts)
    sub = time.perf_counter() - t0
    del o
    return {"value": 1.0 / (3.0 * sub), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"one RK3 substep of box256_y513_shift (nx=nz=256, ny=513) timed in {sub:.1f} s with {threads} threads, "
                      "steps/s = 1/(3 x substep time)"}


def copy_probe(mib):
    """Five launches of a plain elementwise copy (read n bytes, write n bytes):
    what FETCH_SIZE and WRITE_SIZE read for it scales the counters of the
    solver's kernels to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0", dtype=torch.float32)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 5
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 5,
                      "ms_per_launch": ms, "GBps": 8 * n / (ms * 1e-3) / 1e9}), flush=True)


def step_bytes(table, steps):
    return sum(r["bytes"] for r in table) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--accuracy-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--deck", default=DECK)
    ap.add_argument("--chunk", type=int, default=0, help="rows per transform chunk (0: all rows)")
    ap.add_argument("--precision", choices=("both", "mixed", "fp64"), default="both")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run five plain device copies of MIB MiB (the HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this library has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    deck = read_input(args.deck)
    log(f"deck {args.deck}: nx={deck.nx} ny={deck.ny} nz={deck.nz} nxd={deck.nxd} nzd={deck.nzd}")
    t0 = time.perf_counter()
    V0 = hst.initial_field(deck)
    log(f"seeded field in {time.perf_counter() - t0:.1f} s")
    acc = max(args.accuracy_steps, args.warmup + args.steps)
    if args.precision != "both":
        # a single form, for sweeps: no accuracy statement
        p = hst.MIXED if args.precision == "mixed" else hst.FP64
        r = run_form(deck, p, V0, args.warmup, args.steps, 0, chunk=args.chunk)
        tab = r["table"]
        print(json.dumps({"precision": args.precision, "chunk": r["chunk"], "steps_per_s": args.steps / r["elapsed"],
                          "kernels": {x["kernel"]: round(x["ms"] / x["launches"], 4) for x in tab},
                          "kernel_ms_per_step": {x["kernel"]: round(x["ms"] / 2, 3) for x in tab}}), flush=True)
        return
    log("FP64 form")
    r64 = run_form(deck, hst.FP64, V0, args.warmup, args.steps, acc, chunk=args.chunk)
    log(f"FP64: {args.steps / r64['elapsed']:.3f} steps/s")
    log("mixed form")
    rmx = run_form(deck, hst.MIXED, V0, args.warmup, args.steps, acc, chunk=args.chunk)
    log(f"mixed: {args.steps / rmx['elapsed']:.3f} steps/s")
    del V0
    # accuracy of the mixed form against the FP64 form after `acc` steps
    l2 = {}
    for c, name in enumerate(("u", "v", "w")):
        a, b = rmx["V"][c][..., 2:-2], r64["V"][c][..., 2:-2]
        l2[name] = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    names = ("energy", "diss", "uv", "vw", "uu", "vv", "ww", "uw")
    for k, n in enumerate(names):
        l2[n] = float(abs(rmx["stats"][k] - r64["stats"][k]) / abs(r64["stats"][k]))
    rmx["V"] = r64["V"] = None

    sps_mx = args.steps / rmx["elapsed"]
    sps_64 = args.steps / r64["elapsed"]
    ms_step = 1000.0 * rmx["elapsed"] / args.steps
    # roofline of the dominant kernel of the mixed step (HIP events on the library's stream)
    tab = rmx["table"]
    dom = max(tab, key=lambda r: r["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    per_launch = dom["bytes"] / dom["launches"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch of that kernel from the FETCH_SIZE / WRITE_SIZE runs
    # (profiles/traffic_box256_y513_shift_mixed.md), scaled by the copy calibration
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh)
        if t.get("workload") == os.path.basename(args.deck)[:-3] and t.get("form") == "mixed":
            traffic = t.get("bytes_per_launch", {}).get(dom["kernel"])
    bytes_mx = step_bytes(rmx["table"], 2)
    bytes_64 = step_bytes(r64["table"], 2)
    kernels = {r["kernel"]: {"ms_per_launch": round(r["ms"] / r["launches"], 4),
                             "GBps": round(r["bytes"] / r["launches"] / (r["ms"] / r["launches"] * 1e-3) / 1e9, 1),
                             "share": round(r["ms"] / sum(x["ms"] for x in tab), 3)} for r in tab}
    k64 = r64["table"]
    kernels64 = {r["kernel"]: {"ms_per_launch": round(r["ms"] / r["launches"], 4),
                               "GBps": round(r["bytes"] / r["launches"] / (r["ms"] / r["launches"] * 1e-3) / 1e9, 1)}
                 for r in k64}
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU oracle baseline, {threads} threads")
        cpu = cpu_baseline(deck, threads)
        log(f"CPU: {cpu['value']:.4f} steps/s")
    out = {
        "metric": METRIC,
        "value": round(sps_mx, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: seeded divergence-free random field, E(k) ~ k^4 exp(-2 (k/12)^2), u_rms = 1",
        "config": {
            "workload": "box256_y513_shift: nx=nz=256 (257 x 513 Fourier columns), ny=513, nxd=512 (1024 x points), nzd=768, "
                        "Re=8000, S=1, exact_shift, dt=0.001; mixed = FP32 transforms/products/derivative RHS, "
                        "FP64 solves/CN/mean mode/phase",
            "fp64_steps_per_s": round(sps_64, 4),
            "fp64_ms_per_step": round(1000.0 / sps_64, 3),
            "mixed_over_fp64": round(sps_mx / sps_64, 3),
            "rel_l2_mixed_vs_fp64": {k: float(f"{v:.3e}") for k, v in l2.items()},
            "accuracy_steps": rmx["steps_done"],
            "step_hbm_fraction_mixed": round(bytes_mx / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "step_hbm_fraction_fp64": round(bytes_64 / (1.0 / sps_64) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_GB_per_step": {"mixed": round(bytes_mx / 1e9, 2), "fp64": round(bytes_64 / 1e9, 2)},
            "kernels_mixed": kernels,
            "kernels_fp64": kernels64,
            "device_GB": round(rmx["mem"] / 1e9, 1),
            "transform_chunk_rows": rmx["chunk"],
        },
        "roofline": {"kernel": dom["kernel"], "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
