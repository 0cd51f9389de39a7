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


def nproc() -> int:
    """The host threads this process may use, as nproc(1) counts them: the CPU
    affinity mask, capped by OMP_NUM_THREADS and OMP_THREAD_LIMIT."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    for var in ("OMP_NUM_THREADS", "OMP_THREAD_LIMIT"):
        try:
            n = min(n, int(os.environ[var].split(",")[0]))
        except (KeyError, ValueError):
            pass
    return max(n, 1)


def cpu_step(deck, threads):
    """The FP64 CPU oracle on the host: warm-up = the start-up transform and
    CFL number of the seeded field (every array touched once), then one whole
    RK3 step (three substeps: all three rows of the RK3 table, the transforms,
    products, right-hand sides, shift and solves) timed with a monotonic clock.
    Returns (seconds per step, warm-up seconds)."""
    from oracle import oracle

    oracle.set_threads(threads)
    o = oracle.Oracle(deck)
    o.initial_field()
    o.fill_all_ghosts(0.0)
    t0 = time.perf_counter()
    oracle.compute_cfl(o)
    warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.timestep(deck.deltat)
    step = time.perf_counter() - t0
    del o
    return step, warm


def cpu_baseline(deck, threads, small_deck):
    """The CPU baseline, timed with the build's FP64 oracle after a warm-up
    transform: one whole step of the metric deck at nproc threads (the value),
    and one whole step of the
    nx = nz = 32, ny = 65 deck (the CPU-runnable correctness workload) at 1
    thread and at nproc threads: one thread on the metric deck (about 16 times
    the nproc time) or on nx = nz = 128, ny = 257 (over two minutes a step) does
    not fit the benchmark's time limit."""
    step, warm = cpu_step(deck, threads)
    log(f"CPU oracle, {threads} threads: {step:.1f} s per step (warm-up {warm:.1f} s)")
    s1, _ = cpu_step(small_deck, 1)
    log(f"CPU oracle, box32_y65, 1 thread: {s1:.3f} s per step")
    sn, _ = cpu_step(small_deck, threads)
    log(f"CPU oracle, box32_y65, {threads} threads: {sn:.3f} s per step")
    return {"value": 1.0 / step, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"one whole RK3 step of box256_y513_shift (nx=nz=256, ny=513, exact_shift) after a warm-up "
                      f"transform ({warm:.1f} s): {step:.1f} s with {threads} threads (nproc)",
            "box32_y65": {"workload": "nx=nz=32, ny=65, Re=1000, dt=0.005 (one whole RK3 step after a warm-up "
                                      "transform)",
                          "one_thread": {"value": round(1.0 / s1, 4), "unit": "steps/s", "s_per_step": round(s1, 4)},
                          "nproc": {"value": round(1.0 / sn, 4), "unit": "steps/s", "s_per_step": round(sn, 4),
                                    "cores": threads}}}


def copy_probe(gib):
    """The library's bandwidth probe on `gib` GiB buffers, one timed launch per
    grid (read-only, write-only, copy kernels): what FETCH_SIZE and WRITE_SIZE
    read for its copy kernel (gib GiB read, gib GiB written per launch) scales
    the counters of the solver's kernels to bytes."""
    print(json.dumps({"bandwidth_probe": hst.bandwidth_probe(gib, 1)}), flush=True)


def kernel_sources_sha() -> str:
    """sha256 of the HIP sources and headers of the library: the traffic
    figures of profiles/traffic.json hold for the build they were measured on."""
    import hashlib

    csrc = os.path.join(ROOT, "homogenenousshearturbulence_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(csrc)):
        if name.endswith((".hip", ".h")):
            with open(os.path.join(csrc, name), "rb") as fh:
                h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()


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
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run the bandwidth probe once on GIB GiB buffers (the HBM counter calibration) and exit")
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
    log("bandwidth probe (4 GiB streaming kernels)")
    probe = hst.bandwidth_probe(4.0, 5)
    log(f"HBM measured: read {probe['read_GBps']} GB/s, write {probe['write_GBps']} GB/s, copy {probe['copy_GBps']} GB/s")
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
    # (profiles/traffic_*.md), scaled by the copy calibration; null when the
    # kernels changed since (the sources' hash differs)
    traffic = None
    traffic_note = "no counter run for this workload"
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh)
        if t.get("workload") == os.path.basename(args.deck)[:-3] and t.get("form") == "mixed":
            if t.get("sources_sha256") == kernel_sources_sha():
                traffic = t.get("bytes_per_launch", {}).get(dom["kernel"])
                traffic_note = "rocprofv3 FETCH_SIZE/WRITE_SIZE of this build (profiles/traffic_*.md)"
            else:
                traffic_note = "the counter run in profiles/ is of other kernel sources: not used"
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
        threads = args.cpu_threads or nproc()
        log(f"CPU oracle baseline, {threads} threads")
        small = read_input(os.path.join(ROOT, "homogenenousshearturbulence_amd", "decks", "box32_y65.in"))
        cpu = cpu_baseline(deck, threads, small)
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
                     "traffic": traffic, "traffic_source": traffic_note,
                     "measured_ceiling": probe,
                     "frac_of_measured_copy": round(achieved / probe["copy_GBps"], 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
