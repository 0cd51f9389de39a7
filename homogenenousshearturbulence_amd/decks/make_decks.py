"""Writes the five decks as hst.in files.

Every deck: alfa0 = beta0 = 1, ly = 2 pi, S = 1, uniform y, fixed step
(cflmax = 0), seed 1, a fresh seeded field (restart_from_file = .false.),
no snapshots.  The seeded spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) is the square
of the reference's envelope (k/kpeak)^2 exp(1 - (k/kpeak)^2) with kpeak = k0;
`amplitude` is set so that the seeded field has <u_i u_i> = 3 (rms velocity 1
per component), from the exact energy of the seeded modes.

    python -m homogenenousshearturbulence_amd.decks.make_decks
"""
from __future__ import annotations

import math
import os

from homogenenousshearturbulence_amd.deck import Deck, write_deck
from homogenenousshearturbulence_amd import hst

HERE = os.path.dirname(os.path.abspath(__file__))
TWO_PI = 2.0 * math.pi

CONFIGS = {
    # name: (nx=nz, ny, re, kpeak, exact_shift, deltat, nstep, note)
    "box32_y65": (32, 65, 1000.0, 4.0, False, 0.005, 100, "correctness deck: FP64 vs the CPU oracle"),
    # Re 4000 is not resolved on 128 modes: the energy reaches the truncation
    # wavenumbers and the run blows up at t = 0.31 whatever the step (dt = 0.002,
    # 0.001 or the CFL-limited step); 150 steps of 0.002 (t = 0.3) stay finite
    "box128_y257": (128, 257, 4000.0, 8.0, False, 0.002, 150,
                    "FP64 on one GPU; under-resolved at Re 4000: finite to t = 0.3 (150 steps), blows up at t = 0.31"),
    "box256_y513_shift": (256, 513, 8000.0, 12.0, True, 0.001, 100, "the exact-shift deck: mixed vs FP64"),
    "box512_y1025": (512, 1025, 20000.0, 16.0, False, 5e-4, 20, "mixed precision at the largest one-GPU grid"),
    "les32_y65": (32, 65, 1000.0, 8.0, False, 0.002, 400,
                  "coarse run of the learned closure: the grid of box128_y257 filtered by 4, at Re 1000"),
}


def make(name: str) -> Deck:
    n, ny, re, kp, exact, dt, nstep, _ = CONFIGS[name]
    d = Deck(nx=n, ny=ny, nz=n, alfa0=1.0, beta0=1.0, ly=TWO_PI, ystretch=0.0, re=re, S=1.0, linear=False,
             exact_shift=exact, deltat=dt, cflmax=0.0, t_max=1.0e9, nstep=nstep, dt_stat=dt, dt_field=1.0e9,
             dt_save=1.0e9, time=0.0, restart_from_file=False, initial_condition="random", amplitude=1.0, seed=1,
             kpeak=kp)
    q2 = hst.initial_energy(d)
    d.amplitude = math.sqrt(3.0 / q2)
    return d


def main() -> None:
    for name, cfg in CONFIGS.items():
        d = make(name)
        write_deck(d, os.path.join(HERE, name + ".in"),
                   comment=f"{name}: {cfg[7]}.\nnx = nz = {cfg[0]}, ny = {cfg[1]}, Re = {cfg[2]}, kpeak = {cfg[3]}, "
                           f"{cfg[6]} steps of {cfg[5]}; amplitude gives <u_i u_i> = 3.\n"
                           "Written by make_decks.py.")
        print(name, d.amplitude)


if __name__ == "__main__":
    main()
