"""The hst.in deck: a Fortran namelist file with the groups &mesh, &physics,
&time_control, &init and (optional) &damping.

Mirrors read_input (reference src/hst_input.f90:14-103): the same defaults
(:26-42), the same checks (:59-90) and the same derived quantities
(:92-103, fft_fit :107-115).  Only the values the RK3 substep and its driver
use are kept; a deck that switches on a hook outside this library's scope
(Stokes layer, large-scale damping) is rejected with the reason.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import re

PI = 3.141592653589793


class hst_deck(ctypes.Structure):
    """The C struct of include/hst_gpu.h (hst_deck)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("linear", ctypes.c_int32), ("exact_shift", ctypes.c_int32),
        ("seed", ctypes.c_int32), ("tg_mode", ctypes.c_int32),
        ("initial_condition", ctypes.c_int32),
        ("alfa0", ctypes.c_double), ("beta0", ctypes.c_double), ("ly", ctypes.c_double),
        ("ystretch", ctypes.c_double),
        ("re", ctypes.c_double), ("S", ctypes.c_double), ("s2_mean", ctypes.c_double),
        ("s2_amplitude", ctypes.c_double), ("s2_period", ctypes.c_double),
        ("s2_start", ctypes.c_double),
        ("amplitude", ctypes.c_double), ("kpeak", ctypes.c_double),
        ("tg_amplitude", ctypes.c_double),
        ("sl_amplitude", ctypes.c_double),
        ("damp_kpar", ctypes.c_double), ("damp_kperp", ctypes.c_double),
        ("sl_period", ctypes.c_double), ("sl_delta", ctypes.c_double), ("sl_start", ctypes.c_double),
        ("damp_rate", ctypes.c_double), ("damp_edge", ctypes.c_double), ("damp_norm", ctypes.c_double),
        ("damp_angle", ctypes.c_double),
        ("sl_bodyforce", ctypes.c_int32), ("sl_ramp", ctypes.c_int32),
        ("damp_hard", ctypes.c_int32), ("damp_mean", ctypes.c_int32), ("damp_follow", ctypes.c_int32),
    ]


class hst_grid_info(ctypes.Structure):
    _fields_ = [
        ("nxd", ctypes.c_int32), ("nzd", ctypes.c_int32),
        ("nxm", ctypes.c_int32), ("nzm", ctypes.c_int32), ("nyg", ctypes.c_int32),
        ("ncol", ctypes.c_int32),
        ("lx", ctypes.c_double), ("lz", ctypes.c_double), ("dx", ctypes.c_double),
        ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("factor", ctypes.c_double),
        ("ni", ctypes.c_double),
    ]


def fft_fit(n: int) -> bool:
    """True when n is a power of two times at most one factor three (src/hst_input.f90:107-115)."""
    j = n
    while j % 2 == 0:
        j //= 2
    return j in (1, 3)


@dataclasses.dataclass
class Deck:
    # &mesh (defaults: src/hst_input.f90:26-27)
    nx: int = 63
    ny: int = 128
    nz: int = 63
    alfa0: float = 2.0 * PI / 3.0
    beta0: float = 2.0 * PI
    ly: float = 2.0
    ystretch: float = 0.0
    npy: int = 0
    # &physics (:28-30)
    re: float = 1000.0
    S: float = 1.0
    linear: bool = False
    exact_shift: bool = False
    s2_mean: float = 0.0
    s2_amplitude: float = 0.0
    s2_period: float = 0.0
    s2_start: float = 0.0
    sl_amplitude: float = 0.0
    sl_period: float = 1.0
    sl_delta: float = 0.02
    sl_start: float = 0.0
    sl_bodyforce: bool = True
    sl_ramp: bool = False
    # &time_control (:31-36)
    deltat: float = 0.0
    cflmax: float = 1.0
    t_max: float = 100.0
    nstep: int = 1000000
    wall_max: float = 0.0
    dt_stat: float = 0.01
    dt_field: float = 10.0
    dt_save: float = 10.0
    time: float = 0.0
    time_from_restart: bool = False
    restart_from_file: bool = True
    restart_file: str = "Dati.cart.out"
    timing: bool = False
    # &init (:37-41)
    initial_condition: str = "random"
    amplitude: float = 1.0e-3
    seed: int = 1
    kpeak: float = 4.0
    tg_amplitude: float = 1.0
    tg_mode: int = 1
    # &damping (:40-41)
    damp_kpar: float = 0.0
    damp_kperp: float = 0.0
    damp_rate: float = 0.0
    damp_hard: bool = False
    damp_edge: float = 0.0
    damp_norm: float = 2.0
    damp_mean: bool = False
    damp_follow: bool = True
    damp_angle: float = 0.0

    # ---------------------------------------------------------------- derived
    @property
    def ni(self) -> float:
        return 1.0 / self.re

    @property
    def lx(self) -> float:
        return 2.0 * PI / self.alfa0

    @property
    def lz(self) -> float:
        return 2.0 * PI / self.beta0

    @property
    def nxd(self) -> int:
        n = 3 * (self.nx + 1) // 2
        while not fft_fit(n):
            n += 1
        return n

    @property
    def nzd(self) -> int:
        n = 3 * self.nz
        while not fft_fit(n):
            n += 1
        return n

    @property
    def factor(self) -> float:
        return 1.0 / (2.0 * self.nxd * self.nzd)

    @property
    def dt_fixed(self) -> float:
        return self.deltat

    def to_c(self) -> hst_deck:
        d = hst_deck()
        for name, _ in hst_deck._fields_:
            if name == "initial_condition":
                d.initial_condition = 1 if self.initial_condition.strip() == "taylor_green" else 0
            else:
                v = getattr(self, name)
                setattr(d, name, int(v) if isinstance(v, bool) else v)
        return d

    def check(self) -> None:
        """The deck checks of read_input (src/hst_input.f90:59-90) that concern this path."""
        if self.sl_amplitude != 0.0 and (self.s2_amplitude != 0.0 or self.s2_mean != 0.0):
            raise ValueError("the Stokes layer and the spanwise shear S2 cannot both be active")
        if self.sl_amplitude != 0.0 and self.sl_period <= 0.0:
            raise ValueError("sl_period must be positive")
        if (self.damp_kpar > 0.0) != (self.damp_kperp > 0.0):
            raise ValueError("damp_kpar and damp_kperp must both be positive (damping on) or both zero (off)")
        if self.damp_norm < 1.0 or self.damp_edge < 0.0 or self.damp_edge >= 1.0 or self.damp_rate < 0.0:
            raise ValueError("damping needs damp_norm >= 1, 0 <= damp_edge < 1, damp_rate >= 0")
        if self.initial_condition.strip() not in ("random", "taylor_green"):
            raise ValueError("initial_condition must be 'random' or 'taylor_green'")
        if self.tg_mode < 1:
            raise ValueError("tg_mode must be positive")
        if self.ny < 8:
            raise ValueError("ny must be at least 8")
        if self.npy not in (0, 1):
            raise ValueError("one GPU holds every y row: npy must be 0 or 1")


_TOKEN = re.compile(
    r"""\s*(?:
        (?P<str>'[^']*'|"[^"]*")
      | (?P<eq>=)
      | (?P<comma>,)
      | (?P<slash>/)
      | (?P<word>[^\s=,/'"]+)
    )""",
    re.VERBOSE,
)


def _value(tok: str):
    t = tok.strip()
    low = t.lower()
    if low in (".true.", "t", ".t.", "true"):
        return True
    if low in (".false.", "f", ".f.", "false"):
        return False
    if t[0] in "'\"":
        return t[1:-1]
    num = low.replace("d", "e")
    try:
        if re.fullmatch(r"[+-]?\d+", num):
            return int(num)
        return float(num)
    except ValueError as exc:
        raise ValueError(f"cannot read namelist value {tok!r}") from exc


def parse_namelist(text: str) -> dict[str, dict[str, object]]:
    """Groups of a namelist file: {group: {key: value}}; '!' starts a comment."""
    lines = []
    for line in text.splitlines():
        out, quote = [], None
        for ch in line:
            if quote:
                out.append(ch)
                if ch == quote:
                    quote = None
            elif ch in "'\"":
                quote = ch
                out.append(ch)
            elif ch == "!":
                break
            else:
                out.append(ch)
        lines.append("".join(out))
    body = "\n".join(lines)
    groups: dict[str, dict[str, object]] = {}
    for m in re.finditer(r"&(\w+)(.*?)(?:/|&end)", body, re.DOTALL | re.IGNORECASE):
        name = m.group(1).lower()
        toks = [t for t in _TOKEN.finditer(m.group(2))]
        vals: dict[str, object] = {}
        key = None
        i = 0
        while i < len(toks):
            t = toks[i]
            if t.group("word") and i + 1 < len(toks) and toks[i + 1].group("eq"):
                key = t.group("word").lower()
                i += 2
                continue
            if t.group("comma"):
                i += 1
                continue
            if key is None:
                raise ValueError(f"namelist group &{name}: value without a key")
            raw = t.group("str") or t.group("word")
            vals[key] = _value(raw)
            i += 1
        groups[name] = vals
    return groups


_FIELDS = {f.name.lower(): f for f in dataclasses.fields(Deck)}


def read_input(path: str) -> Deck:
    """Reads a deck with the reference defaults (src/hst_input.f90:14-103)."""
    with open(path, "r", encoding="latin-1") as fh:
        return deck_from_text(fh.read())


def deck_from_text(text: str) -> Deck:
    groups = parse_namelist(text)
    for g in ("mesh", "physics", "time_control", "init"):
        if g not in groups:
            raise ValueError(f"the deck has no &{g} group")
    d = Deck()
    for g, vals in groups.items():
        for k, v in vals.items():
            f = _FIELDS.get(k)
            if f is None:
                continue  # keys of the reference this path does not use (line_chunk, transport, ...)
            if f.type in ("bool", bool):
                v = bool(v)
            elif f.type in ("int", int):
                v = int(v)
            elif f.type in ("float", float):
                v = float(v)
            setattr(d, f.name, v)
    d.check()
    return d


def write_deck(d: Deck, path: str, comment: str = "") -> None:
    """Writes a deck the reference's read_input accepts."""
    def fmt(v):
        if isinstance(v, bool):
            return ".true." if v else ".false."
        if isinstance(v, int):
            return str(v)
        if isinstance(v, float):
            return repr(v)
        return f"'{v}'"

    groups = {
        "mesh": ["nx", "ny", "nz", "alfa0", "beta0", "ly", "ystretch"],
        "physics": ["re", "S", "linear", "exact_shift", "s2_mean", "s2_amplitude", "s2_period", "s2_start",
                    "sl_amplitude", "sl_period", "sl_delta", "sl_start", "sl_bodyforce", "sl_ramp"],
        "time_control": ["deltat", "cflmax", "t_max", "nstep", "dt_stat", "dt_field", "dt_save", "time",
                         "time_from_restart", "restart_from_file"],
        "init": ["initial_condition", "amplitude", "seed", "kpeak"],
    }
    if d.damp_kpar > 0.0:
        groups["damping"] = ["damp_kpar", "damp_kperp", "damp_rate", "damp_hard", "damp_edge", "damp_norm",
                             "damp_mean", "damp_follow", "damp_angle"]
    with open(path, "w") as fh:
        for line in comment.splitlines():
            fh.write(f"! {line}\n")
        for g, keys in groups.items():
            items = ", ".join(f"{k} = {fmt(getattr(d, k))}" for k in keys)
            fh.write(f"&{g}  {items} /\n")


def new_timestep(deltat: float, cfl: float, cflmax: float, dt_fixed: float) -> float:
    """The next step length (new_timestep, src/hst.f90:163-169): cflmax/cfl
    when cflmax and the CFL number are positive, never above the deck's fixed
    step dt_fixed when that is set, and exactly dt_fixed without cflmax."""
    if cflmax > 0.0 and cfl > 0.0:
        deltat = cflmax / cfl
    if dt_fixed > 0.0:
        deltat = min(deltat, dt_fixed)
    if dt_fixed > 0.0 and cflmax <= 0.0:
        deltat = dt_fixed
    return deltat


def first_timestep(deltat: float) -> float:
    """The step before the first CFL number is known: deltat = 0 in the deck
    stands for 1 (src/hst.f90:85), which new_timestep then replaces."""
    return 1.0 if deltat == 0.0 else deltat


def crossed(time: float, deltat: float, dt: float) -> bool:
    """The interval test of the driver (src/hst.f90:154-158)."""
    if dt <= 0.0:
        return False
    return math.floor((time + 0.5 * deltat) / dt) > math.floor((time - 0.5 * deltat) / dt)
