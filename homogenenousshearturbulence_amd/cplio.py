"""Velocity field files in the CPL layout (Dati.cart.out, fields/field<n>.fld).

Mirrors the file format of the reference's hst_io (src/hst_io.f90:1-33):
a text header as io.cpl writes it, five raw doubles (time, S, S2, gamma_x,
gamma_y), the line "Vfield=", then the array (0..nx, -nz..nz, -2..ny+1) of
(u, w, v) complex128, C order with the components innermost.  The four ghost
rows are written as plain periodic copies of the interior rows
(write_cpl, src/hst_io.f90:296-310).

Fields here use the reference's in-memory layout V(iy, iz, ix, c) as a numpy
array of shape (3, nx+1, 2nz+1, ny+4) (component, x mode, z mode, row).
"""
from __future__ import annotations

import struct

import numpy as np

# our component behind CPL's (u, v, w): CPL_ORDER (src/hst_io.f90:47)
CPL_ORDER = (0, 2, 1)


def _g17(x: float) -> str:
    """Fortran's G0.17 edit of a double, as str_r writes it (src/hst_io.f90:364-370)."""
    if x == 0.0:
        return "0.0000000000000000"
    ax = abs(x)
    if 0.1 <= ax < 1.0e17:
        # F format with 17 significant digits
        e = int(np.floor(np.log10(ax))) + 1
        d = max(17 - e, 0)
        s = f"{x:.{d}f}"
        return s
    s = f"{x:.16E}"  # mantissa d.dddd... with 17 significant digits
    mant, exp = s.split("E")
    m = mant.replace(".", "")
    sign = ""
    if m.startswith("-"):
        sign, m = "-", m[1:]
    e = int(exp) + 1
    return f"{sign}0.{m}E{e:+d}"


def header(nx: int, ny: int, nz: int, alfa0: float, beta0: float, ystretch: float, re: float,
           deltat: float, t_max: float, dt_field: float, dt_save: float, time: float,
           S: float, s2: float, gamma_y: float) -> bytes:
    """The CPL header of restart_write (src/hst_io.f90:246-257)."""
    TAB, LF = "\t", "\n"
    ht = ystretch if ystretch > 0.0 else -1.0
    text = (f"nx={nx} {TAB}ny={nz} {TAB}nz={ny + 1} {TAB}alpha0={_g17(alfa0)} {TAB}beta0={_g17(beta0)} "
            f"{TAB}htcoeff={_g17(ht)} {TAB}Re={_g17(re)} {TAB}Pr=0.71{LF}"
            f"deltat={_g17(deltat)} {TAB}t_max={_g17(t_max)} {TAB}dt_field={_g17(dt_field)} "
            f"{TAB}dt_save={_g17(dt_save)}{LF}"
            f"t_field={_g17(time)}{LF}"
            f"meanpx=0 {TAB}meanflowx=0 {TAB}meanpy=0 {TAB}meanflowy=0{LF}")
    b = text.encode("latin-1")
    for key, val in (("time", time), ("S", S), ("S2", s2), ("gamma_x", S * time), ("gamma_y", gamma_y)):
        b += f"{key}={LF}".encode() + struct.pack("<d", val) + LF.encode()
    return b + b"Vfield=\n"


def write_field(path: str, V: np.ndarray, head: bytes) -> None:
    """restart_write + write_cpl with one rank (src/hst_io.f90:240-310).

    V: complex128 (3, nx+1, 2nz+1, ny+4) in the reference layout; only the
    interior rows are taken, the file's ghost rows are periodic copies."""
    ncomp, nxm, nzm, nyg = V.shape
    ny = nyg - 4
    buf = np.empty((nxm, nzm, nyg, ncomp), dtype=np.complex128)
    inner = np.stack([V[c][:, :, 2:ny + 2] for c in (CPL_ORDER if ncomp == 3 else (0,))], axis=-1)
    buf[:, :, 2:ny + 2, :] = inner
    buf[:, :, ny + 2:ny + 4, :] = inner[:, :, 0:2, :]
    buf[:, :, 0:2, :] = inner[:, :, ny - 2:ny, :]
    with open(path, "wb") as fh:
        fh.write(head)
        fh.write(buf.tobytes(order="C"))


def read_field(path: str):
    """Reads a velocity file: returns (time, V, sizes) with V in the reference
    layout (3, nx+1, 2nz+1, ny+4), interior rows and the file's ghost rows as
    stored (restart_read, src/hst_io.f90:69-178, same mesh; the reader of
    tests/compare_fields.py:30-46)."""
    b = open(path, "rb").read()
    i = b.find(b"Vfield=\n")
    if i < 0:
        raise ValueError(f"{path}: no Vfield= line (not a CPL field file)")
    head = b[:i].decode("latin-1")
    sizes = {}
    for key in ("nx", "ny", "nz"):
        j = head.find(key + "=")
        sizes[key] = int(head[j + len(key) + 1:].split()[0])
    t = struct.unpack("<d", b[b.find(b"time=\n") + 6:][:8])[0]
    nx, nzc, nyc = sizes["nx"], sizes["ny"], sizes["nz"]
    a = np.frombuffer(b[i + 8:], np.complex128).reshape((nx + 1, 2 * nzc + 1, nyc + 3, 3))
    V = np.empty((3, nx + 1, 2 * nzc + 1, nyc + 3), dtype=np.complex128)
    for k, c in enumerate(CPL_ORDER):
        V[c] = a[..., k]
    return t, V, {"nx": nx, "nz": nzc, "ny": nyc - 1}


def compare_fields(A: np.ndarray, B: np.ndarray) -> float:
    """max|A-B| / max|A| over the interior rows (tests/compare_fields.py:55-59)."""
    d = np.abs(A[..., 2:-2] - B[..., 2:-2]).max()
    m = np.abs(A).max()
    return float(d / m)


def energy(V: np.ndarray) -> float:
    """<u_i u_i> as compare_fields.py computes it (tests/compare_fields.py:49-52)."""
    w = np.where(np.arange(V.shape[1]) == 0, 1.0, 2.0)[None, :, None, None]
    return float((np.abs(V[..., 2:-2]) ** 2 * w).sum() / (V.shape[-1] - 4))
