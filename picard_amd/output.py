"""Text dumps of particles and fields in the reference's on-disk layout
(DumpParticles and DumpFields, Write.f90:2-523).

Every value is written in the Fortran edit descriptor E10.4: a mantissa
0.dddd (four significant digits, rounded to nearest with ties to even on the
exact binary value), then 'E+dd' when the decimal exponent has at most two
digits and '+ddd' otherwise; a negative value drops the leading zero
('-.dddd').  Particles are written seven values per record (positions,
velocities, species), fields one value per record in Fortran array order
(i fastest, then j, then k, guard layers included).

Files: save/particles/particles_IIIII_RRRRR and save/fields/<name>_IIIII_RRRRR
with IIIII the dump counter and RRRRR the MPI rank (0 on one GPU); field names
UE, Ex, Ey, Ez, N1..N<S>, X1..X<S>, Y1..Y<S>, Z1..Z<S>.
"""
from __future__ import annotations

import os

import numpy as np


def format_e10_4(values) -> np.ndarray:
    """E10.4 strings of an array of float64 (NaN and infinities are not
    produced by the time step and are not handled)."""
    a = np.asarray(values, np.float64).ravel()
    sci = np.char.mod("%.3e", np.abs(a))            # 'd.ddde+XX', correctly rounded
    out = np.empty(a.size, dtype="<U10")
    for q, (txt, v) in enumerate(zip(sci.tolist(), a.tolist())):
        mant, ex = txt.split("e")
        digits = mant[0] + mant[2:5]
        e = int(ex) + 1
        if v == 0.0:
            digits, e = "0000", 0
        expo = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else \
            f"{'+' if e >= 0 else '-'}{abs(e):03d}"
        neg = v < 0.0 or (v == 0.0 and np.signbit(v))
        out[q] = ("-." if neg else "0.") + digits + expo
    return out


def particles_text(rp, sp) -> str:
    """DumpParticles' records: real_particles(1:7, p) with slot 7 = species."""
    rp = np.asarray(rp, np.float64)
    blk = rp[:, :7].copy()
    blk[:, 6] = np.asarray(sp, np.float64)          # Write.f90:14
    s = format_e10_4(blk)
    lines = ["".join(s[7 * q:7 * q + 7]) for q in range(blk.shape[0])]
    return "".join(line + "\n" for line in lines)


def field_text(a) -> str:
    """One value per record, Fortran order of a C-order [Z, Y, X] view."""
    s = format_e10_4(np.asarray(a, np.float64))
    return "".join(v + "\n" for v in s.tolist())


def write_particles(path, rp, sp, iteration: int, rank: int = 0) -> str:
    """DumpParticles (Write.f90:2-31) under directory `path`; returns the file name."""
    d = os.path.join(path, "save", "particles")
    os.makedirs(d, exist_ok=True)
    name = os.path.join(d, f"particles_{iteration:05d}_{rank:05d}")
    with open(name, "w") as f:
        f.write(particles_text(rp, sp))
    return name


def write_fields(path, U, N, F, E, iteration: int, rank: int = 0) -> list:
    """DumpFields (Write.f90:34-523) under `path`.  U [Z,Y,X], E [Z,Y,X,3],
    N [Z,Y,X,S], F [Z,Y,X,3,S] (this package's views of the Fortran arrays).
    Returns the file names written."""
    d = os.path.join(path, "save", "fields")
    os.makedirs(d, exist_ok=True)
    S = N.shape[-1]
    items = [("UE", U), ("Ex", E[..., 0]), ("Ey", E[..., 1]), ("Ez", E[..., 2])]
    items += [(f"N{s + 1}", N[..., s]) for s in range(S)]
    for c, tag in enumerate("XYZ"):
        items += [(f"{tag}{s + 1}", F[..., c, s]) for s in range(S)]
    names = []
    for tag, arr in items:
        name = os.path.join(d, f"{tag}_{iteration:05d}_{rank:05d}")
        with open(name, "w") as f:
            f.write(field_text(arr))
        names.append(name)
    return names
