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


def format_e(values, d: int) -> np.ndarray:
    """Ew.d strings (w = d + 6) of an array of float64: mantissa 0.<d digits>
    rounded to nearest on the exact binary value, exponent 'E+dd' or, past
    two digits, '+ddd'; negative values drop the leading zero.  NaN and
    infinities are not produced by the time step and are not handled."""
    a = np.asarray(values, np.float64).ravel()
    sci = np.char.mod(f"%.{d - 1}e", np.abs(a))     # 'd.ddd...e+XX', correctly rounded
    out = np.empty(a.size, dtype=f"<U{d + 6}")
    for q, (txt, v) in enumerate(zip(sci.tolist(), a.tolist())):
        mant, ex = txt.split("e")
        digits = mant[0] + mant[2:d + 1]
        e = int(ex) + 1
        if v == 0.0:
            digits, e = "0" * d, 0
        expo = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else \
            f"{'+' if e >= 0 else '-'}{abs(e):03d}"
        neg = v < 0.0 or (v == 0.0 and np.signbit(v))
        out[q] = ("-." if neg else "0.") + digits + expo
    return out


def format_e10_4(values) -> np.ndarray:
    """E10.4 strings of an array of float64 (see format_e)."""
    return format_e(values, 4)


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


def _i4(v: int) -> str:
    """I4: right-justified in 4 columns, asterisks when it does not fit."""
    t = str(int(v))
    return t.rjust(4) if len(t) <= 4 else "****"


def parameters_text(cfg, d, procs=(1, 1, 1), haser=None) -> str:
    """save/parameters.dat of WriteWarnings (WriteWarnings.f90:130-256): one
    `name = value` record per quantity, reals in E12.6, integers in I4, and
    the logicals in the Fortran runtime's list-directed form (' name =  T').
    `d` is the derived constants (host.derive), `haser` the Haser arrays
    (Qn, vn, Ek, nu_i, nu_d; default: the Config's)."""
    h = haser or dict(Qn=cfg.Qn, vn=cfg.vn, Ek=cfg.Ek, nu_i=cfg.nu_i, nu_d=cfg.nu_d)
    lines = []

    def reals(name, vals):
        for s, t in enumerate(format_e(np.asarray(vals, np.float64)[:8], 6).tolist()):
            lines.append(f"{name}_{s + 1} = {t}")

    reals("n2p", d.n2p)
    reals("mass", d.mass)
    reals("charge", d.charge)
    reals("density", d.density)
    reals("kelvin", d.kelvin)
    reals("Qn", h["Qn"])
    reals("vn", h["vn"])
    reals("Ek", h["Ek"])
    reals("nui", h["nu_i"])
    reals("nud", h["nu_d"])
    for s in range(8):
        lines.append(f"ppc_{s + 1} = {_i4(d.ppc[s])}")
    for s in range(8):
        lines.append(f" gyration_{s + 1} =  {'T' if cfg.gyration[s] else 'F'}")
    for name, vals in (("v0", d.ve[:3]), ("B0", d.B0[:3]), ("dxyz", d.dxyz[:3])):
        for c, t in enumerate(format_e(vals, 6).tolist()):
            lines.append(f"{name}_{c + 1} = {t}")
    bounds = (cfg.gmin[0], cfg.gmax[0], cfg.gmin[1], cfg.gmax[1], cfg.gmin[2], cfg.gmax[2], d.dt)
    for name, t in zip(("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "dt"),
                       format_e(bounds, 6).tolist()):
        lines.append(f"{name} = {t}")
    for name, v in zip(("Nx", "Ny", "Nz", "iprocs", "jprocs", "kprocs"),
                       (cfg.nx, cfg.ny, cfg.nz) + tuple(procs)):
        lines.append(f"{name} = {_i4(v)}")
    for name, f in zip(("xflow", "yflow", "zflow"), cfg.flow):
        lines.append(f" {name} =  {'T' if f else 'F'}")
    return "".join(line + "\n" for line in lines)


def write_parameters(path, cfg, d, procs=(1, 1, 1), haser=None) -> str:
    """save/parameters.dat under `path` (WriteWarnings.f90:130-256, written
    once at setup, picard.f90:488-489); returns the file name."""
    os.makedirs(os.path.join(path, "save"), exist_ok=True)
    name = os.path.join(path, "save", "parameters.dat")
    with open(name, "w") as f:
        f.write(parameters_text(cfg, d, procs, haser))
    return name
