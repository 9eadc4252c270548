"""Host side of the PIC time step: run configuration, the derived constants of
picard.f90, the host particle sources, and a device-resident Domain that calls
the C ABI in the reference's main-loop order.

Layouts match the reference (see include/picard_hip.h): particles are
``float64[n, 8]`` + ``int32[n]`` (real_particles(8,n), int_particles(n)); grids
are C-order ``[Nz+2, Ny+2, Nx+2(, comp)]`` views of the Fortran column-major
arrays."""
from __future__ import annotations

import ctypes as ct
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib

# physical constants, picard.f90:56-66
TAU = 6.2831853071795864769252867665590
EC = 1.60217646e-19
MP = 1.6726216e-27
ME = 9.1093819e-31
EPS0 = 8.854187817e-12
KB = 1.380650e-23
MU0 = 1.25663706143591729538505735331180e-6


@dataclass
class Config:
    """Run parameters: picard.f90 sections 1.1, 2.1.0 and 2.1.1 as data."""
    nx: int = 32
    ny: int = 32
    nz: int = 32
    species: int = 2
    flow: tuple = (True, False, False)
    gmin: tuple = (-160.0, -160.0, -160.0)
    gmax: tuple = (160.0, 160.0, 160.0)
    dt: float | None = None          # None: dt_wpe / omega_pe
    dt_wpe: float = 0.1
    maxerr: float = 1e-6
    maxit: int = 2 ** 20
    mass: tuple = (ME, MP, 18 * MP, 4 * MP, 44 * MP, MP, MP, MP)
    charge3: tuple = (EC,) * 6       # charge(3:8), picard.f90:202-207
    density: tuple = (0.0, 5.0e6, 0, 0, 0, 0, 0, 0)   # density(2:8) used
    ppc: tuple = (0, 8, 0, 0, 0, 0, 0, 0)             # ppc(2:8) used
    kelvin: tuple = (116045.0, 116045.0, 116045.0, 0, 0, 0, 0, 0)
    v0: tuple = (4.0e5, 0.0, 0.0)    # bulk velocity, every species (picard.f90:238-248)
    B0: tuple = (0.0, 0.0, 5.0e-9)
    gyration: tuple = (True,) * 8
    fluid: tuple = (False,) * 8
    max_particles: int = 0           # 0: sized from the plasma
    seed: int = 20261203
    # Haser production (picard.f90:258-302), per species s at index s-1;
    # rate_scale multiplies ParticlesProduce.f90:70's expected count
    Qn: tuple = (0.0,) * 8
    vn: tuple = (0.0,) * 8
    nu_i: tuple = (0.0,) * 8
    nu_d: tuple = (0.0,) * 8
    Ek: tuple = (0.0,) * 8
    rate_scale: float = 1.0


def config1(**kw) -> Config:
    """32^3, dx = 10 m, H+ and e- at 8 per cell, 400 km/s drift, 10 eV, 5 nT, x in/outflow."""
    return Config(**kw)


# Haser source of the 256^3 workload: Q = 1e26 s^-1, vn = 1 km/s and an
# ionisation/loss scale of 1e6 m (nu_i = nu_d = vn / 1e6 m) as stated, counted
# per step (rate_scale = dt).  At the reference's macroparticle weight for a
# source-only species (n2p(3) = n2p(1), picard.f90:394-401) that is 7.9e8 ions
# per step on the 2560 m box, more than half the plasma load, so the rate is
# scaled to 3.3e22 s^-1: about 2^18 heavy ions (and as many electrons) per step.
HASER_Q_STATED = 1.0e26
HASER_Q_USED = 3.3e22


def config2(n: int = 256, **kw) -> Config:
    """n^3 cells of 10 m, H+ and e- at 32 per cell (1.07e9 particles at 256^3), the
    drifting Maxwellian and B of the 32^3 case, x in/outflow, and a Haser source
    of 18 amu ions (+e, one electron each) at the domain centre."""
    half = 5.0 * n
    vn = 1.0e3
    nu = vn / 1.0e6
    base = dict(nx=n, ny=n, nz=n, species=3, gmin=(-half,) * 3, gmax=(half,) * 3,
                ppc=(0, 32, 0, 0, 0, 0, 0, 0),
                Qn=(0.0, 0.0, HASER_Q_USED, 0, 0, 0, 0, 0), vn=(0.0, 0.0, vn, 0, 0, 0, 0, 0),
                nu_i=(0.0, 0.0, nu, 0, 0, 0, 0, 0), nu_d=(0.0, 0.0, nu, 0, 0, 0, 0, 0),
                Ek=(0.0, 0.0, 15.2 * EC, 0, 0, 0, 0, 0))
    base.update(kw)
    cfg = Config(**base)
    cfg.rate_scale = derive(cfg).dt
    return cfg


def config3(n: int = 512, **kw) -> Config:
    """The mixed-precision Poisson case: n^3 cells of 10 m, every boundary
    open (the BCpotentials face potential on x, y and z), solved to relative
    residual 1e-10; the charge is poisson_charge()."""
    half = 5.0 * n
    base = dict(nx=n, ny=n, nz=n, flow=(True, True, True), gmin=(-half,) * 3, gmax=(half,) * 3,
                maxerr=1.0e-10)
    base.update(kw)
    return Config(**base)


def poisson_charge(cfg: Config, seed: int = 20261205, blobs: int = 16,
                   Q: float = HASER_Q_STATED, vn: float = 1.0e3, L: float = 1.0e6) -> np.ndarray:
    """Charge per cell [C] of the Poisson case, the array CalcEpotential reads
    (b = -C / (dV eps0), CalcEpotential.f90:50), [Z+2, Y+2, X+2] with zero
    guards.  `blobs` Gaussian blobs (centres uniform in the interior, sigma
    uniform in 4-20 cells, amplitude uniform in +-1e-3 e n of the plasma
    density n) plus e times the Haser ion density Q / (4 pi vn r^2) exp(-r/L)
    of a source at the origin, r taken as ParticlesProduce.f90:46-58 takes it:
    the cell centre's distance, or, where (4 r)^3 < dV, the mean distance of
    the 8 points at the quarter offsets of the cell."""
    d = derive(cfg)
    rng = np.random.default_rng(seed)
    dx, dy, dz = (float(v) for v in d.dxyz[:3])
    x0 = cfg.gmin[0] + dx * np.arange(cfg.nx)          # lower cell edges
    y0 = cfg.gmin[1] + dy * np.arange(cfg.ny)
    z0 = cfg.gmin[2] + dz * np.arange(cfg.nz)
    Z, Y, X = np.meshgrid(z0, y0, x0, indexing="ij")
    xc, yc, zc = X + 0.5 * dx, Y + 0.5 * dy, Z + 0.5 * dz
    amp = 1.0e-3 * EC * float(cfg.density[1])
    rho = np.zeros_like(xc)
    for _ in range(blobs):
        cx = cfg.gmin[0] + dx * rng.uniform(0.0, cfg.nx)
        cy = cfg.gmin[1] + dy * rng.uniform(0.0, cfg.ny)
        cz = cfg.gmin[2] + dz * rng.uniform(0.0, cfg.nz)
        sig = rng.uniform(4.0, 20.0) * dx
        a = rng.uniform(-amp, amp)
        rho += a * np.exp(-((xc - cx) ** 2 + (yc - cy) ** 2 + (zc - cz) ** 2) / (2.0 * sig * sig))
    r2 = xc * xc + yc * yc + zc * zc
    r1 = np.sqrt(r2)
    near = (4.0 * r1) ** 3 < d.dV
    if near.any():
        acc = np.zeros(int(near.sum()))
        for fx, fy, fz in ((0.25, 0.25, 0.25), (0.25, 0.25, 0.75), (0.25, 0.75, 0.25),
                           (0.75, 0.25, 0.25), (0.25, 0.75, 0.75), (0.75, 0.25, 0.75),
                           (0.75, 0.75, 0.25), (0.75, 0.75, 0.75)):        # the reference's order
            acc = acc + np.sqrt((X[near] + fx * dx) ** 2 + (Y[near] + fy * dy) ** 2 +
                                (Z[near] + fz * dz) ** 2)
        r1[near] = acc * 0.125
        r2[near] = r1[near] ** 2
    rho += EC * Q / (2.0 * TAU * r2 * vn) * np.exp(-r1 / L)
    C = np.zeros((cfg.nz + 2, cfg.ny + 2, cfg.nx + 2))
    C[1:-1, 1:-1, 1:-1] = rho * d.dV
    return C


@dataclass
class Derived:
    dxyz: np.ndarray
    dV: float
    charge: np.ndarray
    density: np.ndarray
    ppc: np.ndarray
    ipp: np.ndarray
    epp: np.ndarray
    hpp: np.ndarray
    n2p: np.ndarray
    eta: np.ndarray
    ve: np.ndarray
    v0: np.ndarray           # [8, 4]
    B0: np.ndarray           # [4]
    E0: np.ndarray           # [4]
    tB: np.ndarray           # [8, 4]
    sB: np.ndarray           # [8, 4]
    mass: np.ndarray
    kelvin: np.ndarray
    dt: float
    omega_pe: float


def _fsum(vals):
    """Fortran SUM: sequential from 0."""
    s = 0.0
    for v in vals:
        s = s + float(v)
    return s


def derive(cfg: Config) -> Derived:
    """picard.f90:333-466 for one rank."""
    n = (cfg.nx, cfg.ny, cfg.nz)
    L = [cfg.gmax[c] - cfg.gmin[c] for c in range(3)]
    dxyz = np.array([L[c] / float(n[c]) for c in range(3)])       # :344-360 (iprocs = 1)
    dV = dxyz[0] * dxyz[1] * dxyz[2]
    mass = np.array(cfg.mass, np.float64)
    charge = np.zeros(8)
    charge[2:] = cfg.charge3
    charge[0] = -EC                                              # :371
    charge[1] = EC
    ppc = np.array(cfg.ppc, np.int64)
    density = np.array(cfg.density, np.float64)
    for s in range(1, 8):                                        # :373
        density[s] = density[s] * float(ppc[s]) / float(max(ppc[s], 1))
    density[0] = _fsum(-density[s] * charge[s] for s in range(1, 8)) * (1.0 / charge[0])
    for s in range(8):                                           # :375-380
        if density[s] <= 0.0 or cfg.species < s + 1:
            ppc[s] = 0
            density[s] = 0.0
    density[0] = _fsum(-density[s] * charge[s] / charge[0] for s in range(1, 8))   # :381
    epp = np.array([max(int(0.5 - charge[s] / charge[0]), 0) for s in range(8)])   # :384
    hpp = np.array([max(int(0.5 - charge[s] / charge[1]), 0) for s in range(8)])   # :385
    hpp[0] = 0
    ipp = np.zeros(8, np.int64)
    ipp[2:] = ppc[2:]                                            # :388
    ipp[1] = ppc[1] - int(np.sum(ipp[2:] * hpp[2:]))
    ipp[0] = 0
    ppc[0] = int(np.sum(ipp[1:] * epp[1:]))                      # :392
    n2p = np.zeros(8)
    for s in range(8):                                           # :394-401
        if ppc[s] > 0:
            n2p[s] = dV * density[s] * (1.0 / float(ppc[s]))
        else:
            n2p[s] = n2p[0]
    eta = charge / mass                                          # :405
    w = [float(ipp[s]) * (n2p[s] * mass[s] + float(epp[s]) * n2p[0] * mass[0] +
                          float(hpp[s]) * n2p[1] * mass[1]) for s in range(8)]
    v0 = np.zeros((8, 4))
    v0[:, :3] = np.asarray(cfg.v0, np.float64)[None, :]
    ve = np.zeros(4)
    den = _fsum(w)
    for c in range(3):                                           # :408-413
        ve[c] = _fsum(w[s] * v0[s, c] for s in range(8)) / den if den != 0.0 else 0.0
    for s in range(8):
        v0[s, 3] = math.sqrt(_fsum(v0[s, c] ** 2 for c in range(3)))
    ve[3] = math.sqrt(_fsum(ve[c] ** 2 for c in range(3)))
    B0 = np.zeros(4)
    B0[:3] = cfg.B0
    B0[3] = math.sqrt(_fsum(B0[c] ** 2 for c in range(3)))
    E0 = np.zeros(4)                                             # :424-427
    E0[0] = ve[2] * B0[1] - ve[1] * B0[2]
    E0[1] = ve[0] * B0[2] - ve[2] * B0[0]
    E0[2] = ve[1] * B0[0] - ve[0] * B0[1]
    E0[3] = math.sqrt(_fsum(E0[c] ** 2 for c in range(3)))
    omega_pe = math.sqrt(density[0] * EC * EC / (EPS0 * mass[0])) if density[0] > 0 else 0.0
    dt = cfg.dt if cfg.dt is not None else cfg.dt_wpe / omega_pe
    tB = np.zeros((8, 4))
    for s in range(8):                                           # :452-461
        if cfg.gyration[s]:
            for c in range(4):
                tB[s, c] = 0.5 * dt * B0[c] * eta[s]
    sB = np.zeros((8, 4))
    for s in range(8):                                           # :463-466
        for c in range(4):
            sB[s, c] = 2.0 * tB[s, c] / (1.0 + tB[s, 3] ** 2)
    return Derived(dxyz=dxyz, dV=dV, charge=charge, density=density, ppc=ppc, ipp=ipp, epp=epp,
                   hpp=hpp, n2p=n2p, eta=eta, ve=ve, v0=v0, B0=B0, E0=E0, tB=tB, sB=sB,
                   mass=mass, kelvin=np.array(cfg.kelvin, np.float64), dt=dt, omega_pe=omega_pe)


# ---------------------------------------------------------------------------
# host particle sources (ParticlesInitialize.f90, BCinflow.f90) with an
# explicitly seeded generator in place of the Fortran runtime's random_number

def _normal(rng, m):
    """Box-Muller as Tools.f90:30-48: sqrt(-2 ln r2) sin(tau r1), r2 > 0."""
    r1 = rng.random(m)
    r2 = rng.random(m)
    bad = r2 <= 0.0
    while np.any(bad):
        r2[bad] = rng.random(int(bad.sum()))
        bad = r2 <= 0.0
    return np.sqrt(-2.0 * np.log(r2)) * np.sin(TAU * r1)


def _maxwellian(rng, m, v0, T, mass):
    """Gaussian(u, v, t, m), Tools.f90:2-28."""
    if T > 0.0 and mass > 0.0:
        a = mass / (2.0 * KB * T)
        pw = math.sqrt(1.0 / a)
        u = np.stack([_normal(rng, m), _normal(rng, m), _normal(rng, m)], axis=1)
        u = u * pw / math.sqrt(2.0)
        return u + np.asarray(v0[:3])[None, :]
    return np.repeat(np.asarray(v0[:3], np.float64)[None, :], m, axis=0)


def _half_rotate(v, ve, tB):
    """The dt/2 rotation with norm restore used by the sources
    (ParticlesInitialize.f90:73-80)."""
    vmin = v - ve[None, :3]
    t = tB
    vmid = np.empty_like(vmin)
    vmid[:, 0] = vmin[:, 0] + vmin[:, 1] * t[2] - vmin[:, 2] * t[1]
    vmid[:, 1] = vmin[:, 1] + vmin[:, 2] * t[0] - vmin[:, 0] * t[2]
    vmid[:, 2] = vmin[:, 2] + vmin[:, 0] * t[1] - vmin[:, 1] * t[0]
    s2 = np.sum(vmid * vmid, axis=1)
    ok = s2 > 0.0
    scale = np.ones_like(s2)
    scale[ok] = np.sqrt(np.sum(vmin[ok] * vmin[ok], axis=1) / s2[ok])
    return vmid * scale[:, None] + ve[None, :3]


def _emit_cells(rng, d: Derived, cfg: Config, corners, species_from):
    """Particles for a list of cell lower corners, in the reference's loop
    order: cell, species ss, ion pp, then its epp electrons and hpp protons."""
    out_r, out_s = [], []
    ncell = corners.shape[0]
    if ncell == 0:
        return np.zeros((0, 8)), np.zeros(0, np.int32)
    for ss in range(species_from, cfg.species + 1):
        s = ss - 1
        nion = int(d.ipp[s])
        if nion == 0:
            continue
        group = 1 + int(d.epp[s]) + int(d.hpp[s])
        m = ncell * nion
        pos = (np.repeat(corners, nion, axis=0) +
               d.dxyz[None, :] * rng.random((m, 3)))
        blk = np.zeros((m, group, 8))
        blk[:, :, 0:3] = pos[:, None, :]
        blk[:, 0, 3:6] = _half_rotate(_maxwellian(rng, m, d.v0[s], d.kelvin[s], d.mass[s]), d.ve,
                                      d.tB[s])
        sp = np.zeros((m, group), np.int32)
        sp[:, 0] = ss
        for e in range(int(d.epp[s])):
            blk[:, 1 + e, 3:6] = _half_rotate(_maxwellian(rng, m, d.v0[s], d.kelvin[0], d.mass[0]),
                                              d.ve, d.tB[0])
            sp[:, 1 + e] = 1
        for h in range(int(d.hpp[s])):
            q = 1 + int(d.epp[s]) + h
            blk[:, q, 3:6] = _half_rotate(_maxwellian(rng, m, d.v0[s], d.kelvin[1], d.mass[1]),
                                          d.ve, d.tB[1])
            sp[:, q] = 2
        out_r.append(blk.reshape(ncell, nion * group, 8))
        out_s.append(sp.reshape(ncell, nion * group))
    if not out_r:
        return np.zeros((0, 8)), np.zeros(0, np.int32)
    # order: cell-major, then species block
    r = np.concatenate(out_r, axis=1).reshape(-1, 8)
    s = np.concatenate(out_s, axis=1).reshape(-1).astype(np.int32)
    return np.ascontiguousarray(r), np.ascontiguousarray(s)


def _cell_corners(cfg: Config, d: Derived, ii, jj, kk):
    """Lower corners x0 = Lmin + dx*(i-2) for 1-based cell indices (arrays)."""
    K, J, I = np.meshgrid(kk, jj, ii, indexing="ij")
    x0 = cfg.gmin[0] + d.dxyz[0] * (I.ravel() - 2)
    y0 = cfg.gmin[1] + d.dxyz[1] * (J.ravel() - 2)
    z0 = cfg.gmin[2] + d.dxyz[2] * (K.ravel() - 2)
    return np.stack([x0, y0, z0], axis=1)


def particles_initialize(cfg: Config, d: Derived, rng):
    """ParticlesInitialize.f90:54-119: every interior cell, co-located ions and electrons."""
    c = _cell_corners(cfg, d, np.arange(2, cfg.nx + 2), np.arange(2, cfg.ny + 2),
                      np.arange(2, cfg.nz + 2))
    return _emit_cells(rng, d, cfg, c, 1)


def bc_inflow(cfg: Config, d: Derived, rng):
    """BCinflow.f90: fresh plasma in the guard cells of every flow face, in the
    reference's face order and cell ranges (+x, -x over the y/z-widened face;
    +y with z up to Nz+1+zflow, -y from 2-zflow; +z with y from 2-yflow, -z up
    to Ny+1+yflow; species from 2 on the y and z faces)."""
    xf, yf, zf = cfg.flow
    yneg = ypos = 1 if yf else 0
    zneg = zpos = 1 if zf else 0
    parts_r, parts_s = [], []

    def add(corners, first_species, upper=None):
        if upper is not None:             # the + faces start at L_max itself (:59, :201, :342)
            corners = corners.copy()
            corners[:, upper] = cfg.gmax[upper]
        r, s = _emit_cells(rng, d, cfg, corners, first_species)
        parts_r.append(r)
        parts_s.append(s)

    jj = np.arange(2 - yneg, cfg.ny + 2 + ypos)
    kk = np.arange(2 - zneg, cfg.nz + 2 + zpos)
    if xf:
        c = _cell_corners(cfg, d, np.array([cfg.nx + 2]), jj, kk)    # x0 = Lx_max (:59)
        add(c, 1, upper=0)
        c = _cell_corners(cfg, d, np.array([1]), jj, kk)             # x0 = Lx_min - dx (:130)
        add(c, 1)
    ii = np.arange(2, cfg.nx + 2)
    if yf:
        add(_cell_corners(cfg, d, ii, np.array([cfg.ny + 2]), np.arange(2, cfg.nz + 2 + zpos)), 2,
            upper=1)
        add(_cell_corners(cfg, d, ii, np.array([1]), np.arange(2 - zneg, cfg.nz + 2)), 2)
    if zf:   # the z faces take one y guard row on opposite sides (BCinflow.f90:337, :407)
        add(_cell_corners(cfg, d, ii, np.arange(2 - yneg, cfg.ny + 2), np.array([cfg.nz + 2])), 2,
            upper=2)
        add(_cell_corners(cfg, d, ii, np.arange(2, cfg.ny + 2 + ypos), np.array([1])), 2)
    if not parts_r:
        return np.zeros((0, 8)), np.zeros(0, np.int32)
    return np.concatenate(parts_r), np.concatenate(parts_s).astype(np.int32)


# ---------------------------------------------------------------------------

def geometry(cfg: Config, d: Derived, max_particles: int, device: int = 0,
             precision: int = _lib.PICARD_FP64, order_free: bool = False) -> _lib.Geometry:
    g = _lib.Geometry()
    g.nx, g.ny, g.nz, g.species = cfg.nx, cfg.ny, cfg.nz, cfg.species
    for c in range(3):
        g.flow[c] = 1 if cfg.flow[c] else 0
        g.lmin[c] = g.gmin[c] = cfg.gmin[c]
        g.lmax[c] = g.gmax[c] = cfg.gmax[c]
        g.dxyz[c] = d.dxyz[c]
    g.precision = precision
    g.deposit = (_lib.PICARD_DEPOSIT_FIXED_POINT if order_free
                 else _lib.PICARD_DEPOSIT_REFERENCE_ORDER)
    g.device = device
    g.max_particles = max_particles
    return g


def species_struct(cfg: Config, d: Derived) -> _lib.Species:
    s = _lib.Species()
    for q in range(8):
        s.charge[q] = d.charge[q]
        s.n2p[q] = d.n2p[q]
        s.eta[q] = d.eta[q]
        s.fluid[q] = 1 if cfg.fluid[q] else 0
        s.gyration[q] = 1 if cfg.gyration[q] else 0
        for c in range(4):
            s.tB[4 * q + c] = d.tB[q, c]
            s.sB[4 * q + c] = d.sB[q, c]
    for c in range(4):
        s.ve[c] = d.ve[c]
        s.B0[c] = d.B0[c]
        s.E0[c] = d.E0[c]
    return s


def haser_struct(cfg: Config, d: Derived | None = None) -> _lib.Haser:
    hs = _lib.Haser()
    for q in range(8):
        hs.Qn[q] = cfg.Qn[q]
        hs.vn[q] = cfg.vn[q]
        hs.nu_i[q] = cfg.nu_i[q]
        hs.nu_d[q] = cfg.nu_d[q]
        hs.Ek[q] = cfg.Ek[q]
    hs.rate_scale = cfg.rate_scale
    return hs


def plasma_struct(d: Derived) -> _lib.Plasma:
    pl = _lib.Plasma()
    for q in range(8):
        pl.ipp[q] = int(d.ipp[q])
        pl.epp[q] = int(d.epp[q])
        pl.hpp[q] = int(d.hpp[q])
        pl.mass[q] = d.mass[q]
        pl.kelvin[q] = d.kelvin[q]
        for c in range(4):
            pl.v0[4 * q + c] = d.v0[q, c]
    return pl


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _ip(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))


class Domain:
    """One device-resident domain; methods follow the reference's subroutines."""

    def __init__(self, cfg: Config, max_particles: int | None = None, device: int = 0,
                 d: Derived | None = None, precision: int = _lib.PICARD_FP64,
                 order_free: bool = False):
        """order_free: the FP64 form with the fixed-point deposit and the FFT
        face sums (PICARD_DEPOSIT_FIXED_POINT) instead of the reference's
        summation orders."""
        self.cfg = cfg
        self.d = derive(cfg) if d is None else d
        if max_particles is None:
            per_cell = int(self.d.ppc[: cfg.species].sum())
            max_particles = int(1.5 * per_cell * (cfg.nx + 2) * (cfg.ny + 2) * (cfg.nz + 2)) + 1024
        self.geo = geometry(cfg, self.d, max_particles, device, precision, order_free)
        self.precision = precision
        self.spc = species_struct(cfg, self.d)
        h = ct.c_void_p()
        check("picard_create", lib().picard_create(ct.byref(self.geo), ct.byref(self.spc),
                                                   ct.byref(h)))
        self.h = h
        self.shape = (cfg.nz + 2, cfg.ny + 2, cfg.nx + 2)

    def close(self):
        if self.h:
            lib().picard_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # particles
    def append(self, rp, sp):
        rp = np.ascontiguousarray(rp, np.float64)
        sp = np.ascontiguousarray(sp, np.int32)
        assert rp.ndim == 2 and rp.shape[1] == 8 and sp.shape == (rp.shape[0],)
        check("picard_particles_append",
              lib().picard_particles_append(self.h, _dp(rp), _ip(sp), rp.shape[0]))

    @property
    def num_local(self) -> int:
        n = ct.c_int64()
        check("picard_particles_count", lib().picard_particles_count(self.h, ct.byref(n)))
        return n.value

    def particles(self, first=0, count=None):
        n = self.num_local
        count = n - first if count is None else count
        rp = np.zeros((count, 8), np.float64)
        sp = np.zeros(count, np.int32)
        check("picard_particles_read",
              lib().picard_particles_read(self.h, first, count, _dp(rp), _ip(sp)))
        return rp, sp

    # fields
    def field(self, which):
        shp = {_lib.FIELD_U: self.shape, _lib.FIELD_C: self.shape,
               _lib.FIELD_E: self.shape + (3,), _lib.FIELD_N: self.shape + (self.cfg.species,),
               _lib.FIELD_D: (2 * self.cfg.nz + 1, 2 * self.cfg.ny + 1, 2 * self.cfg.nx + 1),
               _lib.FIELD_F: self.shape + (3, self.cfg.species)}[which]
        a = np.zeros(shp, np.float64)
        check("picard_field_read", lib().picard_field_read(self.h, which, _dp(a), a.size))
        return a

    def set_field(self, which, a):
        a = np.ascontiguousarray(a, np.float64)
        check("picard_field_write", lib().picard_field_write(self.h, which, _dp(a), a.size))

    # the call surface, picard.f90:564-631
    def advance_position(self, dt=None):
        check("picard_advance_position",
              lib().picard_advance_position(self.h, self.d.dt if dt is None else dt))

    def bc_outflow(self):
        n, r = ct.c_int64(), ct.c_int64()
        check("picard_bc_outflow", lib().picard_bc_outflow(self.h, ct.byref(n), ct.byref(r)))
        return n.value, r.value

    def particles_transfer(self):
        check("picard_particles_transfer", lib().picard_particles_transfer(self.h))

    def particles_gridify(self):
        check("picard_particles_gridify", lib().picard_particles_gridify(self.h))

    def calc_distance(self):
        check("picard_calc_distance", lib().picard_calc_distance(self.h))

    def bc_potentials(self):
        check("picard_bc_potentials", lib().picard_bc_potentials(self.h))

    def calc_epotential(self, maxerr=None, maxit=None):
        sw = ct.c_int64()
        rc = lib().picard_calc_epotential(self.h, self.cfg.maxerr if maxerr is None else maxerr,
                                          self.cfg.maxit if maxit is None else maxit,
                                          ct.byref(sw))
        check("picard_calc_epotential", rc)
        return sw.value

    def calc_efield(self):
        check("picard_calc_efield", lib().picard_calc_efield(self.h))

    def bc_fields(self):
        check("picard_bc_fields", lib().picard_bc_fields(self.h))

    def advance_velocity(self, dt=None):
        check("picard_advance_velocity",
              lib().picard_advance_velocity(self.h, self.d.dt if dt is None else dt))

    def step(self, produced=None):
        st = _lib.StepStats()
        if produced is None:
            rp = np.zeros((0, 8))
            sp = np.zeros(0, np.int32)
        else:
            rp, sp = (np.ascontiguousarray(produced[0], np.float64),
                      np.ascontiguousarray(produced[1], np.int32))
        check("picard_step", lib().picard_step(self.h, self.d.dt, self.cfg.maxerr, self.cfg.maxit,
                                               _dp(rp), _ip(sp), rp.shape[0], ct.byref(st)))
        return st

    def generate_initial(self, seed: int):
        """ParticlesInitialize on the device (Philox-keyed by seed)."""
        pl = plasma_struct(self.d)
        check("picard_generate_initial", lib().picard_generate_initial(self.h, ct.byref(pl), seed))

    def generate_inflow(self, seed: int, step: int):
        """BCinflow on the device for one step."""
        pl = plasma_struct(self.d)
        check("picard_generate_inflow",
              lib().picard_generate_inflow(self.h, ct.byref(pl), seed, step))

    def produce(self, seed: int, step: int) -> int:
        """ParticlesProduce (Haser source) on the device for one step; returns
        the number of particles appended."""
        pl = plasma_struct(self.d)
        hs = haser_struct(self.cfg, self.d)
        n = ct.c_int64()
        check("picard_particles_produce",
              lib().picard_particles_produce(self.h, ct.byref(pl), ct.byref(hs), seed, step,
                                             self.d.dt, ct.byref(n)))
        return n.value

    def set_production(self, seed: int | None):
        """Run the Haser source inside step() (after the position push, as
        picard.f90:566); None switches it off."""
        if seed is None:
            check("picard_set_production", lib().picard_set_production(self.h, None, None, 0))
            return
        pl = plasma_struct(self.d)
        hs = haser_struct(self.cfg, self.d)
        check("picard_set_production",
              lib().picard_set_production(self.h, ct.byref(pl), ct.byref(hs), seed))

    def output_step(self, path, write_fields: bool, write_particles: bool, fields_index: int = 0,
                    particles_index: int = 0, produced=None):
        """One main-loop iteration on an output step (picard.f90:556-628): the
        phases of step() up to BCfields, then AdvanceVelocityImp, the particle
        dump, ParticlesGridifyAll and the field dump, and AdvanceVelocityExp.
        Writes under `path` in the reference's layout; returns the file names."""
        from . import output
        self.advance_position()                                   # :564
        if produced is not None:
            self.append(*produced)                                # :566
        self.bc_outflow()                                         # :569
        self.particles_transfer()                                 # :572
        self.particles_gridify()                                  # :577
        self.bc_potentials()                                      # :581
        sweeps = self.calc_epotential()                           # :584
        self.calc_efield()                                        # :587
        self.bc_fields()                                          # :589
        self.advance_velocity_imp()                               # :603
        names = []
        if write_particles:                                       # :606-608
            rp, sp = self.particles()
            names.append(output.write_particles(path, rp, sp, particles_index))
        if write_fields:                                          # :610-615
            self.particles_gridify_all()
            names += output.write_fields(path, self.field(_lib.FIELD_U), self.field(_lib.FIELD_N),
                                         self.field(_lib.FIELD_F), self.field(_lib.FIELD_E),
                                         fields_index)
        self.advance_velocity_exp()                               # :616
        return names, sweeps

    def particles_gridify_all(self):
        """ParticlesGridifyAll: N, C and the number fluxes (field FIELD_F)."""
        check("picard_particles_gridify_all", lib().picard_particles_gridify_all(self.h))

    def advance_velocity_imp(self, deltat=None):
        """AdvanceVelocityImp (output steps; the main loop passes deltat = dt/2)."""
        deltat = 0.5 * self.d.dt if deltat is None else deltat
        check("picard_advance_velocity_imp",
              lib().picard_advance_velocity_imp(self.h, deltat, self.d.dt))

    def advance_velocity_exp(self, deltat=None):
        """AdvanceVelocityExp (output steps; the main loop passes deltat = dt/2)."""
        deltat = 0.5 * self.d.dt if deltat is None else deltat
        check("picard_advance_velocity_exp",
              lib().picard_advance_velocity_exp(self.h, deltat, self.d.dt))

    def diagnostics(self) -> dict:
        """The diagnostics line (picard.f90:649-672): particle count and the
        particle, electric and magnetic energies, computed on the device."""
        out = _lib.Diag()
        mass = np.ascontiguousarray(self.d.mass, np.float64)
        check("picard_diagnostics", lib().picard_diagnostics(self.h, _dp(mass), ct.byref(out)))
        return dict(num_local=out.num_local, num_global=out.num_global,
                    particle_energy=out.particle_energy, e_energy=out.e_energy,
                    b_energy=out.b_energy)

    def timing(self, enable: bool):
        check("picard_timing_enable", lib().picard_timing_enable(self.h, 1 if enable else 0))

    def split_solver(self, inner: int = 512):
        """FP32 sweeps per FP64 refinement step of the FP32-split solve."""
        check("picard_split_solver", lib().picard_split_solver(self.h, inner))

    def split_sweeps_per_launch(self, sweeps: int = 2):
        """At most this many inner sweeps per launch of the FP32-split solve
        (1, 2 or 3; the library's default is 2)."""
        check("picard_split_sweeps_per_launch",
              lib().picard_split_sweeps_per_launch(self.h, sweeps))

    def timing_read(self):
        ms, n = ct.c_double(), ct.c_int64()
        check("picard_timing_read", lib().picard_timing_read(self.h, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def timing_kernel_clock(self):
        """The timed launches by the kernels' own clock stamps: [ms], count."""
        ms, n = ct.c_double(), ct.c_int64()
        check("picard_timing_kernel_clock",
              lib().picard_timing_kernel_clock(self.h, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def timing_sweeps(self):
        """Jacobi sweeps of the launches timing_read timed."""
        n = ct.c_int64()
        check("picard_timing_sweeps", lib().picard_timing_sweeps(self.h, ct.byref(n)))
        return n.value

    def timing_solve(self):
        """Wall time [ms] of the solve phase of step() since timing(True), and
        the number of solves timed."""
        ms, n = ct.c_double(), ct.c_int64()
        check("picard_timing_solve", lib().picard_timing_solve(self.h, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        check("picard_synchronize", lib().picard_synchronize(self.h))
