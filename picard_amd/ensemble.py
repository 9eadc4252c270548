"""Host side of the ensemble API (include/picard_ensemble.h): many independent
simulations stepped together on one device, and the parameter draws of the
64-simulation workload.

Each simulation is the 32^3 workload's setup at its own size and plasma: a
box of cubic cells (x in/outflow, y and z periodic), H+ and e- at `ppc`
macroparticles per cell each, a drifting Maxwellian, uniform B, dt =
0.1 / omega_pe, Jacobi tolerance 1e-6.  No reference counterpart: the
reference runs one simulation per process (picard.f90:12-684)."""
from __future__ import annotations

import ctypes as ct
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib

EC = 1.60217646e-19          # picard.f90:57
KB = 1.380650e-23            # picard.f90:64
EPS0 = 8.854187817e-12       # picard.f90:65
EV_K = 11604.5               # kelvin per eV (10 eV = 116045 K, the 32^3 workload)


@dataclass
class SimSpec:
    nx: int
    ny: int
    nz: int
    ppc: int
    dx: float
    density: float           # H+ [m^-3]
    kelvin: float
    drift: float             # [m/s] along +x
    B: tuple                 # [T]
    seed: int
    dt_wpe: float = 0.1
    maxerr: float = 1e-6
    maxit: int = 2 ** 20

    def cstruct(self) -> _lib.SimParams:
        p = _lib.SimParams()
        p.nx, p.ny, p.nz, p.ppc = self.nx, self.ny, self.nz, self.ppc
        p.dx, p.density, p.kelvin, p.drift = self.dx, self.density, self.kelvin, self.drift
        for c in range(3):
            p.B[c] = self.B[c]
        p.dt_wpe, p.maxerr, p.maxit = self.dt_wpe, self.maxerr, self.maxit
        p.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        return p

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def debye_length(density: float, kelvin: float) -> float:
    """sqrt(eps0 kB T / (n e^2)) [m]."""
    return math.sqrt(EPS0 * KB * kelvin / (density * EC * EC))


def draw_ensemble(count: int = 64, seed: int = 20261210, nmin: int = 16, nmax: int = 96,
                  ppc=(4, 32)) -> list[SimSpec]:
    """The 64-simulation workload's draws: cubic grids with N log-uniform in
    [nmin, nmax], ppc uniform in 4..32 per species, n uniform in 1-20 cm^-3,
    T uniform in 1-50 eV, drift uniform in 300-700 km/s along x, |B| uniform
    in 1-10 nT along a uniformly random direction; cells of one Debye length
    (the 32^3 workload's 10 m at 10 eV and 5 cm^-3)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        dens = rng.uniform(1.0, 20.0) * 1e6
        T = rng.uniform(1.0, 50.0) * EV_K
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        B = tuple(float(v) for v in u * rng.uniform(1.0, 10.0) * 1e-9)
        out.append(SimSpec(nx=n, ny=n, nz=n, ppc=int(rng.integers(ppc[0], ppc[1] + 1)),
                           dx=debye_length(dens, T), density=dens, kelvin=T,
                           drift=rng.uniform(300.0, 700.0) * 1e3, B=B,
                           seed=int(rng.integers(1, 2 ** 62))))
    return out


def derive(spec: SimSpec) -> _lib.SimDerived:
    """picard.f90:333-466 for one simulation (host only)."""
    d = _lib.SimDerived()
    p = spec.cstruct()
    check("picard_sim_derive", lib().picard_sim_derive(ct.byref(p), ct.byref(d)))
    return d


def footprint(spec: SimSpec) -> tuple[int, int]:
    """(device bytes of the simulation's grids, particle slots it needs)."""
    g, n = ct.c_int64(), ct.c_int64()
    p = spec.cstruct()
    check("picard_sim_footprint", lib().picard_sim_footprint(ct.byref(p), ct.byref(g),
                                                             ct.byref(n)))
    return g.value, n.value


# bytes per particle slot of the ensemble's arena: int16 cell x3, FP32 offset
# x3, FP32 velocity x3, species, slot, id, sort key; two arenas for the sort
PARTICLE_BYTES = 2 * (6 + 12 + 12 + 1 + 2 + 4) + 4


class Ensemble:
    """Simulations stepped together on one device (picard_ensemble_*)."""

    def __init__(self, max_sims: int, max_cells: int, max_particles: int, device: int = 0,
                 inner: int = 512, graphs: bool = True):
        cfg = _lib.EnsembleConfig()
        cfg.device, cfg.max_sims, cfg.max_cells = device, max_sims, max_cells
        cfg.max_particles, cfg.inner, cfg.use_graphs = max_particles, inner, 1 if graphs else 0
        h = ct.c_void_p()
        check("picard_ensemble_create", lib().picard_ensemble_create(ct.byref(cfg), ct.byref(h)))
        self.h = h
        self.specs: dict[int, SimSpec] = {}

    @classmethod
    def sized_for(cls, specs, steps_room: int = 4, **kw) -> "Ensemble":
        """An ensemble with room for `specs` (cells, and particles for their
        loads plus `steps_room` steps of inflow)."""
        cells = sum(s.cells for s in specs)
        parts = 0
        for s in specs:
            _, p = footprint(s)
            parts += p + steps_room * 4 * s.ny * s.nz * s.ppc
        return cls(max_sims=len(specs), max_cells=cells, max_particles=int(parts * 1.02) + 4096, **kw)

    def close(self):
        if self.h:
            lib().picard_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, spec: SimSpec) -> int:
        slot = ct.c_int32()
        p = spec.cstruct()
        check("picard_ensemble_add", lib().picard_ensemble_add(self.h, ct.byref(p), ct.byref(slot)))
        self.specs[slot.value] = spec
        return slot.value

    def remove(self, slot: int):
        check("picard_ensemble_remove", lib().picard_ensemble_remove(self.h, slot))
        self.specs.pop(slot, None)

    def step(self) -> _lib.EnsembleStats:
        st = _lib.EnsembleStats()
        check("picard_ensemble_step", lib().picard_ensemble_step(self.h, ct.byref(st)))
        return st

    def status(self, slot: int) -> _lib.SimStatus:
        st = _lib.SimStatus()
        check("picard_ensemble_status", lib().picard_ensemble_status(self.h, slot, ct.byref(st)))
        return st

    def field(self, slot: int, which: int) -> np.ndarray:
        s = self.specs[slot]
        shp = (s.nz + 2, s.ny + 2, s.nx + 2)
        shp = {_lib.FIELD_E: shp + (3,), _lib.FIELD_N: shp + (2,)}.get(which, shp)
        a = np.zeros(shp, np.float64)
        check("picard_ensemble_field_read",
              lib().picard_ensemble_field_read(self.h, slot, which,
                                               a.ctypes.data_as(ct.POINTER(ct.c_double)), a.size))
        return a

    def particles(self, slot: int):
        """(real_particles [n, 8] with the id in slot 8, species [n]) sorted by id."""
        n = self.status(slot).num_local
        rp = np.zeros((n, 8), np.float64)
        sp = np.zeros(n, np.int32)
        check("picard_ensemble_particles_read",
              lib().picard_ensemble_particles_read(self.h, slot,
                                                   rp.ctypes.data_as(ct.POINTER(ct.c_double)),
                                                   sp.ctypes.data_as(ct.POINTER(ct.c_int32)), n))
        o = np.argsort(rp[:, 7], kind="stable")
        return rp[o], sp[o]

    def tune(self, sweeps: int = 64):
        """The autotuner: returns (chosen tile, per-sweep microseconds of every
        candidate)."""
        tile = _lib.EnsembleTile()
        times = np.zeros(64)
        ncand = ct.c_int32()
        check("picard_ensemble_tune",
              lib().picard_ensemble_tune(self.h, sweeps, ct.byref(tile),
                                         times.ctypes.data_as(ct.POINTER(ct.c_double)), 64,
                                         ct.byref(ncand)))
        return tile, times[: ncand.value]

    def set_tile(self, cells_per_thread: int, blocks_per_cu: int, threads: int = 256):
        t = _lib.EnsembleTile(cells_per_thread, threads, blocks_per_cu, 0)
        check("picard_ensemble_set_tile", lib().picard_ensemble_set_tile(self.h, ct.byref(t)))

    def get_tile(self) -> _lib.EnsembleTile:
        t = _lib.EnsembleTile()
        check("picard_ensemble_get_tile", lib().picard_ensemble_get_tile(self.h, ct.byref(t)))
        return t
