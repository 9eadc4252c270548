"""ctypes bindings for the CPU oracle and for the reference's own compiled subroutines.

TEST INFRASTRUCTURE ONLY.  ``oracle`` restates Picard's PIC time step in C
(FP64, reference evaluation order) so that tests can check the HIP product
path against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package.

Array conventions (numpy, C order), matching the Fortran column-major layout
of the reference:

* ``rp``  real_particles(8, n)  -> ``float64[n, 8]``
* ``sp``  int_particles(n)      -> ``int32[n]`` (species, 1-based)
* ``U``, ``C``                  -> ``float64[Nz+2, Ny+2, Nx+2]``
* ``E``                         -> ``float64[Nz+2, Ny+2, Nx+2, 3]``
* ``N``                         -> ``float64[Nz+2, Ny+2, Nx+2, S]``
* ``D``                         -> ``float64[2Nz+1, 2Ny+1, 2Nx+1]``
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
import sys
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpicard_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpicard_ref.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int32)
_lp = ct.POINTER(ct.c_int64)


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref where the reference is mounted)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


class _Grid(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("species", ct.c_int32),
        ("flow", ct.c_int32 * 3),
        ("lmin", ct.c_double * 3), ("lmax", ct.c_double * 3),
        ("gmin", ct.c_double * 3), ("gmax", ct.c_double * 3),
        ("dxyz", ct.c_double * 3),
    ]


@dataclass
class Grid:
    """One-rank domain: the local box equals the global box."""
    nx: int
    ny: int
    nz: int
    species: int
    flow: tuple = (True, False, False)
    gmin: tuple = (-160.0, -160.0, -160.0)
    gmax: tuple = (160.0, 160.0, 160.0)
    dxyz: tuple = field(default=None)

    def __post_init__(self):
        if self.dxyz is None:
            # picard.f90:339-360 on one rank: Lx_local = Lx, dxyz = Lx_local/Nx_local
            self.dxyz = tuple((self.gmax[c] - self.gmin[c]) / float(n)
                              for c, n in enumerate((self.nx, self.ny, self.nz)))

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    def cstruct(self) -> _Grid:
        g = _Grid()
        g.nx, g.ny, g.nz, g.species = self.nx, self.ny, self.nz, self.species
        for c in range(3):
            g.flow[c] = 1 if self.flow[c] else 0
            g.lmin[c] = self.gmin[c]
            g.lmax[c] = self.gmax[c]
            g.gmin[c] = self.gmin[c]
            g.gmax[c] = self.gmax[c]
            g.dxyz[c] = self.dxyz[c]
        return g


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        G = ct.POINTER(_Grid)
        L.orc_advance_position.argtypes = [_dp, ct.c_double, ct.c_int64]
        L.orc_bc_outflow.argtypes = [G, _dp, _ip, _lp, _lp]
        L.orc_bc_outflow.restype = ct.c_int64
        L.orc_particles_transfer.argtypes = [G, _dp, ct.c_int64]
        L.orc_particles_gridify.argtypes = [G, _dp, _dp, _dp, _ip, ct.c_int64, _dp, _dp]
        L.orc_particles_gridify.restype = ct.c_int64
        L.orc_calc_distance.argtypes = [_dp, ct.c_int32, ct.c_int32, ct.c_int32, _dp, _ip]
        L.orc_bc_potentials.argtypes = [G, _dp, _dp, _dp]
        L.orc_calc_epotential.argtypes = [G, _dp, _dp, ct.c_double, ct.c_int32, _dp]
        L.orc_calc_epotential.restype = ct.c_int64
        L.orc_calc_epotential_pairwise.argtypes = [G, _dp, _dp, ct.c_double, ct.c_int32]
        L.orc_calc_epotential_pairwise.restype = ct.c_int64
        L.orc_calc_efield.argtypes = [G, _dp, _dp]
        L.orc_bc_fields.argtypes = [G, _dp]
        L.orc_advance_velocity.argtypes = [G, _dp, _ip, _dp, _dp, ct.c_double, _dp, _dp, _dp,
                                           _dp, ct.c_int64, _ip, _ip]
        L.orc_particles_gridify_all.argtypes = [G, _dp, _dp, _dp, _dp, _ip, ct.c_int64, _dp, _dp]
        L.orc_particles_gridify_all.restype = ct.c_int64
        L.orc_diagnostics.argtypes = [G, _dp, _ip, ct.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_advance_velocity_ie.argtypes = [ct.c_int, G, _dp, _ip, _dp, _dp, ct.c_double,
                                              ct.c_double, _dp, _dp, _dp, ct.c_int64, _ip, _ip]
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _f8(v, n=8):
    a = np.zeros(n, np.float64)
    v = np.asarray(v, np.float64).ravel()
    a[: v.size] = v
    return a


def _i8(v, n=8):
    a = np.zeros(n, np.int32)
    v = np.asarray(v).astype(np.int32).ravel()
    a[: v.size] = v
    return a


# ---------------------------------------------------------------------------
# oracle entry points (in-place on numpy arrays, like the Fortran subroutines)

def advance_position(rp, dt, n=None):
    n = rp.shape[0] if n is None else n
    lib().orc_advance_position(_d(rp), dt, n)


def bc_outflow(g: Grid, rp, sp, n=None):
    """Returns (new n, removed slots in loop order (1-based))."""
    n = rp.shape[0] if n is None else n
    nn = ct.c_int64(n)
    removed = np.zeros(max(n, 1), np.int64)
    cnt = lib().orc_bc_outflow(ct.byref(g.cstruct()), _d(rp), _i(sp), ct.byref(nn),
                               removed.ctypes.data_as(_lp))
    return nn.value, removed[:cnt].copy()


def particles_transfer(g: Grid, rp, n=None):
    n = rp.shape[0] if n is None else n
    lib().orc_particles_transfer(ct.byref(g.cstruct()), _d(rp), n)


def particles_gridify(g: Grid, rp, sp, charge, n2p, n=None):
    n = rp.shape[0] if n is None else n
    N = np.zeros(g.shape + (g.species,), np.float64)
    C = np.zeros(g.shape, np.float64)
    skipped = lib().orc_particles_gridify(ct.byref(g.cstruct()), _d(N), _d(C), _d(rp), _i(sp), n,
                                          _d(_f8(charge)), _d(_f8(n2p)))
    return N, C, skipped


def calc_distance(nx, ny, nz, dxyz, flow):
    D = np.zeros((2 * nz + 1, 2 * ny + 1, 2 * nx + 1), np.float64)
    lib().orc_calc_distance(_d(D), nx, ny, nz, _d(np.asarray(dxyz, np.float64)),
                            _i(_i8(flow, 3)))
    return D


def bc_potentials(g: Grid, C, D, U=None):
    U = np.zeros(g.shape, np.float64) if U is None else U
    lib().orc_bc_potentials(ct.byref(g.cstruct()), _d(U), _d(C), _d(D))
    return U


def calc_epotential(g: Grid, U, C, maxerr, maxit):
    """In place on U.  Returns (sweeps, final sqrt(r)/sqrt(b))."""
    ratio = ct.c_double(0.0)
    sweeps = lib().orc_calc_epotential(ct.byref(g.cstruct()), _d(U), _d(C), maxerr, maxit,
                                       ct.byref(ratio))
    return sweeps, ratio.value


def calc_epotential_pairwise(g: Grid, U, C, maxerr, maxit):
    return lib().orc_calc_epotential_pairwise(ct.byref(g.cstruct()), _d(U), _d(C), maxerr, maxit)


def calc_efield(g: Grid, U):
    E = np.zeros(g.shape + (3,), np.float64)
    lib().orc_calc_efield(ct.byref(g.cstruct()), _d(E), _d(U))
    return E


def bc_fields(g: Grid, E):
    lib().orc_bc_fields(ct.byref(g.cstruct()), _d(E))


def advance_velocity(g: Grid, rp, sp, E, B0, dt, ve, eta, tB, sB, fluid=None, gyration=None, n=None):
    n = rp.shape[0] if n is None else n
    fluid = _i8([0] * 8 if fluid is None else fluid)
    gyration = _i8([1] * 8 if gyration is None else gyration)
    tB = np.ascontiguousarray(np.asarray(tB, np.float64).reshape(8, 4))
    sB = np.ascontiguousarray(np.asarray(sB, np.float64).reshape(8, 4))
    lib().orc_advance_velocity(ct.byref(g.cstruct()), _d(rp), _i(sp), _d(E), _d(_f8(B0, 4)), dt,
                               _d(_f8(ve, 4)), _d(_f8(eta)), _d(tB), _d(sB), n, _i(fluid),
                               _i(gyration))


def advance_velocity_ie(mode, g: Grid, rp, sp, E, B0, deltat, dt, ve, eta, tB, fluid=None,
                        gyration=None, n=None):
    """AdvanceVelocityImp (mode 1) or AdvanceVelocityExp (mode 2)."""
    n = rp.shape[0] if n is None else n
    fluid = _i8([0] * 8 if fluid is None else fluid)
    gyration = _i8([1] * 8 if gyration is None else gyration)
    tB = np.ascontiguousarray(np.asarray(tB, np.float64).reshape(8, 4))
    lib().orc_advance_velocity_ie(mode, ct.byref(g.cstruct()), _d(rp), _i(sp), _d(E),
                                  _d(_f8(B0, 4)), deltat, dt, _d(_f8(ve, 4)), _d(_f8(eta)),
                                  _d(tB), n, _i(fluid), _i(gyration))


def particles_gridify_all(g: Grid, rp, sp, charge, n2p, n=None):
    """ParticlesGridifyAll: returns (F [Z,Y,X,3,S], N [Z,Y,X,S], C, skipped)."""
    n = rp.shape[0] if n is None else n
    F = np.zeros(g.shape + (3, g.species))
    N = np.zeros(g.shape + (g.species,))
    C = np.zeros(g.shape)
    sk = lib().orc_particles_gridify_all(ct.byref(g.cstruct()), _d(F), _d(N), _d(C),
                                         _d(np.ascontiguousarray(rp)),
                                         _i(np.ascontiguousarray(sp, np.int32)), n,
                                         _d(_f8(charge)), _d(_f8(n2p)))
    return F, N, C, sk


def diagnostics(g: Grid, rp, sp, E, n2p, mass, E0, B0, n=None):
    """(uP, uE, uB) of the diagnostics line, picard.f90:649-672."""
    n = rp.shape[0] if n is None else n
    out = np.zeros(3)
    lib().orc_diagnostics(ct.byref(g.cstruct()), _d(np.ascontiguousarray(rp)),
                          _i(np.ascontiguousarray(sp, np.int32)), n, _d(np.ascontiguousarray(E)),
                          _d(_f8(n2p)), _d(_f8(mass)), _d(_f8(E0, 4)), _d(_f8(B0, 4)), _d(out))
    return tuple(float(v) for v in out)


# ---------------------------------------------------------------------------
# the reference's own MPI-free subroutines, compiled from /root/reference by
# oracle/Makefile into oracle/_ref (absent on machines without the reference)

_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = ct.CDLL(REF_PATH)
    return _ref


def _byref_d(x):
    return ct.byref(ct.c_double(x))


def _byref_i(x):
    return ct.byref(ct.c_int32(int(x)))


def ref_advance_position(rp, dt, n):
    """AdvancePosition(real_particles, dt, max_per_proc, num_local)."""
    R = ref()
    R.advanceposition_(_d(rp), _byref_d(dt), _byref_i(rp.shape[0]), _byref_i(n))


def ref_bc_outflow(g: Grid, rp, sp, n):
    """BCoutflow(...) with Lx_min.. = xmin.. on one rank.  Returns the new num_local."""
    R = ref()
    nl = ct.c_int32(n)
    a = [ct.c_double(v) for v in (g.gmin[0], g.gmax[0], g.gmin[1], g.gmax[1], g.gmin[2], g.gmax[2])]
    R.bcoutflow_(_d(rp), _i(sp), *[ct.byref(x) for x in a], *[ct.byref(ct.c_double(v)) for v in
                 (g.gmin[0], g.gmax[0], g.gmin[1], g.gmax[1], g.gmin[2], g.gmax[2])],
                 _byref_i(rp.shape[0]), ct.byref(nl),
                 *[_byref_i(1 if f else 0) for f in g.flow])
    return nl.value


def ref_calc_distance(nx, ny, nz, dxyz, flow):
    R = ref()
    D = np.zeros((2 * nz + 1, 2 * ny + 1, 2 * nx + 1), np.float64)
    R.calcdistance_(_d(D), _byref_i(nx), _byref_i(ny), _byref_i(nz),
                    _d(np.asarray(dxyz, np.float64)), *[_byref_i(1 if f else 0) for f in flow])
    return D


def ref_advance_velocity_ie(mode, g: Grid, rp, sp, E, B0, deltat, dt, v0, ve, eta, tB, n,
                            fluid=None, gyration=None):
    """AdvanceVelocityImp (mode 1) / AdvanceVelocityExp (mode 2) of the
    reference: (real_particles, int_particles, E, B0, Lx_min, Ly_min, Lz_min,
    dxyz, deltat, dt, v0, ve, eta, tB, num_local, max_per_proc, species,
    Nx_local, Ny_local, Nz_local, fluid, gyration) (AdvanceVelocity.f90:128-130)."""
    R = ref()
    fluid = _i8([0] * 8 if fluid is None else fluid)
    gyration = _i8([1] * 8 if gyration is None else gyration)
    tB = np.ascontiguousarray(np.asarray(tB, np.float64).reshape(8, 4))
    v0 = np.ascontiguousarray(np.asarray(v0, np.float64).reshape(8, 4))
    fn = R.advancevelocityimp_ if mode == 1 else R.advancevelocityexp_
    fn(_d(rp), _i(sp), _d(E), _d(_f8(B0, 4)), _byref_d(g.gmin[0]), _byref_d(g.gmin[1]),
       _byref_d(g.gmin[2]), _d(np.asarray(g.dxyz, np.float64)), _byref_d(deltat), _byref_d(dt),
       _d(v0), _d(_f8(ve, 4)), _d(_f8(eta)), _d(tB), _byref_i(n), _byref_i(rp.shape[0]),
       _byref_i(g.species), _byref_i(g.nx), _byref_i(g.ny), _byref_i(g.nz), _i(fluid),
       _i(gyration))


def ref_bc_fields(g: Grid, E, U, B0, E0):
    R = ref()
    R.bcfields_(_d(E), _d(U), _d(_f8(B0, 4)), _d(_f8(E0, 4)), _byref_i(g.nx), _byref_i(g.ny),
                _byref_i(g.nz), *[_byref_i(1 if f else 0) for f in g.flow], _byref_i(1),
                _byref_i(1), _byref_i(1), _byref_i(0))


def ref_advance_velocity(g: Grid, rp, sp, E, B0, dt, v0, ve, eta, tB, sB, n, fluid=None,
                         gyration=None):
    R = ref()
    fluid = _i8([0] * 8 if fluid is None else fluid)
    gyration = _i8([1] * 8 if gyration is None else gyration)
    tB = np.ascontiguousarray(np.asarray(tB, np.float64).reshape(8, 4))
    sB = np.ascontiguousarray(np.asarray(sB, np.float64).reshape(8, 4))
    v0 = np.ascontiguousarray(np.asarray(v0, np.float64).reshape(8, 4))
    dxyz = np.asarray(g.dxyz, np.float64)
    R.advancevelocity_(_d(rp), _i(sp), _d(E), _d(_f8(B0, 4)), _byref_d(g.gmin[0]),
                       _byref_d(g.gmin[1]), _byref_d(g.gmin[2]), _d(dxyz), _byref_d(dt), _d(v0),
                       _d(_f8(ve, 4)), _d(_f8(eta)), _d(tB), _d(sB), _byref_i(n),
                       _byref_i(rp.shape[0]), _byref_i(g.species), _byref_i(g.nx),
                       _byref_i(g.ny), _byref_i(g.nz), _i(fluid), _i(gyration))


def ref_particles_produce(g: Grid, rp, sp, n, seed, *, kelvin, mass, nu_d, nu_i, Ek, Qn, vn, v0,
                          ve, ppc, epp, hpp, ipp, dt, n2p):
    """Seeds the runtime generator with `seed`, then calls the reference's
    ParticlesProduce(real_particles, int_particles, Lx_min, Ly_min, Lz_min,
    dxyz, kelvin, mass, nu_d, nu_i, Ek, Qn, vn, v0, ve, max_per_proc,
    num_local, ppc, epp, hpp, ipp, species, Nx_local, Ny_local, Nz_local, dt,
    n2p) (ParticlesProduce.f90:1-3).  Returns the new num_local."""
    R = ref()
    R.picard_ref_seed(ct.c_int32(seed))
    nl = ct.c_int32(n)
    v0 = np.ascontiguousarray(np.asarray(v0, np.float64).reshape(8, 4))
    R.particlesproduce_(_d(rp), _i(sp), _byref_d(g.gmin[0]), _byref_d(g.gmin[1]),
                        _byref_d(g.gmin[2]), _d(np.asarray(g.dxyz, np.float64)), _d(_f8(kelvin)),
                        _d(_f8(mass)), _d(_f8(nu_d)), _d(_f8(nu_i)), _d(_f8(Ek)), _d(_f8(Qn)),
                        _d(_f8(vn)), _d(v0), _d(_f8(ve, 4)), _byref_i(rp.shape[0]), ct.byref(nl),
                        _i(_i8(ppc)), _i(_i8(epp)), _i(_i8(hpp)), _i(_i8(ipp)),
                        _byref_i(g.species), _byref_i(g.nx), _byref_i(g.ny), _byref_i(g.nz),
                        _byref_d(dt), _d(_f8(n2p)))
    return nl.value


def _plasma_args(plasma):
    v0 = np.ascontiguousarray(np.asarray(plasma["v0"], np.float64).reshape(8, 4))
    tB = np.ascontiguousarray(np.asarray(plasma["tB4"], np.float64).reshape(8, 4))
    return (_d(_f8(plasma["kelvin"])), _d(_f8(plasma["mass"])), _d(v0), _d(_f8(plasma["ve4"], 4)),
            tB, (_i(_i8(plasma["ppc"])), _i(_i8(plasma["epp"])), _i(_i8(plasma["hpp"])),
                 _i(_i8(plasma["ipp"]))))


def ref_particles_initialize(g: Grid, rp, sp, n, seed, plasma):
    """Seeds the runtime generator with `seed`, then calls the reference's
    ParticlesInitialize(real_particles, int_particles, Lx_min, Ly_min, Lz_min,
    dxyz, kelvin, mass, v0, ve, tB, max_per_proc, num_local, ppc, epp, hpp,
    ipp, species, Nx_local, Ny_local, Nz_local) (ParticlesInitialize.f90:1-2).
    `plasma` holds kelvin, mass, v0 (8x4), ve4, tB4 (8x4), ppc, epp, hpp, ipp.
    Returns the new num_local."""
    R = ref()
    R.picard_ref_seed(ct.c_int32(seed))
    nl = ct.c_int32(n)
    kel, mass, v0, ve, tB, ints = _plasma_args(plasma)
    R.particlesinitialize_(_d(rp), _i(sp), _byref_d(g.gmin[0]), _byref_d(g.gmin[1]),
                           _byref_d(g.gmin[2]), _d(np.asarray(g.dxyz, np.float64)), kel, mass, v0,
                           ve, _d(tB), _byref_i(rp.shape[0]), ct.byref(nl), *ints,
                           _byref_i(g.species), _byref_i(g.nx), _byref_i(g.ny), _byref_i(g.nz))
    return nl.value


def ref_bc_inflow(g: Grid, rp, sp, n, seed, plasma):
    """Seeds the runtime generator with `seed`, then calls the reference's
    BCinflow(real_particles, int_particles, Lx_min, Ly_min, Lz_min, Lx_max,
    Ly_max, Lz_max, dxyz, kelvin, mass, v0, ve, max_per_proc, num_local, ppc,
    epp, hpp, ipp, species, Nx_local, Ny_local, Nz_local, xmin, xmax, ymin,
    ymax, zmin, zmax, tB, xflow, yflow, zflow) (BCinflow.f90:2-4) on one rank
    (local box = global box).  Returns the new num_local."""
    R = ref()
    R.picard_ref_seed(ct.c_int32(seed))
    nl = ct.c_int32(n)
    kel, mass, v0, ve, tB, ints = _plasma_args(plasma)
    fl = [ct.c_int32(1 if f else 0) for f in g.flow]
    R.bcinflow_(_d(rp), _i(sp), _byref_d(g.gmin[0]), _byref_d(g.gmin[1]), _byref_d(g.gmin[2]),
                _byref_d(g.gmax[0]), _byref_d(g.gmax[1]), _byref_d(g.gmax[2]),
                _d(np.asarray(g.dxyz, np.float64)), kel, mass, v0, ve, _byref_i(rp.shape[0]),
                ct.byref(nl), *ints, _byref_i(g.species), _byref_i(g.nx), _byref_i(g.ny),
                _byref_i(g.nz), _byref_d(g.gmin[0]), _byref_d(g.gmax[0]), _byref_d(g.gmin[1]),
                _byref_d(g.gmax[1]), _byref_d(g.gmin[2]), _byref_d(g.gmax[2]), _d(tB),
                ct.byref(fl[0]), ct.byref(fl[1]), ct.byref(fl[2]))
    return nl.value


def ref_uniforms(seed, count):
    """The first `count` draws of the runtime's random_number after seeding
    with `seed` (the stream ref_particles_produce consumes)."""
    R = ref()
    R.picard_ref_seed(ct.c_int32(seed))
    out = np.zeros(count, np.float64)
    R.picard_ref_uniforms(_d(out), ct.c_int64(count))
    return out


def _in_dir(path, fn):
    """Run fn with the working directory set to `path` (the reference's
    writers open save/... relative to it), restoring it afterwards."""
    old = os.getcwd()
    os.makedirs(os.path.join(path, "save", "particles"), exist_ok=True)
    os.makedirs(os.path.join(path, "save", "fields"), exist_ok=True)
    os.chdir(path)
    try:
        return fn()
    finally:
        os.chdir(old)


def ref_dump_particles(rp, sp, n, iteration, myid, path):
    """DumpParticles(real_particles, int_particles, num_local, max_per_proc,
    writeParticles_iteration, myid) (Write.f90:2-31) writing under `path`.
    Returns the file's bytes and the incremented iteration."""
    R = ref()
    it = ct.c_int32(iteration)
    buf = np.ascontiguousarray(rp, np.float64)   # (n, 8) C order == real_particles(8, n)
    _in_dir(path, lambda: R.dumpparticles_(_d(buf), _i(np.ascontiguousarray(sp, np.int32)),
                                           _byref_i(n), _byref_i(buf.shape[0]), ct.byref(it),
                                           _byref_i(myid)))
    name = os.path.join(path, "save", "particles", f"particles_{iteration:05d}_{myid:05d}")
    with open(name, "rb") as f:
        return f.read(), it.value


def ref_dump_fields(U, N, F, E, nx, ny, nz, iteration, species, myid, path):
    """DumpFields(U, N, F, E, Nx_local, Ny_local, Nz_local,
    writeFields_iteration, species, myid) (Write.f90:34-523) writing under
    `path`.  Arrays in this package's C-order views.  Returns {name: bytes}."""
    R = ref()
    it = ct.c_int32(iteration)
    _in_dir(path, lambda: R.dumpfields_(_d(np.ascontiguousarray(U)), _d(np.ascontiguousarray(N)),
                                        _d(np.ascontiguousarray(F)), _d(np.ascontiguousarray(E)),
                                        _byref_i(nx), _byref_i(ny), _byref_i(nz), ct.byref(it),
                                        _byref_i(species), _byref_i(myid)))
    out = {}
    d = os.path.join(path, "save", "fields")
    for name in sorted(os.listdir(d)):
        if name.endswith(f"_{iteration:05d}_{myid:05d}"):
            with open(os.path.join(d, name), "rb") as f:
                out[name] = f.read()
    return out


def _byref_d(x):
    return ct.byref(ct.c_double(float(x)))


def ref_write_warnings(cfg, d, path, procs=(1, 1, 1)):
    """WriteWarnings(mass, charge, eta, kelvin, density, v0, ve, dt, n2p,
    dxyz, B0, Nx, Ny, Nz, xmin..zmax, iprocs, jprocs, kprocs, gyration,
    xflow, yflow, zflow, Qn, vn, Ek, nu_i, nu_d, ppc) (WriteWarnings.f90:4-...)
    as picard.f90:488-489 calls it, writing under `path`.  Returns the text it
    printed on standard output and the bytes of save/parameters.dat and
    save/warnings.dat.  Logicals are passed as 4-byte integers 1/0."""
    import tempfile
    R = ref()
    lg = lambda b: ct.byref(ct.c_int32(1 if b else 0))
    v0 = np.ascontiguousarray(d.v0, np.float64)           # (8, 4) C == v0(4, 8)
    args = [_d(np.ascontiguousarray(d.mass, np.float64)), _d(np.ascontiguousarray(d.charge, np.float64)),
            _d(np.ascontiguousarray(d.eta, np.float64)), _d(np.ascontiguousarray(d.kelvin, np.float64)),
            _d(np.ascontiguousarray(d.density, np.float64)), _d(v0),
            _d(np.ascontiguousarray(d.ve, np.float64)), _byref_d(d.dt),
            _d(np.ascontiguousarray(d.n2p, np.float64)), _d(np.ascontiguousarray(d.dxyz[:3], np.float64)),
            _d(np.ascontiguousarray(d.B0, np.float64)),
            _byref_i(cfg.nx), _byref_i(cfg.ny), _byref_i(cfg.nz),
            _byref_d(cfg.gmin[0]), _byref_d(cfg.gmax[0]), _byref_d(cfg.gmin[1]),
            _byref_d(cfg.gmax[1]), _byref_d(cfg.gmin[2]), _byref_d(cfg.gmax[2]),
            _byref_i(procs[0]), _byref_i(procs[1]), _byref_i(procs[2]),
            _i(np.array([1 if g else 0 for g in cfg.gyration], np.int32)),
            lg(cfg.flow[0]), lg(cfg.flow[1]), lg(cfg.flow[2]),
            _d(_f8(cfg.Qn)), _d(_f8(cfg.vn)), _d(_f8(cfg.Ek)), _d(_f8(cfg.nu_i)), _d(_f8(cfg.nu_d)),
            _i(_i8(d.ppc))]
    keep = args                                            # arrays stay alive for the call
    with tempfile.TemporaryFile() as cap:
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(cap.fileno(), 1)
        try:
            _in_dir(path, lambda: R.writewarnings_(*keep))
            R.picard_ref_flush()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        cap.seek(0)
        text = cap.read()
    out = []
    for name in ("parameters.dat", "warnings.dat"):
        with open(os.path.join(path, "save", name), "rb") as f:
            out.append(f.read())
    return text, out[0], out[1]
