"""Restatement of the reference's host particle sources, ParticlesInitialize
(ParticlesInitialize.f90:54-119) and BCinflow (BCinflow.f90:22-470), with the
Box-Muller Maxwellian of Tools.f90:1-48, as scalar Python loops for small
grids.

TEST INFRASTRUCTURE ONLY (see the package docstring): the checker of the
device sources, never part of the product path.

The uniform draws come from a provider, so the same restatement checks both
the reference (the runtime's one sequential random_number stream:
``SequentialDraws``) and the device form (one counter-based stream per ion,
keyed by the source call and the ion's slot within it: ``PerIonDraws``).
Transcendentals are Python's math.log/sin/sqrt (the C library's), evaluated
in the reference's operand order.
"""
from __future__ import annotations

import math

import numpy as np

TAU = 6.2831853071795864769252867665590      # Tools.f90:35
KB = 1.380650e-23                            # Tools.f90:9


class SequentialDraws:
    """random_number as the reference calls it: one flat stream for the whole
    call, whatever the face or slot."""

    def __init__(self, u):
        self.u = np.asarray(u, np.float64)
        self.pos = 0

    def ion(self, face, local):
        return self

    def uniform(self):
        if self.pos >= self.u.size:
            raise IndexError("uniform stream exhausted")
        v = float(self.u[self.pos])
        self.pos += 1
        return v


class _Seq:
    def __init__(self, u):
        self.u = u
        self.pos = 0

    def uniform(self):
        if self.pos >= len(self.u):
            raise IndexError("per-ion stream exhausted")
        v = float(self.u[self.pos])
        self.pos += 1
        return v


class PerIonDraws:
    """One stream per ion: `fn(face, local)` returns the stream's uniforms (an
    array long enough for the ion and its companions), where `face` numbers
    the source's cell blocks (0 for ParticlesInitialize; +x, -x, +y, -y, +z,
    -z as 0..5 for BCinflow) and `local` is the ion's slot counted from the
    block's first particle."""

    def __init__(self, fn):
        self.fn = fn

    def ion(self, face, local):
        return _Seq(self.fn(face, local))


def normal(st) -> float:
    """Tools.f90:30-48: r1, then r2 until r2 > 0 with -2 ln r2 >= 0;
    sqrt(-2 ln r2) sin(tau r1)."""
    r1 = st.uniform()
    while True:
        r2 = st.uniform()
        if r2 > 0.0 and -2.0 * math.log(r2) >= 0.0:
            break
    return math.sqrt(-2.0 * math.log(r2)) * math.sin(TAU * r1)


def gaussian(st, v, t, m):
    """Tools.f90:1-28: a drifting Maxwellian velocity (bulk v, temperature t,
    mass m), or v itself when t or m is not positive."""
    if t > 0.0 and m > 0.0:
        a = m / (2.0 * KB * t)
        pw = math.sqrt(1.0 / a)                # dsqrt(a**(-1)), :14
        u = [normal(st), normal(st), normal(st)]
        s2 = math.sqrt(2.0)
        return [u[c] * pw / s2 + v[c] for c in range(3)]
    return [float(v[0]), float(v[1]), float(v[2])]


def half_rotate(u, ve, tB):
    """ParticlesInitialize.f90:76-84 (BCinflow.f90:73-81): the dt/2 rotation
    about B in the ve frame with the norm restored."""
    vmin = [u[c] - ve[c] for c in range(3)]
    vmid = [vmin[0] + vmin[1] * tB[2] - vmin[2] * tB[1],
            vmin[1] + vmin[2] * tB[0] - vmin[0] * tB[2],
            vmin[2] + vmin[0] * tB[1] - vmin[1] * tB[0]]
    # squares as products (x**2 of a real with an integer exponent is x*x)
    smid = vmid[0] * vmid[0] + vmid[1] * vmid[1] + vmid[2] * vmid[2]
    if smid > 0.0:
        f = math.sqrt((vmin[0] * vmin[0] + vmin[1] * vmin[1] + vmin[2] * vmin[2]) / smid)
        vmid = [vmid[c] * f for c in range(3)]
    return [vmid[c] + ve[c] for c in range(3)]


def _cell(out, draws, face, local0, x0, y0, z0, dxyz, first_species, plasma):
    """The particles of one cell (species loop, ions, their electrons and
    protons) appended to `out`; returns the slots used."""
    species, ipp, epp, hpp = plasma["species"], plasma["ipp"], plasma["epp"], plasma["hpp"]
    kelvin, mass, v0, ve, tB = (plasma["kelvin"], plasma["mass"], plasma["v0"], plasma["ve"],
                                plasma["tB"])
    used = 0
    for ss in range(first_species, species + 1):
        s = ss - 1
        for _ in range(int(ipp[s])):
            st = draws.ion(face, local0 + used)
            x = x0 + dxyz[0] * st.uniform()
            y = y0 + dxyz[1] * st.uniform()
            z = z0 + dxyz[2] * st.uniform()
            u = half_rotate(gaussian(st, v0[s], kelvin[s], mass[s]), ve, tB[s])
            out.append((ss, x, y, z, u))
            for _e in range(int(epp[s])):
                ue = half_rotate(gaussian(st, v0[s], kelvin[0], mass[0]), ve, tB[0])
                out.append((1, x, y, z, ue))
            for _h in range(int(hpp[s])):
                uh = half_rotate(gaussian(st, v0[s], kelvin[1], mass[1]), ve, tB[1])
                out.append((2, x, y, z, uh))
            used += 1 + int(epp[s]) + int(hpp[s])
    return used


def _emit(out, draws, face, cells, dxyz, first_species, plasma):
    local = 0
    for (x0, y0, z0) in cells:
        local += _cell(out, draws, face, local, x0, y0, z0, dxyz, first_species, plasma)


def _pack(out):
    n = len(out)
    rp = np.zeros((n, 8))
    sp = np.zeros(n, np.int32)
    for q, (s, x, y, z, u) in enumerate(out):
        rp[q, :3] = (x, y, z)
        rp[q, 3:6] = u
        sp[q] = s
    return rp, sp


def particles_initialize(g, plasma, draws):
    """ParticlesInitialize.f90:54-119: every interior cell (kk, jj, ii loops,
    x fastest), lower edge Lmin + dxyz*(i-2), species 1..species, ipp ions
    each followed by epp electrons and hpp protons at the ion's position."""
    dx = g.dxyz
    cells = [(g.gmin[0] + dx[0] * (ii - 2), g.gmin[1] + dx[1] * (jj - 2),
              g.gmin[2] + dx[2] * (kk - 2))
             for kk in range(2, g.nz + 2) for jj in range(2, g.ny + 2)
             for ii in range(2, g.nx + 2)]
    out = []
    _emit(out, draws, 0, cells, dx, 1, plasma)
    return _pack(out)


def bc_inflow(g, plasma, draws):
    """BCinflow.f90 on one rank (the local box is the global box, so every
    flow face is owned, :22-50): guard cells of +x (x0 = Lx_max, :53-121) and
    -x (x0 = Lx_min - dx, :124-191), jj and kk widened by the y/z flow flags,
    species from 1; then +y (:194-263; kk up to Nz+1+zflow_pos), -y (:264-333;
    kk from 2-zflow_neg), +z (:334-403; jj from 2-yflow_neg), -z (:404-470;
    jj up to Ny+1+yflow_pos), species from 2."""
    dx = g.dxyz
    xf, yf, zf = (bool(f) for f in g.flow)
    yn = yp = 1 if yf else 0
    zn = zp = 1 if zf else 0
    lx, ly, lz = g.gmin
    ux, uy, uz = g.gmax
    out = []
    if xf:
        for face, x0 in ((0, ux), (1, lx - dx[0])):
            cells = [(x0, ly + dx[1] * (jj - 2), lz + dx[2] * (kk - 2))
                     for kk in range(2 - zn, g.nz + 2 + zp) for jj in range(2 - yn, g.ny + 2 + yp)]
            _emit(out, draws, face, cells, dx, 1, plasma)
    if yf:
        cells = [(lx + dx[0] * (ii - 2), uy, lz + dx[2] * (kk - 2))
                 for kk in range(2, g.nz + 2 + zp) for ii in range(2, g.nx + 2)]
        _emit(out, draws, 2, cells, dx, 2, plasma)
        cells = [(lx + dx[0] * (ii - 2), ly - dx[1], lz + dx[2] * (kk - 2))
                 for kk in range(2 - zn, g.nz + 2) for ii in range(2, g.nx + 2)]
        _emit(out, draws, 3, cells, dx, 2, plasma)
    if zf:
        cells = [(lx + dx[0] * (ii - 2), ly + dx[1] * (jj - 2), uz)
                 for jj in range(2 - yn, g.ny + 2) for ii in range(2, g.nx + 2)]
        _emit(out, draws, 4, cells, dx, 2, plasma)
        cells = [(lx + dx[0] * (ii - 2), ly + dx[1] * (jj - 2), lz - dx[2])
                 for jj in range(2, g.ny + 2 + yp) for ii in range(2, g.nx + 2)]
        _emit(out, draws, 5, cells, dx, 2, plasma)
    return _pack(out)
