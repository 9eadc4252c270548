"""Restatement of the reference's Haser-model source, ParticlesProduce
(ParticlesProduce.f90:1-113), in numpy/Python loops for small grids.

TEST INFRASTRUCTURE ONLY (see the package docstring): the checker of the
device source, never part of the product path.

The uniform draws are supplied by a provider, so the same restatement checks
both the reference (the runtime's sequential random_number stream:
``SequentialDraws``) and the device form (one counter-based stream per
(cell, species) and per ion: ``CellStreamDraws``).
"""
from __future__ import annotations

import math

import numpy as np

TAU = 6.2831853071795864769252867665590


class SequentialDraws:
    """random_number as the reference calls it: one flat stream, rnd for
    every (cell, species) with n2p > 0 (ParticlesProduce.f90:69), then rndp(7)
    for every created ion (:74)."""

    def __init__(self, u):
        self.u = np.asarray(u, np.float64)
        self.pos = 0

    def _take(self, k):
        if self.pos + k > self.u.size:
            raise IndexError("uniform stream exhausted")
        out = self.u[self.pos:self.pos + k]
        self.pos += k
        return out

    def count(self, cs, ss):
        return float(self._take(1)[0])

    def ion(self, cs, ss, pp):
        return self._take(7)


class CellStreamDraws:
    """Draws keyed per (cell, species) index cs: count(cs) and ion(cs, pp)
    come from caller-supplied functions (e.g. a Philox restatement)."""

    def __init__(self, count_fn, ion_fn, producing):
        self.count_fn = count_fn
        self.ion_fn = ion_fn
        self.producing = producing      # 1-based species that can produce

    def count(self, cs, ss):
        return self.count_fn(cs, ss)

    def ion(self, cs, ss, pp):
        return self.ion_fn(cs, ss, pp)


def _sq(a):
    a = float(a)
    return a * a


def speeds(mass, Ek, epp, hpp, species):
    """ue, uh, ui of ParticlesProduce.f90:26-36."""
    ue, uh, ui = np.zeros(8), np.zeros(8), np.zeros(8)
    m = mass
    for s in range(species):
        ue[s] = math.sqrt(2.0 * m[s] * Ek[s] * (1.0 / (m[0] * (m[s] + m[0]))))
        uh[s] = math.sqrt(2.0 * m[s] * Ek[s] * (1.0 / (m[1] * (m[s] + m[1]))))
        if epp[s] > 0:
            ui[s] = math.sqrt(2.0 * m[0] * Ek[s] * (1.0 / (m[s] * (m[s] + m[0]))))
        elif hpp[s] > 0:
            ui[s] = math.sqrt(2.0 * m[1] * Ek[s] * (1.0 / (m[s] * (m[s] + m[1]))))
    return ue, uh, ui


def cell_distance(x0, y0, z0, dx, dy, dz):
    """r1, r2 of ParticlesProduce.f90:41-59: distance of the cell centre from
    the origin, or the mean over the 8 quarter points when (4 r1)^3 < dV."""
    dV = dx * dy * dz
    r2 = _sq(x0 + 0.5 * dx) + _sq(y0 + 0.5 * dy) + _sq(z0 + 0.5 * dz)
    r1 = math.sqrt(r2)
    if (4.0 * r1) * (4.0 * r1) * (4.0 * r1) < dV:
        acc = 0.0
        for fx, fy, fz in ((0.25, 0.25, 0.25), (0.25, 0.25, 0.75), (0.25, 0.75, 0.25),
                           (0.75, 0.25, 0.25), (0.25, 0.75, 0.75), (0.75, 0.25, 0.75),
                           (0.75, 0.75, 0.25), (0.75, 0.75, 0.75)):
            acc = acc + math.sqrt(_sq(x0 + fx * dx) + _sq(y0 + fy * dy) + _sq(z0 + fz * dz))
        r1 = acc * 0.125
        r2 = r1 * r1
    return r1, r2


def expected_count(r1, r2, s, Qn, vn, nu_d, nu_i, n2p, dV, rate_scale=1.0):
    """nu_i * nden * dV / n2p of ParticlesProduce.f90:62-70 (times rate_scale)."""
    if vn[s] > 0.0:
        nden = Qn[s] / (2.0 * TAU * r2 * vn[s]) * math.exp(-nu_d[s] * r1 / vn[s])
    else:
        nden = 0.0
    return nu_i[s] * nden * dV / n2p[s] * rate_scale


def produce(nx, ny, nz, lmin, dxyz, species, *, mass, Ek, Qn, vn, nu_d, nu_i, n2p, epp, hpp, dt,
            draws, rate_scale=1.0, producing_only=False):
    """Particles appended by one ParticlesProduce call, as (rp[n, 8], sp[n]).

    Loops as ParticlesProduce.f90:39-107: cells kk, jj, ii (x fastest), species
    ss, ions pp.  With ``producing_only`` the count draw is skipped for
    species that cannot produce (vn, n2p, nu_i or Qn zero): the device form
    draws only for producing species; the reference draws whenever n2p > 0."""
    dx, dy, dz = dxyz
    dV = dx * dy * dz
    ue, uh, ui = speeds(mass, Ek, epp, hpp, species)
    prod = [s for s in range(species)
            if vn[s] > 0.0 and n2p[s] > 0.0 and nu_i[s] > 0.0 and Qn[s] > 0.0]
    rows, spc = [], []
    cell = 0
    for kk in range(2, nz + 2):
        for jj in range(2, ny + 2):
            for ii in range(2, nx + 2):
                x0 = lmin[0] + dx * (ii - 2)
                y0 = lmin[1] + dy * (jj - 2)
                z0 = lmin[2] + dz * (kk - 2)
                r1, r2 = cell_distance(x0, y0, z0, dx, dy, dz)
                for s in range(species):
                    if producing_only:
                        if s not in prod:
                            continue
                        cs = cell * len(prod) + prod.index(s)
                    else:
                        cs = None
                    if not n2p[s] > 0.0:
                        continue
                    lam = expected_count(r1, r2, s, Qn, vn, nu_d, nu_i, n2p, dV, rate_scale)
                    num = int(lam + draws.count(cs, s + 1))
                    for pp in range(num):
                        rnd = draws.ion(cs, s + 1, pp)
                        x = np.array([x0 + dx * rnd[0], y0 + dy * rnd[1], z0 + dz * rnd[2]])
                        rn = math.sqrt(_sq(x[0]) + _sq(x[1]) + _sq(x[2]))
                        dn = math.sqrt(_sq(rnd[3]) + _sq(rnd[4]) + _sq(rnd[5]))
                        v = np.array([vn[s] * x[c] / rn + ui[s] * rnd[3 + c] / dn for c in range(3)])
                        x = np.array([x[c] + v[c] * dt * rnd[6] for c in range(3)])
                        rows.append(np.concatenate([x, v, [0.0, 0.0]]))
                        spc.append(s + 1)
                        for kind, cnt, u in ((1, epp[s], ue[s]), (2, hpp[s], uh[s])):
                            for _ in range(int(cnt)):
                                w = np.array([v[c] - (u + ui[s]) * rnd[3 + c] / dn
                                              for c in range(3)])
                                y = np.array([x[c] - v[c] * dt * rnd[6] + w[c] * dt * rnd[6]
                                              for c in range(3)])
                                rows.append(np.concatenate([y, w, [0.0, 0.0]]))
                                spc.append(kind)
                cell += 1
    if not rows:
        return np.zeros((0, 8)), np.zeros(0, np.int32)
    return np.array(rows), np.array(spc, np.int32)
