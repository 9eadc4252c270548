"""Pins the oracle functions whose reference files need MPI (and so cannot be
built here) by closed forms, invariants and independent formulations:

* ParticlesTransfer  -- wrap invariants and exact +-L shifts
* ParticlesGridify   -- textbook CIC (node positions + hat weights), mass
                        conservation, periodic/flow guard folds, C == 0 for
                        co-located ions and electrons
* BCpotentials       -- point charge from explicit coordinates and images
* CalcEpotential     -- single Fourier mode: iterates (1-lambda^k) u* and the
                        predicted stop sweep including the stale-halo residual
* CalcEfield         -- exact central difference of a quadratic, periodic
                        copies and zero flow guards
"""
import itertools

import numpy as np
import pytest

import oracle
from tests.helpers import EPS0, random_particles

FLOWS = list(itertools.product([False, True], repeat=3))


# ---------------------------------------------------------------------------
# ParticlesTransfer

def test_transfer_wrap_invariants(orc):
    rng = np.random.default_rng(1)
    g = oracle.Grid(6, 6, 6, 2, flow=(True, False, False))
    rp, _ = random_particles(rng, g, 4000, spread=3.0)
    before = rp.copy()
    oracle.particles_transfer(g, rp)
    L = np.array(g.gmax) - np.array(g.gmin)
    for c in range(3):
        d = rp[:, c] - before[:, c]
        moved = d != 0
        # a move is exactly +-L (both operands exact multiples of dx here)
        assert np.all(np.isin(np.round(d[moved] / L[c]), [-1.0, 1.0]))
        assert np.all(rp[:, c] >= g.gmin[c]) and np.all(rp[:, c] <= g.gmax[c])
        inside = (before[:, c] >= g.gmin[c]) & (before[:, c] <= g.gmax[c])
        assert np.array_equal(rp[inside, c], before[inside, c])
    again = rp.copy()
    oracle.particles_transfer(g, again)
    assert np.array_equal(again, rp)                     # idempotent
    assert np.array_equal(rp[:, 3:], before[:, 3:])      # velocities, slots 7-8 untouched


# ---------------------------------------------------------------------------
# ParticlesGridify

def _textbook_cic(g, x, wrap):
    """Node weights of one particle: nodes at Lmin + (n - 1.5) dx (1-based n),
    hat function of width dx; periodic dims wrap onto interior nodes 2..N+1."""
    out = {}
    idx, wts = [], []
    for c in range(3):
        n = (g.nx, g.ny, g.nz)[c]
        dx = g.dxyz[c]
        t = (x[c] - g.gmin[c]) / dx + 1.5     # continuous node coordinate
        lo = int(np.floor(t))
        f = t - lo
        nodes = [lo, lo + 1]
        w = [1.0 - f, f]
        if wrap[c]:
            nodes = [2 + (m - 2) % n for m in nodes]
        idx.append(nodes)
        wts.append(w)
    for a in range(2):
        for b in range(2):
            for d in range(2):
                key = (idx[0][a], idx[1][b], idx[2][d])
                out[key] = out.get(key, 0.0) + wts[0][a] * wts[1][b] * wts[2][d]
    return out


@pytest.mark.parametrize("flow", [(False, False, False), (True, False, False), (True, True, True)])
def test_gridify_single_particles_textbook(orc, flow):
    rng = np.random.default_rng(2)
    g = oracle.Grid(5, 6, 4, 2, flow=flow)
    n2p = np.array([3.5, 1.25] + [0.0] * 6)
    charge = np.array([-1.0, 1.0] + [0.0] * 6)
    for t in range(40):
        x = [g.gmin[c] + (g.gmax[c] - g.gmin[c]) * rng.random() for c in range(3)]
        if t < 6:   # hug the faces: guard deposits, folds and wraps
            c = t % 3
            x[c] = g.gmin[c] + 0.1 * g.dxyz[c] if t < 3 else g.gmax[c] - 0.2 * g.dxyz[c]
        rp = np.zeros((1, 8))
        rp[0, :3] = x
        sp = np.array([2], np.int32)
        N, C, skipped = oracle.particles_gridify(g, rp, sp, charge, n2p)
        assert skipped == 0
        expect = _textbook_cic(g, x, [not f for f in flow])
        got = N[..., 1]
        # compare on interior nodes (guards of periodic dims are copies; of flow dims, zero)
        for k in range(2, g.nz + 2):
            for j in range(2, g.ny + 2):
                for i in range(2, g.nx + 2):
                    # a flow-dim guard weight folds into the first interior node
                    e = 0.0
                    for (a, b, d), w in expect.items():
                        aa, bb, dd = a, b, d
                        if flow[0]:
                            aa = 2 if a == 1 else (g.nx + 1 if a == g.nx + 2 else a)
                        if flow[1]:
                            bb = 2 if b == 1 else (g.ny + 1 if b == g.ny + 2 else b)
                        if flow[2]:
                            dd = 2 if d == 1 else (g.nz + 1 if d == g.nz + 2 else d)
                        if (aa, bb, dd) == (i, j, k):
                            e += w
                    v = got[k - 1, j - 1, i - 1]
                    assert abs(v - e * n2p[1]) <= 1e-13 * n2p[1], (t, i, j, k, v, e)
        assert np.all(N[..., 0] == 0.0)
        # guard handling: periodic guards are copies of their images, flow guards are zero
        for c, f in enumerate(flow):
            ax_ = 2 - c
            lo = np.take(got, 0, axis=ax_)
            hi = np.take(got, got.shape[ax_] - 1, axis=ax_)
            if f:
                assert np.all(lo == 0.0) and np.all(hi == 0.0)
            else:
                n = got.shape[ax_] - 2
                assert np.array_equal(lo, np.take(got, n, axis=ax_))
                assert np.array_equal(hi, np.take(got, 1, axis=ax_))
        assert np.array_equal(C, N[..., 1] * charge[1])


@pytest.mark.parametrize("flow", FLOWS)
def test_gridify_mass_and_neutrality(orc, flow):
    rng = np.random.default_rng(3)
    g = oracle.Grid(6, 5, 7, 2, flow=flow)
    n = 3000
    ions, _ = random_particles(rng, g, n)
    # interleave co-located electron/ion pairs (ParticlesInitialize.f90:83-86)
    rp = np.repeat(ions, 2, axis=0)
    sp = np.tile(np.array([2, 1], np.int32), n)
    n2p = np.array([7.0e5, 7.0e5] + [0.0] * 6)
    charge = np.array([-1.60217646e-19, 1.60217646e-19] + [0.0] * 6)
    N, C, skipped = oracle.particles_gridify(g, rp, sp, charge, n2p)
    assert skipped == 0
    inner = N[1:-1, 1:-1, 1:-1]
    for s in range(2):
        total = inner[..., s].sum()
        assert abs(total - n2p[s] * n) <= 1e-12 * n2p[s] * n
    assert np.array_equal(N[..., 0], N[..., 1])
    assert np.all(C == 0.0)


# ---------------------------------------------------------------------------
# BCpotentials

def _node_pos(g, c, m):
    """Position of 1-based node m along dim c (cell-centred nodes)."""
    return g.gmin[c] + (m - 1.5) * g.dxyz[c]


def _point_potential(g, q, cell, face_node):
    """q/(4 pi eps0) * [1/r + the nearest periodic images kept by CalcDistance]."""
    delta = [abs(_node_pos(g, c, cell[c]) - _node_pos(g, c, face_node[c])) for c in range(3)]
    L = [g.gmax[c] - g.gmin[c] for c in range(3)]
    per = [not f for f in g.flow]
    xf, yf, zf = g.flow
    # image combinations (which dims use L - delta), per CalcDistance.f90:48-78
    if xf and not yf and not zf:
        combos = [(0, 1, 0), (0, 0, 1), (0, 1, 1)]
    elif xf and not yf:
        combos = [(0, 1, 0)]
    elif xf and not zf:
        combos = [(0, 0, 1)]
    elif xf:
        combos = []
    elif yf and not zf:
        combos = [(1, 0, 0), (0, 0, 1), (1, 0, 1)]
    elif yf:
        combos = [(1, 0, 0)]
    elif zf:
        combos = [(1, 0, 0), (0, 1, 0), (1, 1, 0)]
    else:
        combos = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)]
    for cmb in combos:
        assert all(per[c] or not cmb[c] for c in range(3))
    tot = 1.0 / np.sqrt(sum(d * d for d in delta))
    for cmb in combos:
        dd = [(L[c] - delta[c]) if cmb[c] else delta[c] for c in range(3)]
        tot += 1.0 / np.sqrt(sum(d * d for d in dd))
    return q * tot / (4.0 * np.pi * EPS0)


@pytest.mark.parametrize("flow", [f for f in FLOWS if any(f)])
def test_bc_potentials_point_charge(orc, flow):
    g = oracle.Grid(5, 4, 6, 2, flow=flow, gmin=(-25.0, -20.0, -30.0), gmax=(25.0, 20.0, 30.0))
    D = oracle.calc_distance(g.nx, g.ny, g.nz, g.dxyz, flow)
    q = 3.0e-16
    for cell in [(2, 2, 2), (4, 3, 6), (6, 5, 7)]:
        C = np.zeros(g.shape)
        C[cell[2] - 1, cell[1] - 1, cell[0] - 1] = q
        U = oracle.bc_potentials(g, C, D)
        dims = (g.nx, g.ny, g.nz)
        checked = 0
        for c in range(3):
            if not flow[c]:
                continue
            for face in (1, dims[c] + 2):
                a, b = [d for d in range(3) if d != c]
                for ma in range(2, dims[a] + 2):
                    for mb in range(2, dims[b] + 2):
                        node = [0, 0, 0]
                        node[c], node[a], node[b] = face, ma, mb
                        got = U[node[2] - 1, node[1] - 1, node[0] - 1]
                        want = _point_potential(g, q, cell, node)
                        assert abs(got - want) <= 2e-14 * abs(want), (cell, node, got, want)
                        checked += 1
        assert checked > 0
        # everything that is not a flow-face interior node is zero
        mask = np.ones(g.shape, bool)
        for c in range(3):
            if flow[c]:
                a, b = [d for d in range(3) if d != c]
                for face in (0, (g.nx, g.ny, g.nz)[c] + 1):
                    sl = [slice(1, -1)] * 3
                    sl[2 - c] = face
                    mask[tuple(sl)] = False
        assert np.all(U[mask] == 0.0)


def test_bc_potentials_linear(orc):
    rng = np.random.default_rng(4)
    g = oracle.Grid(5, 4, 6, 2, flow=(True, True, False))
    D = oracle.calc_distance(g.nx, g.ny, g.nz, g.dxyz, g.flow)
    C1 = rng.normal(size=g.shape) * 1e-15
    C2 = rng.normal(size=g.shape) * 1e-15
    U1 = oracle.bc_potentials(g, C1, D)
    U2 = oracle.bc_potentials(g, C2, D)
    U12 = oracle.bc_potentials(g, C1 + C2, D)
    scale = np.abs(U1).max() + np.abs(U2).max()
    assert np.abs(U12 - U1 - U2).max() <= 1e-12 * scale


# ---------------------------------------------------------------------------
# CalcEpotential

def _mode_setup(nx=9, ny=4, nz=3, m=1, amp=1.0):
    g = oracle.Grid(nx, ny, nz, 2, flow=(True, False, False),
                    gmin=(-5.0 * nx, -5.0 * ny, -5.0 * nz), gmax=(5.0 * nx, 5.0 * ny, 5.0 * nz))
    dx, dy, dz = g.dxyz
    ii = np.arange(1, nx + 1)
    v = np.sin(np.pi * m * ii / (nx + 1))
    b = amp * v                                   # b = -C/(dV eps0)
    dV = dx * dy * dz
    C = np.zeros(g.shape)
    C[1:-1, 1:-1, 1:-1] = (-b * dV * EPS0)[None, None, :]
    den = 1.0 / ((dx * dy) ** 2 + (dx * dz) ** 2 + (dy * dz) ** 2)
    ax, ay, az = 0.5 * den * (dy * dz) ** 2, 0.5 * den * (dx * dz) ** 2, 0.5 * den * (dx * dy) ** 2
    af = 0.5 * den * (dx * dy * dz) ** 2
    th = np.pi * m / (nx + 1)
    lam = 2 * ax * np.cos(th) + 2 * ay + 2 * az
    return g, C, v, b, amp, lam, af, th


def test_epotential_mode_iterates(orc):
    g, C, v, b, amp, lam, af, th = _mode_setup()
    for maxit in (0, 1, 5, 20):
        U = np.zeros(g.shape)
        sweeps, _ = oracle.calc_epotential(g, U, C, 0.0, maxit)
        K = 2 * maxit + 1                        # CalcEpotential.f90:58 with no exit
        assert sweeps == K
        ck = -af * amp * (1.0 - lam ** K) / (1.0 - lam)
        want = ck * v
        got = U[1:-1, 1:-1, 1:-1]
        assert np.allclose(got, want[None, None, :], rtol=1e-12, atol=1e-12 * abs(ck))
        # periodic halos are copies, flow faces keep their (zero) values
        assert np.array_equal(U[0, 1:-1, 1:-1], U[-2, 1:-1, 1:-1])
        assert np.array_equal(U[1:-1, 0, 1:-1], U[1:-1, -2, 1:-1])
        assert np.all(U[:, :, 0] == 0.0) and np.all(U[:, :, -1] == 0.0)


def _predicted_stop(g, v, amp, lam, af, th, maxerr, maxit):
    dx, dy, dz = g.dxyz
    ny, nz = g.ny, g.nz
    c = [0.0]
    for k in range(1, 2 * maxit + 2):
        c.append(lam * c[-1] - af * amp)
    jw = np.array([(1 if j == 2 else 0) + (1 if j == ny + 1 else 0) for j in range(2, ny + 2)], float)
    kw = np.array([(1 if k == 2 else 0) + (1 if k == nz + 1 else 0) for k in range(2, nz + 2)], float)
    stale = jw[None, :] / dy ** 2 + kw[:, None] / dz ** 2    # [nz, ny]
    vv = np.sum(v * v)
    bnorm = np.sqrt(amp * amp * vv * ny * nz)
    ratios = [1.0]
    for k in range(1, 2 * maxit + 2):
        prev = c[k - 1]
        per_yz = amp - (2 * np.cos(th) - 2) * c[k] / dx ** 2 - stale * (prev - c[k])
        ratios.append(np.sqrt(vv * np.sum(per_yz ** 2)) / bnorm)
    for ss in range(1, maxit + 1):
        if ratios[ss - 1] <= maxerr:
            return ss - 1, ratios
    return 2 * maxit + 1, ratios


@pytest.mark.parametrize("maxerr", [1e-2, 1e-4, 1e-6])
def test_epotential_mode_stop_rule(orc, maxerr):
    g, C, v, b, amp, lam, af, th = _mode_setup(nx=11, ny=3, nz=4, m=2, amp=2.5)
    maxit = 4000
    want, ratios = _predicted_stop(g, v, amp, lam, af, th, maxerr, maxit)
    assert want < 2 * maxit + 1
    # the prediction is not a near-tie
    assert abs(ratios[want] - maxerr) > 1e-6 * maxerr and ratios[want - 1] > maxerr * (1 + 1e-6)
    U = np.zeros(g.shape)
    sweeps, ratio = oracle.calc_epotential(g, U, C, maxerr, maxit)
    assert sweeps == want
    assert abs(ratio - ratios[want]) <= 1e-9 * ratios[want]
    U2 = np.zeros(g.shape)
    assert oracle.calc_epotential_pairwise(g, U2, C, maxerr, maxit) == want


def test_epotential_zero_charge_exits_before_first_sweep(orc):
    g = oracle.Grid(6, 6, 6, 2)
    U = np.zeros(g.shape)
    U[:, :, 0] = 3.0            # whatever is in U stays
    before = U.copy()
    sweeps, _ = oracle.calc_epotential(g, U, np.zeros(g.shape), 1e-6, 100)
    assert sweeps == 0
    assert np.array_equal(U, before)


# ---------------------------------------------------------------------------
# CalcEfield + BCfields

@pytest.mark.parametrize("flow", FLOWS)
def test_efield_quadratic_and_halos(orc, flow):
    g = oracle.Grid(5, 6, 4, 2, flow=flow, gmin=(-20.0, -24.0, -16.0), gmax=(20.0, 24.0, 16.0))
    z, y, x = np.meshgrid(*[np.array([_node_pos(g, c, m) for m in range(1, n + 3)])
                            for c, n in ((2, g.nz), (1, g.ny), (0, g.nx))], indexing="ij")
    a = (0.25, -0.5, 0.125)
    bcoef = (1.5, 2.0, -3.0)
    U = a[0] * x * x + a[1] * y * y + a[2] * z * z + bcoef[0] * x + bcoef[1] * y + bcoef[2] * z
    E = oracle.calc_efield(g, U)
    inner = (slice(1, -1),) * 3
    for c, coord in enumerate((x, y, z)):
        want = -(2 * a[c] * coord + bcoef[c])
        got = E[..., c]
        assert np.allclose(got[inner], want[inner], rtol=1e-13, atol=1e-13)
    for c in range(3):
        axis = 2 - c
        lo = np.take(E, 0, axis=axis)
        hi = np.take(E, E.shape[axis] - 1, axis=axis)
        if flow[c]:
            assert np.all(lo == 0.0) and np.all(hi == 0.0)
        else:
            n = E.shape[axis] - 2
            assert np.array_equal(lo, np.take(E, n, axis=axis))
            assert np.array_equal(hi, np.take(E, 1, axis=axis))
    E2 = E.copy()
    oracle.bc_fields(g, E2)
    assert np.array_equal(E2, E)        # on one rank BCfields re-zeroes zeros


def test_diagnostics_closed_forms(orc):
    """The diagnostics energies (picard.f90:649-672) on fields and particles
    with closed-form sums: uniform E, constant B0, per-species speeds."""
    g = oracle.Grid(5, 4, 3, 3)
    dV = g.dxyz[0] * g.dxyz[1] * g.dxyz[2]
    ncell = 5 * 4 * 3
    E = np.zeros(g.shape + (3,))
    E[1:-1, 1:-1, 1:-1] = (1.5, -2.0, 0.25)
    E[0] = 100.0                          # guard planes do not count
    E0 = (0.5, 0.0, -0.25, 0.0)
    B0 = (1e-9, -2e-9, 5e-9, 0.0)
    rng = np.random.default_rng(4)
    n = 300
    sp = rng.integers(1, 4, n).astype(np.int32)
    speed = np.array([3.0e5, 1.0e4, 2.0e3])[sp - 1]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rp = np.zeros((n, 8))
    rp[:, 3:6] = d * speed[:, None]
    n2p = [2.0e8, 3.0e8, 5.0e7] + [0.0] * 5
    mass = [9.1e-31, 1.67e-27, 3.0e-26] + [0.0] * 5
    uP, uE, uB = oracle.diagnostics(g, rp, sp, E, n2p, mass, E0, B0)
    eps0, mu0 = 8.854187817e-12, 1.25663706143591729538505735331180e-6
    want_P = 0.5 * sum(n2p[s] * mass[s] * np.sum(speed[sp == s + 1] ** 2) for s in range(3))
    want_E = 0.5 * eps0 * dV * ncell * ((1.5 + 0.5) ** 2 + 2.0 ** 2 + 0.0 ** 2)
    want_B = 0.5 / mu0 * dV * ncell * (1e-18 + 4e-18 + 25e-18)
    assert abs(uP - want_P) <= 1e-12 * want_P
    assert abs(uE - want_E) <= 1e-12 * want_E
    assert abs(uB - want_B) <= 1e-12 * want_B


@pytest.mark.parametrize("flow", [(True, False, False), (False, True, True), (True, True, True)])
def test_gridify_all_invariants(orc, flow):
    """ParticlesGridifyAll: N and C equal ParticlesGridify's bit for bit; the
    interior sum of F(s, c) is n2p(s) * sum of v_c over species s (guard folds
    conserve, to rounding); particles with one common velocity give F = N v."""
    g = oracle.Grid(6, 5, 4, 2, flow=flow)
    rng = np.random.default_rng(8)
    n = 3000
    rp = np.zeros((n, 8))
    for c in range(3):
        rp[:, c] = rng.uniform(g.gmin[c], g.gmax[c], n)
    rp[:, 3:6] = rng.normal(0, 1e5, (n, 3))
    sp = rng.integers(1, 3, n).astype(np.int32)
    charge = [-1.6e-19, 1.6e-19] + [0.0] * 6
    n2p = [3.0e8, 2.0e8] + [0.0] * 6
    F, N, C, sk = oracle.particles_gridify_all(g, rp, sp, charge, n2p)
    N0, C0, sk0 = oracle.particles_gridify(g, rp, sp, charge, n2p)
    assert sk == sk0 == 0
    assert np.array_equal(N, N0) and np.array_equal(C, C0)
    inner = (slice(1, -1),) * 3
    for s in range(2):
        for c in range(3):
            # the eight CIC weights of a particle sum to one (F1..F8 are
            # volume fractions), so every particle adds n2p * v_c in total
            want = n2p[s] * rp[sp == s + 1, 3 + c].sum()
            got = F[inner][..., c, s].sum()
            assert abs(got - want) <= 1e-12 * n2p[s] * np.abs(rp[:, 3 + c]).sum()
    rp2 = rp.copy()
    rp2[:, 3:6] = (2.0e5, -3.0e4, 7.0e3)
    F2, N2, _, _ = oracle.particles_gridify_all(g, rp2, sp, charge, n2p)
    for c in range(3):
        assert np.allclose(F2[..., c, :], N2 * rp2[0, 3 + c], rtol=1e-13, atol=0)
