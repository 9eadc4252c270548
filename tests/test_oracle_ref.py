"""Pins the CPU oracle against the reference's own subroutines, compiled from
/root/reference by oracle/Makefile into oracle/_ref (MPI-free files only),
and against the committed golden vectors those subroutines produced.

Bit-exact equality is required everywhere in this file."""
import itertools
import os

import numpy as np
import pytest

import oracle
from tests.helpers import boris_tables, random_particles

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")

FLOWS = list(itertools.product([False, True], repeat=3))


@needs_ref
def test_advance_position_vs_ref(orc):
    rng = np.random.default_rng(11)
    g = oracle.Grid(8, 8, 8, 2)
    rp, _ = random_particles(rng, g, 5000)
    a, b = rp.copy(), rp.copy()
    oracle.ref_advance_position(a, 7.9e-7, 4000)
    oracle.advance_position(b, 7.9e-7, 4000)
    assert np.array_equal(a, b)
    assert np.array_equal(a[4000:], rp[4000:])   # only 1..num_local move


@needs_ref
@pytest.mark.parametrize("flow", FLOWS)
def test_bc_outflow_vs_ref(orc, flow):
    rng = np.random.default_rng(12)
    g = oracle.Grid(6, 7, 5, 2, flow=flow)
    n = 3000
    rp, sp = random_particles(rng, g, n, spread=2.0)
    # a few far outside the "two-neighbour" band (deleted regardless of flow)
    rp[::97, 1] = g.gmax[1] + 1.5 * (g.gmax[1] - g.gmin[1])
    # exact boundary hits are kept (strict comparisons)
    rp[5, 0] = g.gmin[0]
    rp[6, 2] = g.gmax[2]
    a, sa = rp.copy(), sp.copy()
    b, sb = rp.copy(), sp.copy()
    na = oracle.ref_bc_outflow(g, a, sa, n)
    nb, removed = oracle.bc_outflow(g, b, sb, n)
    assert na == nb
    assert na < n
    assert len(removed) == n - nb
    assert np.array_equal(a[:na], b[:nb])
    assert np.array_equal(sa[:na], sb[:nb])


@needs_ref
@pytest.mark.parametrize("flow", FLOWS)
def test_calc_distance_vs_ref(orc, flow):
    for (nx, ny, nz, dx) in [(4, 5, 3, (10.0, 10.0, 10.0)), (6, 4, 5, (3.0, 7.5, 11.25))]:
        a = oracle.ref_calc_distance(nx, ny, nz, dx, flow)
        b = oracle.calc_distance(nx, ny, nz, dx, flow)
        assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("flow", FLOWS)
def test_bc_fields_vs_ref(orc, flow):
    rng = np.random.default_rng(13)
    g = oracle.Grid(5, 6, 4, 2, flow=flow)
    E = rng.normal(size=g.shape + (3,))
    U = rng.normal(size=g.shape)
    a, b = E.copy(), E.copy()
    oracle.ref_bc_fields(g, a, U, [0, 0, 5e-9, 5e-9], [0, 2e-3, 0, 2e-3])
    oracle.bc_fields(g, b)
    assert np.array_equal(a, b)


def _velocity_case(seed, flow, fluid, gyration, species=3):
    rng = np.random.default_rng(seed)
    g = oracle.Grid(6, 5, 7, species, flow=flow)
    n = 4000
    rp, sp = random_particles(rng, g, n, species=species)
    E = rng.normal(0.0, 1e-3, size=g.shape + (3,))
    dt = 7.927e-7
    B0 = np.array([1.1e-9, -2.0e-9, 5e-9])
    mass = np.array([9.1093819e-31, 1.6726216e-27, 18 * 1.6726216e-27] + [1.6726216e-27] * 5)
    charge = np.array([-1.60217646e-19] + [1.60217646e-19] * 7)
    eta = charge / mass
    tB, sB = boris_tables(dt, B0, eta, gyration)
    ve = np.array([3.1e5, -1e4, 2e3, 0.0])
    ve[3] = np.sqrt(np.sum(ve[:3] ** 2))
    B4 = np.array([B0[0], B0[1], B0[2], np.sqrt(np.sum(B0 ** 2))])
    v0 = np.zeros((8, 4))
    return g, rp, sp, E, B4, dt, v0, ve, eta, tB, sB, n


@needs_ref
@pytest.mark.parametrize("flow", [(True, False, False), (False, False, False), (True, True, True)])
@pytest.mark.parametrize("mode", ["boris", "fluid", "nogyration"])
def test_advance_velocity_vs_ref(orc, flow, mode):
    fluid = [0] * 8
    gyration = [1] * 8
    if mode == "fluid":
        fluid = [1, 0, 1, 0, 0, 0, 0, 0]
    if mode == "nogyration":
        gyration = [0, 1, 0, 1, 1, 1, 1, 1]
        fluid = [0, 0, 1, 0, 0, 0, 0, 0]
    g, rp, sp, E, B4, dt, v0, ve, eta, tB, sB, n = _velocity_case(21, flow, fluid, gyration)
    a, b = rp.copy(), rp.copy()
    oracle.ref_advance_velocity(g, a, sp, E, B4, dt, v0, ve, eta, tB, sB, n, fluid, gyration)
    oracle.advance_velocity(g, b, sp, E, B4, dt, ve, eta, tB, sB, fluid, gyration, n)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, rp)


@needs_ref
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("kind", ["boris", "fluid"])
def test_advance_velocity_imp_exp_vs_ref(orc, mode, kind):
    """AdvanceVelocityImp (1) / AdvanceVelocityExp (2) with deltat = dt/2, as
    the main loop calls them on output steps (picard.f90:603, :616)."""
    fluid = [1, 0, 1, 0, 0, 0, 0, 0] if kind == "fluid" else [0] * 8
    gyration = [1] * 8
    g, rp, sp, E, B4, dt, v0, ve, eta, tB, sB, n = _velocity_case(33 + mode, (True, False, False),
                                                                  fluid, gyration)
    a, b = rp.copy(), rp.copy()
    oracle.ref_advance_velocity_ie(mode, g, a, sp, E, B4, 0.5 * dt, dt, v0, ve, eta, tB, n,
                                   fluid, gyration)
    oracle.advance_velocity_ie(mode, g, b, sp, E, B4, 0.5 * dt, dt, ve, eta, tB, fluid,
                               gyration, n)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, rp)


def test_imp_exp_closed_forms(orc):
    """With E = 0 both halves only rotate: |v - ve| is kept exactly (to
    rounding) and Imp followed by Exp rotates by 2*atan(|tB|) in total, the
    angle of one Boris step; with B = 0 the pair adds dt*eta*E."""
    g, rp, sp, E, B4, dt, v0, ve, eta, tB, sB, n = _velocity_case(5, (True, False, False),
                                                                  [0] * 8, [1] * 8)
    E0 = np.zeros_like(E)
    a = rp.copy()
    oracle.advance_velocity_ie(1, g, a, sp, E0, B4, 0.5 * dt, dt, ve, eta, tB, n=n)
    oracle.advance_velocity_ie(2, g, a, sp, E0, B4, 0.5 * dt, dt, ve, eta, tB, n=n)
    u0 = rp[:n, 3:6] - ve[None, :3]
    u1 = a[:n, 3:6] - ve[None, :3]
    assert np.allclose(np.linalg.norm(u1, axis=1), np.linalg.norm(u0, axis=1), rtol=1e-14)
    bhat = B4[:3] / B4[3]
    p0 = u0 - np.outer(u0 @ bhat, bhat)
    p1 = u1 - np.outer(u1 @ bhat, bhat)
    ang = np.arctan2(np.linalg.norm(np.cross(p0, p1), axis=1), np.sum(p0 * p1, axis=1))
    tb = np.linalg.norm(tB.reshape(8, 4)[sp[:n] - 1, :3], axis=1)
    # each half rotates by atan(|tB'|) with |tB'| = |tB| (deltat = dt/2): 2 atan in total
    assert np.allclose(ang, 2 * np.arctan(tb), rtol=1e-6, atol=1e-15)
    # B = 0: pure kicks, dt*eta*E in total
    tz = np.zeros((8, 4))
    b = rp.copy()
    Eu = np.zeros_like(E)
    Eu[..., 0] = 2e-3
    oracle.advance_velocity_ie(1, g, b, sp, Eu, B4 * 0, 0.5 * dt, dt, ve, eta, tz, n=n)
    oracle.advance_velocity_ie(2, g, b, sp, Eu, B4 * 0, 0.5 * dt, dt, ve, eta, tz, n=n)
    dv = b[:n, 3] - rp[:n, 3]
    assert np.allclose(dv, dt * eta[sp[:n] - 1] * 2e-3, rtol=1e-9)


# ---------------------------------------------------------------------------
# golden vectors written by tests/golden/make_golden.py from oracle/_ref

def _golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} missing")
    return np.load(path, allow_pickle=False)


def test_golden_advance_position(orc):
    z = _golden("ref_advance_position.npz")
    rp = z["rp_in"].copy()
    oracle.advance_position(rp, float(z["dt"]), int(z["n"]))
    assert np.array_equal(rp, z["rp_out"])


def test_golden_bc_outflow(orc):
    z = _golden("ref_bc_outflow.npz")
    for t in range(int(z["ncase"])):
        flow = tuple(bool(f) for f in z[f"flow{t}"])
        g = oracle.Grid(*[int(v) for v in z["dims"]], 2, flow=flow)
        rp, sp = z["rp_in"].copy(), z["sp_in"].copy()
        n, _ = oracle.bc_outflow(g, rp, sp, rp.shape[0])
        assert n == int(z[f"n{t}"])
        assert np.array_equal(rp[:n], z[f"rp_out{t}"])
        assert np.array_equal(sp[:n], z[f"sp_out{t}"])


def test_golden_calc_distance(orc):
    z = _golden("ref_calc_distance.npz")
    for t in range(int(z["ncase"])):
        nx, ny, nz = [int(v) for v in z[f"dims{t}"]]
        D = oracle.calc_distance(nx, ny, nz, z[f"dxyz{t}"], z[f"flow{t}"])
        assert np.array_equal(D, z[f"D{t}"])


def test_golden_advance_velocity(orc):
    z = _golden("ref_advance_velocity.npz")
    g = oracle.Grid(*[int(v) for v in z["dims"]], 3, flow=tuple(bool(f) for f in z["flow"]))
    rp = z["rp_in"].copy()
    oracle.advance_velocity(g, rp, z["sp"], z["E"], z["B0"], float(z["dt"]), z["ve"], z["eta"],
                            z["tB"], z["sB"], z["fluid"], z["gyration"], int(z["n"]))
    assert np.array_equal(rp, z["rp_out"])


def _produce_from_golden(z, t):
    from oracle import haser
    nx, ny, nz = [int(v) for v in z[f"dims{t}"]]
    gmin, gmax = z[f"gmin{t}"], z[f"gmax{t}"]
    dxyz = tuple((gmax[c] - gmin[c]) / float(n) for c, n in enumerate((nx, ny, nz)))
    draws = haser.SequentialDraws(z[f"u{t}"])
    rp, sp = haser.produce(nx, ny, nz, tuple(gmin), dxyz, int(z[f"species{t}"]),
                           mass=z[f"mass{t}"], Ek=z[f"Ek{t}"], Qn=z[f"Qn{t}"], vn=z[f"vn{t}"],
                           nu_d=z[f"nu{t}"], nu_i=z[f"nu{t}"], n2p=z[f"n2p{t}"],
                           epp=z[f"epp{t}"], hpp=z[f"hpp{t}"], dt=float(z[f"dt{t}"]), draws=draws)
    return rp, sp, draws


def test_golden_particles_produce():
    """The Haser restatement, fed the runtime uniform stream the reference
    consumed, reproduces the reference's appended particles bit for bit and
    consumes exactly that stream."""
    z = _golden("ref_particles_produce.npz")
    for t in range(int(z["ncase"])):
        rp, sp, draws = _produce_from_golden(z, t)
        assert draws.pos == z[f"u{t}"].size
        assert np.array_equal(sp, z[f"sp_out{t}"])
        assert np.array_equal(rp[:, :6], z[f"rp_out{t}"][:, :6])


def test_golden_particles_produce_invariants():
    """Closed-form checks of the same outputs: each ion's velocity minus the
    radial outflow vn*r/|r| has norm ui and lies in the positive octant (the
    reference draws the direction from uniforms in [0,1)); each companion
    differs from its ion by (u + ui) along that direction; the charge of every
    group cancels."""
    from oracle import haser
    z = _golden("ref_particles_produce.npz")
    for t in range(int(z["ncase"])):
        rp, sp = z[f"rp_out{t}"], z[f"sp_out{t}"]
        S = int(z[f"species{t}"])
        m, Ek, epp, hpp = z[f"mass{t}"], z[f"Ek{t}"], z[f"epp{t}"], z[f"hpp{t}"]
        ue, uh, ui = haser.speeds(m, Ek, epp, hpp, S)
        dt = float(z[f"dt{t}"])
        q = 0
        while q < len(sp):
            s = int(sp[q]) - 1
            assert s >= 2
            x, v = rp[q, :3], rp[q, 3:6]
            d = None
            # the ion's direction draw d is recovered from its companions
            group = 1 + int(epp[s]) + int(hpp[s])
            assert group >= 2
            for k in range(1, group):
                u = (ue[s] if k <= epp[s] else uh[s]) + ui[s]
                dv = v - rp[q + k, 3:6]
                assert abs(np.linalg.norm(dv) - u) <= 1e-9 * u
                assert np.all(dv >= -1e-9 * u)
                assert int(sp[q + k]) == (1 if k <= epp[s] else 2)
                d = dv / u
                # companion displacement is parallel to the velocity change
                dx = rp[q + k, :3] - x
                cross = np.cross(dx, dv)
                assert np.linalg.norm(cross) <= 1e-9 * (np.linalg.norm(dx) * u + 1e-30)
                assert np.linalg.norm(dx) <= u * dt * (1 + 1e-12)
            radial = v - ui[s] * d
            rn = np.linalg.norm(radial)
            assert abs(rn - z[f"vn{t}"][s]) <= 1e-9 * rn
            q += group


def test_golden_advance_velocity_imp_exp(orc):
    z = _golden("ref_advance_velocity_ie.npz")
    g = oracle.Grid(*[int(v) for v in z["dims"]], 3, flow=tuple(bool(f) for f in z["flow"]))
    dt = float(z["dt"])
    for mode in (1, 2):
        rp = z["rp_in"].copy()
        oracle.advance_velocity_ie(mode, g, rp, z["sp"], z["E"], z["B0"], 0.5 * dt, dt, z["ve"],
                                   z["eta"], z["tB"], z["fluid"], z["gyration"], int(z["n"]))
        assert np.array_equal(rp, z[f"rp_out{mode}"])


def _sources_from_golden(z, t):
    from oracle import sources
    nx, ny, nz, S = [int(v) for v in z[f"dims{t}"]]
    g = oracle.Grid(nx, ny, nz, S, flow=tuple(bool(f) for f in z[f"flow{t}"]),
                    gmin=tuple(z[f"gmin{t}"]), gmax=tuple(z[f"gmax{t}"]))
    pl = dict(species=S, ipp=z[f"ipp{t}"], epp=z[f"epp{t}"], hpp=z[f"hpp{t}"],
              kelvin=z[f"kelvin{t}"], mass=z[f"mass{t}"], v0=z[f"v0{t}"][:, :3],
              ve=z[f"ve4{t}"][:3], tB=z[f"tB4{t}"][:, :3])
    draws = sources.SequentialDraws(z[f"u{t}"])
    fn = sources.particles_initialize if str(z[f"kind{t}"]) == "initialize" else sources.bc_inflow
    rp, sp = fn(g, pl, draws)
    return rp, sp, draws


def test_golden_sources():
    """ParticlesInitialize and BCinflow: the restatement, fed the runtime
    uniform stream the reference consumed, reproduces the reference's
    particles bit for bit (slot order, species, positions, velocities) and
    consumes exactly that stream, at every flow combination stored (the
    z faces' cell ranges depend on the y flow, BCinflow.f90:337, :407)."""
    z = _golden("ref_sources.npz")
    kinds = set()
    for t in range(int(z["ncase"])):
        rp, sp, draws = _sources_from_golden(z, t)
        kinds.add(str(z[f"kind{t}"]))
        assert draws.pos == z[f"u{t}"].size, t
        assert np.array_equal(sp, z[f"sp_out{t}"]), t
        assert np.array_equal(rp[:, :6], z[f"rp_out{t}"]), t
    assert kinds == {"initialize", "inflow"}


def test_golden_sources_invariants():
    """Closed-form checks of the same outputs: companions sit at their ion's
    position; every particle lies in its guard or interior cell; with T = 0
    and tB = 0 the velocity reduces to v0 (Tools.f90:24-26), checked by
    rerunning the restatement with those tables."""
    from oracle import sources
    z = _golden("ref_sources.npz")
    for t in range(int(z["ncase"])):
        rp, sp = z[f"rp_out{t}"], z[f"sp_out{t}"]
        epp, hpp = z[f"epp{t}"], z[f"hpp{t}"]
        q = 0
        while q < len(sp):                   # groups: ion, epp electrons, hpp protons
            s = int(sp[q]) - 1
            grp = [1] * int(epp[s]) + [2] * int(hpp[s])
            assert list(sp[q + 1:q + 1 + len(grp)]) == grp
            assert np.all(rp[q + 1:q + 1 + len(grp), :3] == rp[q, :3])
            q += 1 + len(grp)
        nx, ny, nz, S = [int(v) for v in z[f"dims{t}"]]
        lo, hi = z[f"gmin{t}"], z[f"gmax{t}"]
        dx = (hi - lo) / np.array([nx, ny, nz])
        assert np.all(rp[:, :3] >= lo - dx) and np.all(rp[:, :3] <= hi + dx)
        g = oracle.Grid(nx, ny, nz, S, flow=tuple(bool(f) for f in z[f"flow{t}"]),
                        gmin=tuple(lo), gmax=tuple(hi))
        pl = dict(species=S, ipp=z[f"ipp{t}"], epp=z[f"epp{t}"], hpp=z[f"hpp{t}"],
                  kelvin=np.zeros(8), mass=z[f"mass{t}"], v0=z[f"v0{t}"][:, :3],
                  ve=z[f"ve4{t}"][:3], tB=np.zeros((8, 3)))
        fn = sources.particles_initialize if str(z[f"kind{t}"]) == "initialize" else sources.bc_inflow
        rc, sc = fn(g, pl, sources.SequentialDraws(z[f"u{t}"]))
        assert np.array_equal(sc, sp)
        assert np.array_equal(rc[:, 3:6], np.tile(z[f"v0{t}"][0, :3], (len(sp), 1)))


@needs_ref
@pytest.mark.parametrize("flow", [(True, False, False), (True, True, True), (False, True, True),
                                  (False, False, True)])
def test_sources_vs_ref(orc, flow):
    """Live: BCinflow and ParticlesInitialize of the reference build against
    the restatement on the same runtime stream, bit for bit."""
    from oracle import sources
    from picard_amd import host
    cfg = host.Config(nx=4, ny=5, nz=3, flow=flow, gmin=(-20.0, -25.0, -15.0),
                      gmax=(20.0, 25.0, 15.0), ppc=(0, 2, 0, 0, 0, 0, 0, 0))
    d = host.derive(cfg)
    g = oracle.Grid(4, 5, 3, 2, flow=flow, gmin=cfg.gmin, gmax=cfg.gmax)
    pl = dict(species=2, ipp=d.ipp, epp=d.epp, hpp=d.hpp, ppc=d.ppc, kelvin=d.kelvin,
              mass=d.mass, v0=d.v0, ve4=d.ve, tB4=d.tB)
    p2 = dict(pl, v0=d.v0[:, :3], ve=d.ve[:3], tB=d.tB[:, :3])
    for ref_fn, fn in ((oracle.ref_particles_initialize, sources.particles_initialize),
                       (oracle.ref_bc_inflow, sources.bc_inflow)):
        rp = np.zeros((5000, 8))
        sp = np.zeros(5000, np.int32)
        n = ref_fn(g, rp, sp, 0, 99, pl)
        draws = sources.SequentialDraws(oracle.ref_uniforms(99, 100000))
        r2, s2 = fn(g, p2, draws)
        assert np.array_equal(s2, sp[:n]) and np.array_equal(r2[:, :6], rp[:n, :6])
