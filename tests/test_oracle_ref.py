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
