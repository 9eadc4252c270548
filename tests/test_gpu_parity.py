"""GPU parity: every entry point of libpicard_hip.so against the oracle on the
same seeded inputs, bit for bit (FP64 form, reference-order deposit), plus
multi-step runs of the whole main loop."""
import itertools
import os

import numpy as np
import pytest

import oracle
from picard_amd import _lib, host
from tests.helpers import boris_tables, random_particles
from tests.pic_driver import OracleSim, oracle_grid, small_config

pytestmark = pytest.mark.gpu

FLOWS = list(itertools.product([False, True], repeat=3))


def _say(msg):
    print(f"[gpu-parity] {msg}", flush=True)


def _domain(cfg, cap=200000):
    return host.Domain(cfg, max_particles=cap)


def _eq(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: gpu={a[i]!r} "
                             f"oracle={b[i]!r}")


def test_advance_position_outflow_transfer():
    for flow in FLOWS:
        _say(f"position/outflow/transfer flow={flow}")
        cfg = small_config(6, flow=flow)
        g = oracle_grid(cfg)
        rng = np.random.default_rng(100 + sum(f << i for i, f in enumerate(flow)))
        rp, sp = random_particles(rng, g, 5000, spread=2.0)
        rp[::89, 2] = g.gmax[2] + 1.7 * (g.gmax[2] - g.gmin[2])
        dom = _domain(cfg)
        dom.append(rp, sp)
        dt = 7.9e-7
        dom.advance_position(dt)
        a = rp.copy()
        sa = sp.copy()
        oracle.advance_position(a, dt)
        got, gs = dom.particles()
        _eq(got, a, "AdvancePosition")
        n_gpu, removed = dom.bc_outflow()
        n_orc, rem_list = oracle.bc_outflow(g, a, sa)
        assert n_gpu == n_orc and removed == len(rem_list)
        got, gs = dom.particles()
        _eq(got, a[:n_orc], "BCoutflow particles")
        _eq(gs, sa[:n_orc], "BCoutflow species")
        dom.particles_transfer()
        oracle.particles_transfer(g, a, n_orc)
        got, _ = dom.particles()
        _eq(got, a[:n_orc], "ParticlesTransfer")
        dom.close()


@pytest.mark.parametrize("flow", FLOWS)
def test_gridify(flow):
    _say(f"gridify flow={flow}")
    cfg = small_config(7, flow=flow, species=3, density=(0, 5e6, 1e6, 0, 0, 0, 0, 0),
                       ppc=(0, 4, 2, 0, 0, 0, 0, 0))
    d = host.derive(cfg)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(7)
    rp, sp = random_particles(rng, g, 20000, species=3)
    rp[:7, 0] = g.gmax[0]          # boundary hits use guard nodes
    rp[7:14, 1] = g.gmin[1]
    dom = _domain(cfg)
    dom.append(rp, sp)
    dom.particles_gridify()
    N, C, skipped = oracle.particles_gridify(g, rp, sp, d.charge, d.n2p)
    assert skipped == 0
    _eq(dom.field(_lib.FIELD_N), N, "N")
    _eq(dom.field(_lib.FIELD_C), C, "C")
    dom.close()


@pytest.mark.parametrize("flow", FLOWS)
def test_distance_potentials_efield(flow):
    _say(f"distance/potentials/efield flow={flow}")
    cfg = small_config(6, flow=flow)
    cfg.ny = 5
    cfg.gmin = (-30.0, -25.0, -30.0)
    cfg.gmax = (30.0, 25.0, 30.0)
    g = oracle_grid(cfg)
    dom = _domain(cfg)
    dom.calc_distance()
    D = oracle.calc_distance(cfg.nx, cfg.ny, cfg.nz, g.dxyz, flow)
    _eq(dom.field(_lib.FIELD_D), D, "D")
    rng = np.random.default_rng(8)
    C = rng.normal(size=g.shape) * 1e-16
    dom.set_field(_lib.FIELD_C, C)
    dom.bc_potentials()
    U = oracle.bc_potentials(g, C, D)
    _eq(dom.field(_lib.FIELD_U), U, "U after BCpotentials")
    Ur = rng.normal(size=g.shape)
    dom.set_field(_lib.FIELD_U, Ur)
    dom.calc_efield()
    E = oracle.calc_efield(g, Ur)
    _eq(dom.field(_lib.FIELD_E), E, "E")
    dom.bc_fields()
    oracle.bc_fields(g, E)
    _eq(dom.field(_lib.FIELD_E), E, "E after BCfields")
    dom.close()


@pytest.mark.parametrize("flow", [(True, False, False), (True, True, False), (True, True, True),
                                  (False, True, False)])
@pytest.mark.parametrize("case", ["bcpot", "random_halo", "maxit"])
@pytest.mark.parametrize("nx", [8, 7])
def test_epotential(flow, case, nx):
    # even nx runs the two-cells-per-lane pass, odd nx the one-cell pass
    _say(f"epotential flow={flow} case={case} nx={nx}")
    cfg = small_config(8, flow=flow)
    cfg.nx = nx
    cfg.nz = 6
    cfg.gmin = (-5.0 * nx, -40.0, -30.0)
    cfg.gmax = (5.0 * nx, 40.0, 30.0)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(9)
    C = rng.normal(size=g.shape) * 1e-16
    D = oracle.calc_distance(cfg.nx, cfg.ny, cfg.nz, g.dxyz, flow)
    U0 = oracle.bc_potentials(g, C, D)
    maxerr, maxit = 1e-6, 2 ** 20
    if case == "random_halo":
        U0 = U0 + rng.normal(size=g.shape) * 1e-3
    if case == "maxit":
        maxerr, maxit = 1e-12, 7
    dom = _domain(cfg)
    dom.set_field(_lib.FIELD_C, C)
    dom.set_field(_lib.FIELD_U, U0)
    sw = ct_sweeps = None
    sw = dom.calc_epotential(maxerr, maxit)
    U = U0.copy()
    ct_sweeps, _ = oracle.calc_epotential(g, U, C, maxerr, maxit)
    assert sw == ct_sweeps, (sw, ct_sweeps)
    if case == "maxit":
        assert sw == 2 * maxit + 1
    _eq(dom.field(_lib.FIELD_U), U, "U after CalcEpotential")
    dom.close()


@pytest.mark.parametrize("flow", [(False, True, True), (False, False, False), (True, True, True)])
@pytest.mark.parametrize("nx", [130, 256, 258])
def test_epotential_long_rows(flow, nx):
    # rows longer than one wave's 128 cells: the pairs at either end of a
    # wave's segment take their x-neighbours from memory, and a segment may
    # hold a single live pair (nx = 130, 258)
    _say(f"epotential long rows flow={flow} nx={nx}")
    cfg = small_config(6, flow=flow)
    cfg.nx, cfg.ny, cfg.nz = nx, 5, 4
    cfg.gmin = (-5.0 * nx, -25.0, -20.0)
    cfg.gmax = (5.0 * nx, 25.0, 20.0)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(nx)
    C = rng.normal(size=g.shape) * 1e-16
    D = oracle.calc_distance(cfg.nx, cfg.ny, cfg.nz, g.dxyz, flow)
    U0 = oracle.bc_potentials(g, C, D) + rng.normal(size=g.shape) * 1e-3
    for maxerr, maxit in ((1e-12, 9), (1e-4, 2 ** 20)):
        dom = _domain(cfg)
        dom.set_field(_lib.FIELD_C, C)
        dom.set_field(_lib.FIELD_U, U0)
        sw = dom.calc_epotential(maxerr, maxit)
        U = U0.copy()
        ct_sweeps, _ = oracle.calc_epotential(g, U, C, maxerr, maxit)
        assert sw == ct_sweeps, (sw, ct_sweeps)
        _eq(dom.field(_lib.FIELD_U), U, f"U after CalcEpotential (maxit {maxit})")
        dom.close()


@pytest.mark.parametrize("mode", ["boris", "fluid", "nogyration"])
def test_advance_velocity(mode):
    _say(f"advance velocity mode={mode}")
    fluid = [0] * 8
    gyration = [1] * 8
    if mode == "fluid":
        fluid = [1, 0, 1, 0, 0, 0, 0, 0]
    if mode == "nogyration":
        gyration = [0, 1, 0, 1, 1, 1, 1, 1]
        fluid = [0, 0, 1, 0, 0, 0, 0, 0]
    cfg = small_config(6, flow=(True, False, True), species=3,
                       density=(0, 5e6, 1e6, 0, 0, 0, 0, 0), ppc=(0, 4, 2, 0, 0, 0, 0, 0),
                       fluid=tuple(bool(f) for f in fluid), gyration=tuple(bool(f) for f in gyration),
                       B0=(1.1e-9, -2e-9, 5e-9))
    d = host.derive(cfg)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(10)
    rp, sp = random_particles(rng, g, 20000, species=3)
    E = rng.normal(0.0, 1e-3, size=g.shape + (3,))
    dom = _domain(cfg)
    dom.append(rp, sp)
    dom.set_field(_lib.FIELD_E, E)
    dom.advance_velocity()
    a = rp.copy()
    oracle.advance_velocity(g, a, sp, E, d.B0, d.dt, d.ve, d.eta, d.tB, d.sB, fluid, gyration)
    got, _ = dom.particles()
    _eq(got, a, "AdvanceVelocity")
    dom.close()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("kind", ["boris", "fluid"])
def test_advance_velocity_imp_exp(mode, kind):
    """AdvanceVelocityImp (1) / AdvanceVelocityExp (2) at deltat = dt/2,
    bit-exact against the oracle (FP64 form)."""
    _say(f"advance velocity imp/exp mode={mode} {kind}")
    fluid = [1, 0, 1, 0, 0, 0, 0, 0] if kind == "fluid" else [0] * 8
    gyration = [1] * 8
    cfg = small_config(6, flow=(True, False, True), species=3,
                       density=(0, 5e6, 1e6, 0, 0, 0, 0, 0), ppc=(0, 4, 2, 0, 0, 0, 0, 0),
                       fluid=tuple(bool(f) for f in fluid), gyration=tuple(bool(f) for f in gyration),
                       B0=(1.1e-9, -2e-9, 5e-9))
    d = host.derive(cfg)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(11 + mode)
    rp, sp = random_particles(rng, g, 20000, species=3)
    E = rng.normal(0.0, 1e-3, size=g.shape + (3,))
    dom = _domain(cfg)
    dom.append(rp, sp)
    dom.set_field(_lib.FIELD_E, E)
    if mode == 1:
        dom.advance_velocity_imp()
    else:
        dom.advance_velocity_exp()
    a = rp.copy()
    oracle.advance_velocity_ie(mode, g, a, sp, E, d.B0, 0.5 * d.dt, d.dt, d.ve, d.eta, d.tB,
                               fluid, gyration)
    got, _ = dom.particles()
    _eq(got, a, f"AdvanceVelocity{'Imp' if mode == 1 else 'Exp'}")
    dom.close()


@pytest.mark.parametrize("flow", [(True, False, False), (False, True, True), (True, True, True)])
def test_gridify_all(flow):
    """ParticlesGridifyAll: N, C and fluxes F bit-exact against the oracle
    (FP64 form); the split form's fixed-point fluxes within 1e-6."""
    _say(f"gridify_all flow={flow}")
    cfg = small_config(7, flow=flow, species=3, density=(0, 5e6, 1e6, 0, 0, 0, 0, 0),
                       ppc=(0, 4, 2, 0, 0, 0, 0, 0))
    d = host.derive(cfg)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(40)
    rp, sp = random_particles(rng, g, 30000, species=3)
    F0, N0, C0, sk = oracle.particles_gridify_all(g, rp, sp, d.charge, d.n2p)
    assert sk == 0
    dom = _domain(cfg)
    dom.append(rp, sp)
    dom.particles_gridify_all()
    _eq(dom.field(_lib.FIELD_N), N0, "GridifyAll N")
    _eq(dom.field(_lib.FIELD_C), C0, "GridifyAll C")
    _eq(dom.field(_lib.FIELD_F), F0, "GridifyAll F")
    dom.close()
    sdom = host.Domain(cfg, max_particles=100000, d=d, precision=_lib.PICARD_FP32_SPLIT)
    sdom.append(rp, sp)
    sdom.particles_gridify_all()
    Fs = sdom.field(_lib.FIELD_F)
    err = np.linalg.norm(Fs - F0) / np.linalg.norm(F0)
    _say(f"split flux rel err {err:.2e}")
    assert err <= 1e-6
    sdom.close()


def test_output_step_matches_oracle(tmp_path):
    """An output step (picard.f90:600-616): Imp push, particle dump,
    ParticlesGridifyAll, field dump, Exp push; the FP64 form's particles and
    every dumped file equal the oracle's, after one ordinary step."""
    from picard_amd import output
    _say("output step")
    cfg = small_config(8, ppc=4)
    d = host.derive(cfg)
    rng = np.random.default_rng(71)
    rp, sp = host.particles_initialize(cfg, d, rng)
    inflow = [host.bc_inflow(cfg, d, rng) for _ in range(2)]
    sim = OracleSim(cfg, d, rp, sp)
    dom = _domain(cfg)
    dom.append(rp, sp)
    dom.calc_distance()
    sim.append(*inflow[0])
    sim.step()
    dom.append(*inflow[0])
    dom.step()
    # oracle output step
    sim.append(*inflow[1])
    g = sim.g
    oracle.advance_position(sim.rp, d.dt, sim.n)
    n_new, _ = oracle.bc_outflow(g, sim.rp, sim.sp, sim.n)
    sim.n = n_new
    sim.rp, sim.sp = np.ascontiguousarray(sim.rp[:n_new]), np.ascontiguousarray(sim.sp[:n_new])
    oracle.particles_transfer(g, sim.rp, sim.n)
    N, C, _ = oracle.particles_gridify(g, sim.rp, sim.sp, d.charge, d.n2p)
    U = oracle.bc_potentials(g, C, sim.D)
    oracle.calc_epotential(g, U, C, cfg.maxerr, cfg.maxit)
    E = oracle.calc_efield(g, U)
    oracle.bc_fields(g, E)
    fl, gy = [int(f) for f in cfg.fluid], [int(f) for f in cfg.gyration]
    oracle.advance_velocity_ie(1, g, sim.rp, sim.sp, E, d.B0, 0.5 * d.dt, d.dt, d.ve, d.eta, d.tB,
                               fl, gy)
    ref_dir = tmp_path / "oracle"
    output.write_particles(str(ref_dir), sim.rp, sim.sp, 0)
    F, N2, C2, _ = oracle.particles_gridify_all(g, sim.rp, sim.sp, d.charge, d.n2p)
    output.write_fields(str(ref_dir), U, N2, F, E, 0)
    oracle.advance_velocity_ie(2, g, sim.rp, sim.sp, E, d.B0, 0.5 * d.dt, d.dt, d.ve, d.eta, d.tB,
                               fl, gy)
    # device output step
    dom.append(*inflow[1])
    names, _ = dom.output_step(str(tmp_path / "gpu"), True, True)
    got, gsp = dom.particles()
    _eq(got, sim.rp, "particles after the output step")
    assert np.array_equal(gsp, sim.sp)
    assert len(names) == 1 + 4 + 4 * cfg.species        # particles, UE, E, N, X, Y, Z
    for name in names:
        rel = os.path.relpath(name, str(tmp_path / "gpu"))
        with open(name, "rb") as a, open(ref_dir / rel, "rb") as b:
            assert a.read() == b.read(), rel
    dom.close()


def _run_phases(cfg, steps, label):
    """Each subroutine of the main loop in turn on both sides, compared after
    every phase, so a divergence is reported at the phase that causes it."""
    d = host.derive(cfg)
    g = oracle_grid(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    orc = OracleSim(cfg, d, rp, sp)
    dom = host.Domain(cfg, d=d)
    dom.append(rp, sp)
    dom.calc_distance()
    _eq(dom.field(_lib.FIELD_D), orc.D, f"{label} D")
    fl = [int(f) for f in cfg.fluid]
    gy = [int(f) for f in cfg.gyration]
    for it in range(1, steps + 1):
        inflow = host.bc_inflow(cfg, d, rng)
        orc.append(*inflow)
        dom.append(*inflow)
        oracle.advance_position(orc.rp, d.dt, orc.n)
        dom.advance_position()
        _eq(dom.particles()[0], orc.rp, f"{label} step {it} AdvancePosition")
        n_new, rem = oracle.bc_outflow(g, orc.rp, orc.sp, orc.n)
        orc.n = n_new
        orc.rp = np.ascontiguousarray(orc.rp[:n_new])
        orc.sp = np.ascontiguousarray(orc.sp[:n_new])
        ng, rg = dom.bc_outflow()
        assert (ng, rg) == (n_new, len(rem)), (label, it, ng, rg, n_new, len(rem))
        got, gs = dom.particles()
        _eq(got, orc.rp, f"{label} step {it} BCoutflow")
        _eq(gs, orc.sp, f"{label} step {it} BCoutflow species")
        oracle.particles_transfer(g, orc.rp, orc.n)
        dom.particles_transfer()
        _eq(dom.particles()[0], orc.rp, f"{label} step {it} ParticlesTransfer")
        orc.N, orc.C, _ = oracle.particles_gridify(g, orc.rp, orc.sp, d.charge, d.n2p, orc.n)
        dom.particles_gridify()
        _eq(dom.field(_lib.FIELD_N), orc.N, f"{label} step {it} N")
        _eq(dom.field(_lib.FIELD_C), orc.C, f"{label} step {it} C")
        orc.U = oracle.bc_potentials(g, orc.C, orc.D)
        dom.bc_potentials()
        _eq(dom.field(_lib.FIELD_U), orc.U, f"{label} step {it} BCpotentials")
        so, _ = oracle.calc_epotential(g, orc.U, orc.C, cfg.maxerr, cfg.maxit)
        sg = dom.calc_epotential()
        _say(f"{label} step {it}: n={ng} removed={rg} sweeps gpu={sg} oracle={so}")
        assert sg == so, (label, it, sg, so)
        _eq(dom.field(_lib.FIELD_U), orc.U, f"{label} step {it} CalcEpotential")
        orc.E = oracle.calc_efield(g, orc.U)
        oracle.bc_fields(g, orc.E)
        dom.calc_efield()
        dom.bc_fields()
        _eq(dom.field(_lib.FIELD_E), orc.E, f"{label} step {it} E")
        oracle.advance_velocity(g, orc.rp, orc.sp, orc.E, d.B0, d.dt, d.ve, d.eta, d.tB, d.sB,
                                fl, gy, orc.n)
        dom.advance_velocity()
        _eq(dom.particles()[0], orc.rp, f"{label} step {it} AdvanceVelocity")
    dom.close()


def _run_both(cfg, steps, label, check_every=1):
    """The fused picard_step entry point against the oracle loop."""
    d = host.derive(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    orc = OracleSim(cfg, d, rp, sp)
    dom = host.Domain(cfg, d=d)
    dom.append(rp, sp)
    dom.calc_distance()
    for it in range(1, steps + 1):
        inflow = host.bc_inflow(cfg, d, rng)                  # picard.f90:560
        orc.append(*inflow)
        dom.append(*inflow)
        so = orc.step()
        st = dom.step()
        _say(f"{label} step {it}: n={st.num_local} removed={st.removed} sweeps={st.sweeps}")
        assert (st.num_local, st.removed, st.sweeps) == (so["num_local"], so["removed"],
                                                         so["sweeps"])
        if it % check_every == 0 or it == steps:
            got, gs = dom.particles()
            _eq(gs, orc.sp, f"{label} step {it} species")
            _eq(got, orc.rp, f"{label} step {it} particles")
            _eq(dom.field(_lib.FIELD_N), orc.N, f"{label} step {it} N")
            _eq(dom.field(_lib.FIELD_C), orc.C, f"{label} step {it} C")
            _eq(dom.field(_lib.FIELD_U), orc.U, f"{label} step {it} U")
            _eq(dom.field(_lib.FIELD_E), orc.E, f"{label} step {it} E")
    dom.close()


def test_phases_small_xflow():
    _run_phases(small_config(10, ppc=4, flow=(True, False, False)), 4, "10^3 xflow phases")


def test_phases_small_xyzflow():
    _run_phases(small_config(8, ppc=3, flow=(True, True, True), steps_seed=78), 3,
                "8^3 xyzflow phases")


def test_multistep_small_xflow():
    _run_both(small_config(10, ppc=4, flow=(True, False, False)), 12, "10^3 xflow")


def test_multistep_small_xyflow():
    cfg = small_config(8, ppc=3, flow=(True, True, False), steps_seed=77)
    _run_both(cfg, 8, "8^3 xyflow")


def test_multistep_config1_prefix():
    """The 32^3 workload (H+ and e- at 8 per cell, 400 km/s drift) for 2 steps."""
    cfg = host.config1()
    _run_both(cfg, 2, "config1", check_every=1)
