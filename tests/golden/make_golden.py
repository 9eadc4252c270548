"""Writes golden input/output vectors for the reference's MPI-free hot-path
subroutines (AdvancePosition, BCoutflow, CalcDistance, AdvanceVelocity),
computed by the reference's own Fortran compiled into oracle/_ref.

Run in the container where /root/reference is mounted:
    make -C oracle && python tests/golden/make_golden.py [parameters|sources]
The .npz files are data only (inputs and the reference's outputs)."""
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests.helpers import boris_tables, random_particles  # noqa: E402


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built; run `make -C oracle` where the reference is")
    rng = np.random.default_rng(20261202)

    # AdvancePosition
    g = oracle.Grid(6, 6, 6, 2)
    rp, _ = random_particles(rng, g, 600)
    out = rp.copy()
    oracle.ref_advance_position(out, 7.927e-7, 500)
    np.savez_compressed(os.path.join(HERE, "ref_advance_position.npz"), rp_in=rp, rp_out=out,
                        dt=np.float64(7.927e-7), n=np.int64(500))

    # BCoutflow, every flow-flag combination
    g = oracle.Grid(5, 6, 4, 2)
    rp, sp = random_particles(rng, g, 700, spread=2.0)
    rp[::53, 2] = g.gmin[2] - 1.6 * (g.gmax[2] - g.gmin[2])
    d = dict(rp_in=rp, sp_in=sp, dims=np.array([5, 6, 4]))
    flows = list(itertools.product([0, 1], repeat=3))
    for t, flow in enumerate(flows):
        gg = oracle.Grid(5, 6, 4, 2, flow=tuple(bool(f) for f in flow))
        a, sa = rp.copy(), sp.copy()
        n = oracle.ref_bc_outflow(gg, a, sa, a.shape[0])
        d[f"flow{t}"] = np.array(flow, np.int32)
        d[f"n{t}"] = np.int64(n)
        d[f"rp_out{t}"] = a[:n]
        d[f"sp_out{t}"] = sa[:n]
    d["ncase"] = np.int64(len(flows))
    np.savez_compressed(os.path.join(HERE, "ref_bc_outflow.npz"), **d)

    # CalcDistance
    d = {}
    cases = [((4, 5, 3), (10.0, 10.0, 10.0)), ((6, 4, 5), (3.0, 7.5, 11.25))]
    t = 0
    for dims, dx in cases:
        for flow in flows:
            d[f"dims{t}"] = np.array(dims, np.int32)
            d[f"dxyz{t}"] = np.array(dx, np.float64)
            d[f"flow{t}"] = np.array(flow, np.int32)
            d[f"D{t}"] = oracle.ref_calc_distance(*dims, dx, flow)
            t += 1
    d["ncase"] = np.int64(t)
    np.savez_compressed(os.path.join(HERE, "ref_calc_distance.npz"), **d)

    # AdvanceVelocity (Boris branch, three species)
    flow = (1, 0, 0)
    g = oracle.Grid(5, 4, 6, 3, flow=tuple(bool(f) for f in flow))
    rp, sp = random_particles(rng, g, 500, species=3)
    E = rng.normal(0.0, 1e-3, size=g.shape + (3,))
    dt = 7.927e-7
    B0 = np.array([1.1e-9, -2.0e-9, 5e-9])
    B4 = np.array([B0[0], B0[1], B0[2], np.sqrt(np.sum(B0 ** 2))])
    mass = np.array([9.1093819e-31, 1.6726216e-27, 18 * 1.6726216e-27] + [1.6726216e-27] * 5)
    charge = np.array([-1.60217646e-19] + [1.60217646e-19] * 7)
    eta = charge / mass
    fluid = np.zeros(8, np.int32)
    gyration = np.ones(8, np.int32)
    tB, sB = boris_tables(dt, B0, eta, gyration)
    ve = np.array([4.0e5, 0.0, 0.0, 4.0e5])
    out = rp.copy()
    oracle.ref_advance_velocity(g, out, sp, E, B4, dt, np.zeros((8, 4)), ve, eta, tB, sB, 500,
                                fluid, gyration)
    np.savez_compressed(os.path.join(HERE, "ref_advance_velocity.npz"), rp_in=rp, rp_out=out,
                        sp=sp, E=E, B0=B4, dt=np.float64(dt), ve=ve, eta=eta, tB=tB, sB=sB,
                        fluid=fluid, gyration=gyration, n=np.int64(500),
                        dims=np.array([5, 4, 6]), flow=np.array(flow, np.int32))
    # AdvanceVelocityImp / AdvanceVelocityExp with deltat = dt/2
    d = dict(rp_in=rp, sp=sp, E=E, B0=B4, dt=np.float64(dt), ve=ve, eta=eta, tB=tB,
             fluid=fluid, gyration=gyration, n=np.int64(500), dims=np.array([5, 4, 6]),
             flow=np.array(flow, np.int32))
    for mode in (1, 2):
        out = rp.copy()
        oracle.ref_advance_velocity_ie(mode, g, out, sp, E, B4, 0.5 * dt, dt, np.zeros((8, 4)),
                                       ve, eta, tB, 500, fluid, gyration)
        d[f"rp_out{mode}"] = out
    np.savez_compressed(os.path.join(HERE, "ref_advance_velocity_ie.npz"), **d)
    produce_cases()
    dump_cases()
    source_cases()
    print("golden vectors written to", HERE)


def _spread(rng, shape):
    """Values over many decades and both signs, with exact zeros and ties."""
    a = rng.normal(size=shape) * 10.0 ** rng.integers(-120, 120, size=shape)
    flat = a.reshape(-1)
    flat[::17] = 0.0
    flat[5::23] = -0.0
    flat[3::29] = 0.015625
    flat[7::31] = -9.99995e10
    return a


def dump_cases():
    """DumpParticles / DumpFields of the reference (Write.f90) on spread
    values: the inputs and the files' bytes."""
    import tempfile
    rng = np.random.default_rng(99)
    n = 40
    rp = _spread(rng, (n, 8))
    sp = rng.integers(1, 4, n).astype(np.int32)
    nx, ny, nz, S = 4, 3, 2, 3
    shp = (nz + 2, ny + 2, nx + 2)
    U = _spread(rng, shp)
    E = _spread(rng, shp + (3,))
    N = _spread(rng, shp + (S,))
    F = _spread(rng, shp + (3, S))
    d = dict(rp=rp, sp=sp, U=U, E=E, N=N, F=F, dims=np.array([nx, ny, nz, S]))
    with tempfile.TemporaryDirectory() as t:
        text, it = oracle.ref_dump_particles(rp, sp, n, 7, 0, t)
        d["particles_txt"] = np.frombuffer(text, np.uint8)
        files = oracle.ref_dump_fields(U, N, F, E, nx, ny, nz, 2, S, 0, t)
    d["field_names"] = np.array(sorted(files))
    for name, b in files.items():
        d["txt_" + name] = np.frombuffer(b, np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_dump.npz"), **d)
    print(f"dump: {len(files)} field files")


def produce_cases():
    """ParticlesProduce: two Haser cases, the reference's output and the
    runtime uniform stream it consumed."""
    from oracle import haser
    from picard_amd import host
    d_out = {}
    cases = [
        # heavy ion (18 amu, +e: one electron per ion), origin at a cell corner
        dict(dims=(8, 8, 8), gmin=(-40.0,) * 3, gmax=(40.0,) * 3, species=3,
             charge3=(1.602e-19,) * 6, target=400.0),
        # origin at a cell centre (the 8-point distance branch); a +2e ion
        # (two electrons) and a -e ion (one proton)
        dict(dims=(8, 6, 7), gmin=(-45.0, -35.0, -35.0), gmax=(35.0, 25.0, 35.0), species=4,
             charge3=(2 * 1.60217646e-19, -1.60217646e-19) + (1.60217646e-19,) * 4, target=300.0),
    ]
    for t, cs in enumerate(cases):
        cfg = host.Config(nx=cs["dims"][0], ny=cs["dims"][1], nz=cs["dims"][2],
                          species=cs["species"], gmin=cs["gmin"], gmax=cs["gmax"],
                          charge3=cs["charge3"], ppc=(0, 4, 0, 0, 0, 0, 0, 0))
        d = host.derive(cfg)
        vn = np.zeros(8)
        nu = np.zeros(8)
        Ek = np.zeros(8)
        Qn = np.zeros(8)
        for s in range(2, cs["species"]):
            vn[s] = 629.0 + 100.0 * s
            nu[s] = 1.26e-7 * (s + 1)
            Ek[s] = 15.2 * 1.60217646e-19
            Qn[s] = 1.0
        g = oracle.Grid(*cs["dims"], cs["species"], gmin=cs["gmin"], gmax=cs["gmax"])
        kw = dict(mass=d.mass, Ek=Ek, Qn=Qn, vn=vn, nu_d=nu, nu_i=nu, n2p=d.n2p, epp=d.epp,
                  hpp=d.hpp, dt=d.dt)
        # scale Qn so that about `target` ions are produced
        lam = 0.0
        for kk in range(cfg.nz):
            for jj in range(cfg.ny):
                for ii in range(cfg.nx):
                    x0 = cfg.gmin[0] + d.dxyz[0] * ii
                    y0 = cfg.gmin[1] + d.dxyz[1] * jj
                    z0 = cfg.gmin[2] + d.dxyz[2] * kk
                    r1, r2 = haser.cell_distance(x0, y0, z0, *d.dxyz)
                    for s in range(cs["species"]):
                        if d.n2p[s] > 0:
                            lam += haser.expected_count(r1, r2, s, Qn, vn, nu, nu, d.n2p, d.dV)
        Qn = Qn * (cs["target"] / lam)
        kw["Qn"] = Qn
        seed = 4711 + t
        cap = 20000
        rp = np.zeros((cap, 8))
        sp = np.zeros(cap, np.int32)
        n0 = 5
        rp[:n0] = 1.0
        sp[:n0] = 2
        n = oracle.ref_particles_produce(g, rp, sp, n0, seed, kelvin=d.kelvin, v0=d.v0, ve=d.ve,
                                         ppc=d.ppc, ipp=d.ipp, **kw)
        u = oracle.ref_uniforms(seed, 200000)
        draws = haser.SequentialDraws(u)
        haser.produce(cfg.nx, cfg.ny, cfg.nz, cfg.gmin, tuple(d.dxyz), cs["species"],
                      draws=draws, **kw)
        d_out.update({
            f"dims{t}": np.array(cs["dims"], np.int32), f"gmin{t}": np.array(cs["gmin"]),
            f"gmax{t}": np.array(cs["gmax"]), f"species{t}": np.int32(cs["species"]),
            f"mass{t}": d.mass, f"Ek{t}": Ek, f"Qn{t}": Qn, f"vn{t}": vn, f"nu{t}": nu,
            f"n2p{t}": d.n2p, f"epp{t}": d.epp, f"hpp{t}": d.hpp, f"dt{t}": np.float64(d.dt),
            f"u{t}": u[:draws.pos], f"n0{t}": np.int64(n0), f"rp_out{t}": rp[n0:n],
            f"sp_out{t}": sp[n0:n]})
        print(f"produce case {t}: {n - n0} particles, {draws.pos} uniforms")
    d_out["ncase"] = np.int64(len(cases))
    np.savez_compressed(os.path.join(HERE, "ref_particles_produce.npz"), **d_out)


def source_plasma(cfg, d):
    """The plasma tables the sources read (picard.f90:371-466 derived)."""
    return dict(species=cfg.species, ipp=d.ipp, epp=d.epp, hpp=d.hpp, ppc=d.ppc,
                kelvin=d.kelvin, mass=d.mass, v0=d.v0, ve4=d.ve, tB4=d.tB)


def source_cases():
    """ParticlesInitialize and BCinflow of the reference on small grids (two
    plasmas: H+ and e-; and a 4-species plasma with a +2e and a -e ion) at
    several flow-flag combinations: the inputs, the runtime uniform stream
    each call consumed, and the reference's particles."""
    from oracle import sources
    from picard_amd import host
    cases = [dict(dims=(5, 4, 3), species=2, flows=[(True, False, False), (True, True, True),
                                                    (False, True, True), (False, False, True)]),
             dict(dims=(4, 3, 3), species=4, flows=[(True, True, False), (True, False, True)])]
    out = {}
    t = 0
    for k, cs in enumerate(cases):
        nx, ny, nz = cs["dims"]
        kw = dict(nx=nx, ny=ny, nz=nz, species=cs["species"],
                  gmin=(-5.0 * nx, -6.0 * ny, -4.0 * nz), gmax=(5.0 * nx, 6.0 * ny, 4.0 * nz))
        if cs["species"] == 4:
            kw.update(ppc=(0, 4, 2, 1, 0, 0, 0, 0), density=(0.0, 5.0e6, 1.0e6, 5.0e5, 0, 0, 0, 0),
                      charge3=(2 * host.EC, -host.EC) + (host.EC,) * 4,
                      kelvin=(116045.0, 58022.0, 2.0e5, 3.0e4, 0, 0, 0, 0))
        else:
            kw.update(ppc=(0, 3, 0, 0, 0, 0, 0, 0))
        for q, flow in enumerate([None] + cs["flows"]):
            cfg = host.Config(flow=flow or (False, False, False), **kw)
            d = host.derive(cfg)
            g = oracle.Grid(nx, ny, nz, cfg.species, flow=cfg.flow, gmin=cfg.gmin,
                            gmax=cfg.gmax)
            pl = source_plasma(cfg, d)
            cap = 20000
            rp = np.zeros((cap, 8))
            sp = np.zeros(cap, np.int32)
            seed = 4321 + 17 * t
            fn = oracle.ref_particles_initialize if flow is None else oracle.ref_bc_inflow
            n = fn(g, rp, sp, 0, seed, pl)
            u = oracle.ref_uniforms(seed, 400000)
            draws = sources.SequentialDraws(u)
            p2 = dict(pl, v0=d.v0[:, :3], ve=d.ve[:3], tB=d.tB[:, :3])
            r2, s2 = (sources.particles_initialize if flow is None else sources.bc_inflow)(
                g, p2, draws)
            assert np.array_equal(s2, sp[:n]) and np.array_equal(r2[:, :6], rp[:n, :6])
            out[f"kind{t}"] = np.array("initialize" if flow is None else "inflow")
            out[f"dims{t}"] = np.array([nx, ny, nz, cfg.species], np.int32)
            out[f"gmin{t}"] = np.array(cfg.gmin)
            out[f"gmax{t}"] = np.array(cfg.gmax)
            out[f"flow{t}"] = np.array(cfg.flow, np.int32)
            for key in ("ipp", "epp", "hpp", "ppc"):
                out[f"{key}{t}"] = np.asarray(pl[key], np.int32)
            for key in ("kelvin", "mass", "v0", "ve4", "tB4"):
                out[f"{key}{t}"] = np.asarray(pl[key], np.float64)
            out[f"u{t}"] = u[: draws.pos]
            out[f"rp_out{t}"] = rp[:n, :6].copy()
            out[f"sp_out{t}"] = sp[:n].copy()
            print(f"source case {t} ({out[f'kind{t}']}, flow {cfg.flow}): {n} particles, "
                  f"{draws.pos} uniforms")
            t += 1
    out["ncase"] = np.int64(t)
    np.savez_compressed(os.path.join(HERE, "ref_sources.npz"), **out)


def parameters_cases():
    """WriteWarnings' save/parameters.dat (WriteWarnings.f90:130-256) for
    three setups: the 32^3 case with picard.f90's own Haser parameters
    (picard.f90:255-300), the 256^3 workload, and one with mixed gyration
    flags, an I4 field that overflows, three-digit exponents, negative values
    and a 2 x 1 x 3 process grid.  Stores each setup's inputs (JSON) and the
    file's bytes."""
    import json
    import tempfile
    from picard_amd import host
    ec = host.EC
    ref_haser = dict(Qn=[0.0, 0.0, 400.0 ** (-3) * 1.99e26, 0, 0, 0, 0, 0],
                     vn=[0.0, 0.0, 629.0, 0, 0, 0, 0, 0],
                     Ek=[0.0, 0.0, 1.0 * 1.52e1 * ec, 0, 0, 0, 0, 0],
                     nu_i=[0.0, 0.0, 1.26e-7, 0, 0, 0, 0, 0],
                     nu_d=[0.0, 0.0, 1.26e-7, 0, 0, 0, 0, 0])
    cases = [
        (dict(), ref_haser, (1, 1, 1)),
        (dict(n=256), None, (1, 1, 1)),
        (dict(nx=12345, ny=7, nz=3, species=4, flow=(False, True, True),
              gmin=(-1.0e-3, -5.0, 2.0), gmax=(5.0e-3, 5.0, 9.0e5),
              ppc=(0, 3000, 7, 11, 0, 0, 0, 0), v0=(-4.0e5, 2.5e3, -1.0e-200),
              B0=(-1.0e-9, 3.0e-301, 7.0e-9),
              charge3=(-2.0 * ec, 3.0 * ec, ec, ec, ec, ec),
              gyration=(True, False, True, False, False, True, True, False)),
         dict(Qn=[1e-300, 2.5e28, 0, 0, 0, 0, 0, 0], vn=[0, -1.5, 629.0, 0, 0, 0, 0, 0],
              Ek=[0] * 8, nu_i=[0, 0, 0, 1e+300, 0, 0, 0, 0], nu_d=[0] * 8), (2, 1, 3)),
    ]
    out = {}
    for t, (kw, haser, procs) in enumerate(cases):
        if kw.get("n"):
            cfg = host.config2(kw["n"])
        else:
            cfg = host.Config(**kw)
        if haser:
            for k, v in haser.items():
                setattr(cfg, k, tuple(float(x) for x in v))
        d = host.derive(cfg)
        with tempfile.TemporaryDirectory() as tmp:
            _, par, _ = oracle.ref_write_warnings(cfg, d, tmp, procs)
        desc = {k: (list(v) if isinstance(v, tuple) else v) for k, v in vars(cfg).items()}
        out[f"config{t}"] = np.array(json.dumps(desc))
        out[f"procs{t}"] = np.array(procs, np.int32)
        out[f"parameters{t}"] = np.frombuffer(par, np.uint8)
        print(f"parameters case {t}: {len(par)} bytes")
    out["ncase"] = np.int64(len(cases))
    np.savez_compressed(os.path.join(HERE, "ref_parameters.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1:] in (["parameters"], ["sources"]):
        if not oracle.ref_available():
            raise SystemExit("oracle/_ref is not built; run `make -C oracle` where the reference is")
        parameters_cases() if sys.argv[1] == "parameters" else source_cases()
    else:
        main()
        parameters_cases()
