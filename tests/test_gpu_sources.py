"""Device-side sources (ParticlesInitialize / BCinflow on the GPU): counts and
slots follow the reference's loop order, ions and their electrons are
co-located, positions fill their cells, velocities are the drifting
Maxwellian, and a seed reproduces the same particles."""
import numpy as np
import pytest

from picard_amd import _lib, host

pytestmark = pytest.mark.gpu


def _cfg():
    return host.Config(nx=12, ny=10, nz=8, flow=(True, True, False), gmin=(-60.0, -50.0, -40.0),
                       gmax=(60.0, 50.0, 40.0), ppc=(0, 6, 0, 0, 0, 0, 0, 0))


def test_generate_initial_statistics_and_determinism():
    print("[gpu-sources] initial", flush=True)
    cfg = _cfg()
    d = host.derive(cfg)
    a = host.Domain(cfg, max_particles=400000, d=d)
    a.generate_initial(1234)
    rp, sp = a.particles()
    assert rp.shape[0] == cfg.nx * cfg.ny * cfg.nz * 12
    assert np.array_equal(rp[0::2, :3], rp[1::2, :3])
    assert np.all(sp[0::2] == 2) and np.all(sp[1::2] == 1)
    # cell-major order: the first 12 particles sit in cell (2,2,2)
    assert np.all(rp[:12, 0] >= cfg.gmin[0]) and np.all(rp[:12, 0] < cfg.gmin[0] + 10.0)
    for c in range(3):
        assert rp[:, c].min() >= cfg.gmin[c] and rp[:, c].max() <= cfg.gmax[c]
    vi = rp[sp == 2, 3:6]
    ve = rp[sp == 1, 3:6]
    assert abs(vi[:, 0].mean() - 4e5) < 500.0
    sig_i = np.sqrt(host.KB * 116045.0 / host.MP)
    sig_e = np.sqrt(host.KB * 116045.0 / host.ME)
    assert abs(vi[:, 2].std() / sig_i - 1) < 0.02
    assert abs(ve[:, 2].std() / sig_e - 1) < 0.02
    b = host.Domain(cfg, max_particles=400000, d=d)
    b.generate_initial(1234)
    rb, sb = b.particles()
    assert np.array_equal(rb, rp) and np.array_equal(sb, sp)
    c = host.Domain(cfg, max_particles=400000, d=d)
    c.generate_initial(1235)
    assert not np.array_equal(c.particles()[0], rp)
    for dom in (a, b, c):
        dom.close()


def test_generate_inflow_faces():
    print("[gpu-sources] inflow", flush=True)
    cfg = _cfg()
    d = host.derive(cfg)
    dom = host.Domain(cfg, max_particles=400000, d=d)
    dom.generate_inflow(99, 1)
    rp, sp = dom.particles()
    ref_r, ref_s = host.bc_inflow(cfg, d, np.random.default_rng(0))
    assert rp.shape == ref_r.shape                     # same cells, same counts
    assert np.array_equal(sp, ref_s)                   # same species order
    dx = d.dxyz
    # each particle lies in the guard cell the host source would use for its slot
    lo_ref = np.floor((ref_r[:, :3] - np.array(cfg.gmin)) / dx)
    lo_gpu = np.floor((rp[:, :3] - np.array(cfg.gmin)) / dx)
    assert np.array_equal(lo_ref, lo_gpu)
    dom.generate_inflow(99, 2)
    r2, _ = dom.particles(rp.shape[0])
    assert not np.array_equal(r2[:, :3], rp[:, :3])    # a new draw each step
    dom.close()


HASER_KEY_XOR = 0x5A5E12D4C0FFEE11


def _haser_cfg(species=3, gmin=(-40.0,) * 3, gmax=(40.0,) * 3, dims=(8, 8, 8), scale=1.0):
    cfg = host.Config(nx=dims[0], ny=dims[1], nz=dims[2], species=species, gmin=gmin, gmax=gmax,
                      ppc=(0, 4, 0, 0, 0, 0, 0, 0),
                      charge3=(2 * host.EC, -host.EC) + (host.EC,) * 4 if species == 4
                      else (host.EC,) * 6)
    vn = [0.0] * 8
    nu = [0.0] * 8
    Ek = [0.0] * 8
    Qn = [0.0] * 8
    for s in range(2, species):
        vn[s] = 629.0 + 100.0 * s
        nu[s] = 1.26e-7 * (s + 1)
        Ek[s] = 15.2 * host.EC
        Qn[s] = 4.0e19 * scale
    cfg.vn, cfg.nu_i, cfg.nu_d, cfg.Ek, cfg.Qn = tuple(vn), tuple(nu), tuple(nu), tuple(Ek), tuple(Qn)
    cfg.rate_scale = 1.0
    return cfg


def _haser_expect(cfg, d, seed, step):
    """The oracle's ParticlesProduce restatement driven by the device's
    Philox streams (count: (cs, 0, 2*step); ion: (cs, pp, 2*step+1))."""
    from oracle import haser
    from tests.philox import uniforms
    key = (seed ^ HASER_KEY_XOR) & 0xFFFFFFFFFFFFFFFF

    def count_fn(cs, ss):
        return float(uniforms(key, cs & 0xFFFFFFFF, cs >> 32, 2 * step, 1)[0])

    def ion_fn(cs, ss, pp):
        return uniforms(key, cs, pp, 2 * step + 1, 7)

    draws = haser.CellStreamDraws(count_fn, ion_fn, None)
    return haser.produce(cfg.nx, cfg.ny, cfg.nz, cfg.gmin, tuple(d.dxyz), cfg.species,
                         mass=d.mass, Ek=cfg.Ek, Qn=cfg.Qn, vn=cfg.vn, nu_d=cfg.nu_d,
                         nu_i=cfg.nu_i, n2p=d.n2p, epp=d.epp, hpp=d.hpp, dt=d.dt, draws=draws,
                         rate_scale=cfg.rate_scale, producing_only=True)


@pytest.mark.parametrize("case", ["corner", "centre"])
def test_haser_produce_matches_oracle(case):
    """Device ParticlesProduce against the oracle restatement on the same
    Philox draws: same counts, order and species; positions and velocities
    equal to rounding (device exp/sqrt are not the host's)."""
    print(f"[gpu-sources] haser {case}", flush=True)
    if case == "corner":
        cfg = _haser_cfg()
    else:
        cfg = _haser_cfg(species=4, gmin=(-45.0, -35.0, -35.0), gmax=(35.0, 25.0, 35.0),
                         dims=(8, 6, 7))
    d = host.derive(cfg)
    rp_ref, sp_ref = _haser_expect(cfg, d, 77, 3)
    assert 100 < len(sp_ref) < 20000
    dom = host.Domain(cfg, max_particles=50000, d=d)
    n = dom.produce(77, 3)
    rp, sp = dom.particles()
    print(f"[gpu-sources] haser {case}: {n} particles", flush=True)
    assert n == len(sp_ref) == len(sp)
    assert np.array_equal(sp, sp_ref)
    err = np.abs(rp[:, :6] - rp_ref[:, :6]) / np.maximum(np.abs(rp_ref[:, :6]), 1.0)
    assert err.max() <= 1e-13, err.max()
    assert np.array_equal(rp[:, 7], np.arange(n, dtype=np.float64))   # ids in slot order
    # split form: same particles through the integer-cell + FP32-offset store
    dom2 = host.Domain(cfg, max_particles=50000, d=d, precision=_lib.PICARD_FP32_SPLIT)
    assert dom2.produce(77, 3) == n
    r2, s2 = dom2.particles()
    assert np.array_equal(s2, sp_ref)
    assert np.abs(r2[:, :3] - rp_ref[:, :3]).max() <= 1e-5 * float(d.dxyz[0])
    assert np.abs(r2[:, 3:6] - rp_ref[:, 3:6]).max() <= 1e-6 * np.abs(rp_ref[:, 3:6]).max()
    dom.close()
    dom2.close()


def test_haser_in_step():
    """picard_step runs the source after the position push (picard.f90:566):
    a step with production on appends what produce() appends for that step,
    counted 1, 2, ... by the steps."""
    print("[gpu-sources] haser in step", flush=True)
    cfg = _haser_cfg(scale=0.5)
    d = host.derive(cfg)
    a = host.Domain(cfg, max_particles=100000, d=d)
    a.calc_distance()
    a.set_production(99)
    st1 = a.step()
    st2 = a.step()
    b = host.Domain(cfg, max_particles=100000, d=d)
    n1 = b.produce(99, 1)
    n2 = b.produce(99, 2)
    assert n1 > 0 and n2 > 0
    assert st1.num_local + st1.removed == n1
    assert st2.num_local + st2.removed == st1.num_local + n2
    a.close()
    b.close()


def _philox_draws(seed, tags):
    """The device's per-ion streams: key = seed, counter (slot within the
    source block, 0, tag of the block, block index); `tags` maps the
    restatement's face number to the block's tag."""
    from oracle import sources
    from tests.philox import uniforms

    def fn(face, local):
        return uniforms(seed, np.uint64(local & 0xFFFFFFFF), np.uint64(local >> 32), tags[face],
                        48)
    return sources.PerIonDraws(fn)


def _plasma(cfg, d):
    return dict(species=cfg.species, ipp=d.ipp, epp=d.epp, hpp=d.hpp, kelvin=d.kelvin,
                mass=d.mass, v0=d.v0[:, :3], ve=d.ve[:3], tB=d.tB[:, :3])


def _four_species(flow):
    return host.Config(nx=6, ny=5, nz=4, species=4, flow=flow, gmin=(-30.0, -27.5, -18.0),
                       gmax=(30.0, 27.5, 18.0), ppc=(0, 4, 2, 1, 0, 0, 0, 0),
                       density=(0.0, 5.0e6, 1.0e6, 5.0e5, 0, 0, 0, 0),
                       charge3=(2 * host.EC, -host.EC) + (host.EC,) * 4,
                       kelvin=(116045.0, 58022.0, 2.0e5, 3.0e4, 0, 0, 0, 0))


def _compare_sources(rp, sp, ro, so, d, label):
    assert rp.shape[0] == ro.shape[0], label
    assert np.array_equal(sp, so), label
    # positions: lower edge + dxyz * u in IEEE arithmetic on both sides
    assert np.array_equal(rp[:, :3], ro[:, :3]), label
    # velocities: the device's log/sin/sqrt against the C library's
    err = np.abs(rp[:, 3:6] - ro[:, 3:6]) / np.maximum(np.abs(ro[:, 3:6]).max(axis=1,
                                                                           keepdims=True), 1.0)
    assert err.max() <= 1e-13, (label, err.max())
    assert np.array_equal(rp[:, 7], np.arange(rp.shape[0], dtype=np.float64)), label


@pytest.mark.parametrize("flow", [(True, False, False), (True, True, True), (False, True, True),
                                  (True, False, True)])
def test_device_sources_match_oracle(flow):
    """Device ParticlesInitialize and BCinflow against the oracle's
    restatement (bit-exact against the reference build on the runtime's
    stream) driven by the device's Philox streams: counts, slot order and
    species exactly, positions exactly, velocities to 1e-13 of the
    particle's speed scale; the split form stores the same particles."""
    from oracle import sources
    print(f"[gpu-sources] oracle pin flow={flow}", flush=True)
    cfg = _four_species(flow)
    d = host.derive(cfg)
    g = _oracle_grid(cfg)
    pl = _plasma(cfg, d)
    seed, step = 20261207, 3
    ro, so = sources.particles_initialize(g, pl, _philox_draws(seed, {0: 0xA11}))
    dom = host.Domain(cfg, max_particles=200000, d=d)
    dom.generate_initial(seed)
    rp, sp = dom.particles()
    _compare_sources(rp, sp, ro, so, d, "initialize")
    n0 = rp.shape[0]
    tags = {f: 0x1000 + 8 * step + f for f in range(6)}
    ri, si = sources.bc_inflow(g, pl, _philox_draws(seed, tags))
    dom.generate_inflow(seed, step)
    rp2, sp2 = dom.particles(n0)
    rp2[:, 7] -= n0                      # ids continue from the initial load
    _compare_sources(rp2, sp2, ri, si, d, "inflow")
    sd = host.Domain(cfg, max_particles=200000, d=d, precision=_lib.PICARD_FP32_SPLIT)
    sd.generate_initial(seed)
    sd.generate_inflow(seed, step)
    r3, s3 = sd.particles()
    assert np.array_equal(s3, np.concatenate([so, si]))
    ref = np.concatenate([ro, ri])
    assert np.abs(r3[:, :3] - ref[:, :3]).max() <= 1e-5 * float(d.dxyz.max())
    assert np.abs(r3[:, 3:6] - ref[:, 3:6]).max() <= 1e-6 * np.abs(ref[:, 3:6]).max()
    dom.close()
    sd.close()


def _oracle_grid(cfg):
    import oracle
    return oracle.Grid(cfg.nx, cfg.ny, cfg.nz, cfg.species, flow=cfg.flow, gmin=cfg.gmin,
                       gmax=cfg.gmax)
