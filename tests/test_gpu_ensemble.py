"""The ensemble (include/picard_ensemble.h) on the GPU.

* Each simulation against the single-domain FP64 form (bit-exact against the
  oracle) run from the same device-generated plasma (same seed, same Philox
  streams for the initial load and every step's inflow): potential within
  1e-5 and densities within 1e-4 relative, E within 1e-4, per step.
* A simulation's result is the same, bit for bit (fields, and particles
  matched by id), alone and in a mix of others, whatever the slot, with
  others joining and leaving between steps, under other tile shapes, with and
  without graphs, and with the autotuner run between steps.
* Capacity refusals are statuses, not faults."""
import numpy as np
import pytest

from picard_amd import _lib, ensemble, host

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-ensemble] {msg}", flush=True)


def _spec(n, seed, ppc=4, dens=5e6, T=116045.0, drift=4e5, B=(0.0, 0.0, 5e-9), ny=None, nz=None):
    return ensemble.SimSpec(nx=n, ny=ny or n, nz=nz or n, ppc=ppc,
                            dx=ensemble.debye_length(dens, T), density=dens, kelvin=T,
                            drift=drift, B=B, seed=seed)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


FIELDS = (_lib.FIELD_U, _lib.FIELD_E, _lib.FIELD_N, _lib.FIELD_C)


def _state(ens, slot):
    rp, sp = ens.particles(slot)
    return [ens.field(slot, w) for w in FIELDS] + [rp, sp]


def _same(a, b, label):
    names = ("U", "E", "N", "C", "particles", "species")
    for x, y, nm in zip(a, b, names):
        assert x.shape == y.shape, (label, nm, x.shape, y.shape)
        assert np.array_equal(x, y), (label, nm, float(np.abs(x - y).max()))


def test_sim_matches_fp64_domain():
    """One simulation (20 x 18 x 16 cells, oblique B) in the ensemble and the
    same configuration in the single-domain FP64 form, 4 steps from the same
    device-generated plasma."""
    spec = _spec(20, 777, ppc=6, B=(2e-9, -3e-9, 4e-9), ny=18, nz=16)
    ens = ensemble.Ensemble.sized_for([spec], steps_room=6)
    slot = ens.add(spec)
    hx = [0.5 * spec.dx * n for n in (spec.nx, spec.ny, spec.nz)]
    cfg = host.Config(nx=spec.nx, ny=spec.ny, nz=spec.nz, species=2, flow=(True, False, False),
                      gmin=tuple(-h for h in hx), gmax=tuple(hx),
                      density=(0.0, spec.density, 0, 0, 0, 0, 0, 0),
                      ppc=(0, spec.ppc, 0, 0, 0, 0, 0, 0),
                      kelvin=(spec.kelvin,) * 2 + (0.0,) * 6, v0=(spec.drift, 0.0, 0.0),
                      B0=tuple(spec.B))
    d = host.derive(cfg)
    dom = host.Domain(cfg, d=d)
    dom.generate_initial(spec.seed)
    dom.calc_distance()
    # the initial loads are the same particles
    r0, s0 = ens.particles(slot)
    rd, sd = dom.particles()
    assert np.array_equal(s0, sd)
    assert np.abs(r0[:, :3] - rd[:, :3]).max() <= 1e-5 * spec.dx
    assert np.array_equal(r0[:, 7], rd[:, 7])
    for it in range(1, 5):
        st = ens.step()
        dom.generate_inflow(spec.seed, it)
        sf = dom.step()
        e = {nm: _rel(ens.field(slot, w), dom.field(w)) for nm, w in
             (("phi", _lib.FIELD_U), ("E", _lib.FIELD_E), ("N", _lib.FIELD_N))}
        es = ens.status(slot)
        _say(f"vs fp64 step {it}: n ens {es.num_local} fp64 {sf.num_local}, sweeps ens "
             f"{es.sweeps} fp64 {sf.sweeps}, phi {e['phi']:.2e} E {e['E']:.2e} N {e['N']:.2e}")
        assert st.sims == 1 and es.steps == it
        assert e["phi"] <= 1e-5 and e["N"] <= 1e-4 and e["E"] <= 1e-4, (it, e)
        assert abs(es.num_local - sf.num_local) <= 4
        assert abs(es.sweeps - sf.sweeps) <= 12, (es.sweeps, sf.sweeps)
    dom.close()
    ens.close()


def _run_alone(spec, steps, **kw):
    ens = ensemble.Ensemble.sized_for([spec], steps_room=steps + 2, **kw)
    slot = ens.add(spec)
    for _ in range(steps):
        ens.step()
    out = _state(ens, slot)
    sweeps = ens.status(slot).sweeps
    ens.close()
    return out, sweeps


def test_alone_and_in_the_mix():
    a = _spec(16, 11, ppc=4)
    others = [_spec(24, 12, ppc=5, dens=2e6, T=3e5, drift=6e5, B=(1e-9, 0, -2e-9)),
              _spec(12, 13, ppc=8, dens=1.2e7, T=2e4, drift=3.5e5),
              _spec(20, 14, ppc=4, ny=10, nz=14)]
    steps = 3
    alone, sw = _run_alone(a, steps)
    _say(f"alone: {steps} steps, last solve {sw} sweeps")
    mix = ensemble.Ensemble.sized_for([a] + others, steps_room=steps + 2)
    s1 = mix.add(others[0])
    s2 = mix.add(others[1])
    sa = mix.add(a)                               # not the first slot
    mix.step()
    mix.tune(8)                                   # autotuner between steps
    _say(f"mix: tile after tuning {mix.get_tile().cells_per_thread} cells/thread, "
         f"{mix.get_tile().blocks_per_cu} blocks/CU")
    mix.remove(s1)                                # one leaves ...
    s3 = mix.add(others[2])                       # ... one joins
    for _ in range(steps - 1):
        st = mix.step()
        assert st.sims == 3
    _same(_state(mix, sa), alone, "in the mix")
    assert mix.status(sa).sweeps == sw
    assert mix.status(s3).steps == steps - 1 and mix.status(s2).steps == steps
    mix.close()


@pytest.mark.parametrize("cpt,bpc,graphs", [(1, 2, False), (2, 8, True), (4, 4, False), (8, 2, True)])
def test_tile_shapes_and_graphs_same_result(cpt, bpc, graphs):
    a = _spec(18, 21, ppc=4, B=(0.0, 3e-9, 3e-9))
    b = _spec(14, 22, ppc=6, dens=3e6, T=5e4)
    steps = 2
    ref_a, _ = _run_alone(a, steps)
    ens = ensemble.Ensemble.sized_for([b, a], steps_room=steps + 2, graphs=graphs)
    ens.set_tile(cpt, bpc)
    ens.add(b)
    sa = ens.add(a)
    for _ in range(steps):
        ens.step()
    _same(_state(ens, sa), ref_a, f"tile {cpt}x{bpc} graphs={graphs}")
    ens.close()


def test_autotuner_candidates():
    specs = [_spec(16 + 4 * q, 30 + q, ppc=4) for q in range(4)]
    ens = ensemble.Ensemble.sized_for(specs)
    for s in specs:
        ens.add(s)
    tile, times = ens.tune(16)
    _say(f"tuner: {len(times)} candidates, us per sweep {np.round(times, 2).tolist()}, "
         f"chosen {tile.cells_per_thread}x{tile.blocks_per_cu}")
    assert len(times) == 12 and np.all(times > 0)
    assert tile.cells_per_thread in (1, 2, 4, 8) and 1 <= tile.blocks_per_cu <= 8
    t = ens.get_tile()
    assert (t.cells_per_thread, t.blocks_per_cu) == (tile.cells_per_thread, tile.blocks_per_cu)
    st = ens.step()
    assert st.sims == 4 and st.noconv == 0
    ens.close()


def test_capacity_statuses():
    a = _spec(16, 41, ppc=4)
    g, parts = ensemble.footprint(a)
    ens = ensemble.Ensemble(max_sims=1, max_cells=a.cells, max_particles=parts + 100)
    slot = ens.add(a)
    with pytest.raises(_lib.PicardError) as ex:          # no slot left
        ens.add(_spec(16, 42))
    assert ex.value.rc == _lib.PICARD_ERR_CAPACITY
    st = ens.step()                                      # one step's inflow fits
    assert st.sims == 1
    ens.remove(slot)
    ens.close()
    small = ensemble.Ensemble(max_sims=4, max_cells=a.cells - 1, max_particles=10 * parts)
    with pytest.raises(_lib.PicardError) as ex:          # not enough cells
        small.add(a)
    assert ex.value.rc == _lib.PICARD_ERR_CAPACITY
    small.close()
    tight = ensemble.Ensemble(max_sims=4, max_cells=4 * a.cells, max_particles=parts - 1)
    with pytest.raises(_lib.PicardError) as ex:          # not enough particle slots
        tight.add(a)
    assert ex.value.rc == _lib.PICARD_ERR_CAPACITY
    tight.close()
