"""Host side of the ensemble (no device work): the derived constants of a
simulation (picard_sim_derive) are the single-domain host's (host.derive,
picard.f90:333-466) bit for bit, the workload's parameter draws stay in their
stated ranges, and out-of-range parameters are refused."""
import ctypes as ct

import numpy as np
import pytest

from picard_amd import _lib, ensemble, host


def _config_of(spec):
    hx = [0.5 * spec.dx * n for n in (spec.nx, spec.ny, spec.nz)]
    return host.Config(nx=spec.nx, ny=spec.ny, nz=spec.nz, species=2, flow=(True, False, False),
                       gmin=tuple(-h for h in hx), gmax=tuple(hx),
                       density=(0.0, spec.density, 0, 0, 0, 0, 0, 0),
                       ppc=(0, spec.ppc, 0, 0, 0, 0, 0, 0),
                       kelvin=(spec.kelvin,) * 2 + (0.0,) * 6, v0=(spec.drift, 0.0, 0.0),
                       B0=tuple(spec.B), dt_wpe=spec.dt_wpe)


@pytest.mark.parametrize("k", range(6))
def test_sim_derive_matches_host(k):
    spec = ensemble.draw_ensemble(6, seed=5)[k]
    d = ensemble.derive(spec)
    ref = host.derive(_config_of(spec))
    assert d.dxyz == ref.dxyz[0] == ref.dxyz[1] == ref.dxyz[2]
    assert d.dV == ref.dV
    for s in range(2):
        assert d.charge[s] == ref.charge[s]
        assert d.n2p[s] == ref.n2p[s]
        assert d.eta[s] == ref.eta[s]
        for c in range(3):
            assert d.tB[s][c] == ref.tB[s, c]
            assert d.sB[s][c] == ref.sB[s, c]
        assert d.tB4[s] == ref.tB[s, 3]
    for c in range(3):
        assert d.ve[c] == ref.ve[c]
    assert d.dt == ref.dt and d.omega_pe == ref.omega_pe


def test_draws_in_range():
    specs = ensemble.draw_ensemble(64)
    assert len(specs) == 64 and len({s.seed for s in specs}) == 64
    for s in specs:
        assert 16 <= s.nx <= 96 and s.nx == s.ny == s.nz
        assert 4 <= s.ppc <= 32
        assert 1e6 <= s.density <= 20e6
        assert 1 * ensemble.EV_K <= s.kelvin <= 50 * ensemble.EV_K
        assert 3e5 <= s.drift <= 7e5
        assert 1e-9 <= np.linalg.norm(s.B) <= 10e-9 * (1 + 1e-12)
        assert s.dx == pytest.approx(ensemble.debye_length(s.density, s.kelvin))
    # the 32^3 workload's cell: 10 m at 10 eV and 5 cm^-3 (picard.f90 defaults)
    assert ensemble.debye_length(5e6, 116045.0) == pytest.approx(10.5, rel=0.01)


def test_footprint_and_refusals():
    spec = ensemble.draw_ensemble(1, seed=3)[0]
    g, n = ensemble.footprint(spec)
    assert n == spec.cells * 2 * spec.ppc + 4 * spec.ny * spec.nz * spec.ppc
    assert g > spec.cells * 100
    bad = ensemble.SimSpec(nx=4, ny=16, nz=16, ppc=4, dx=10.0, density=5e6, kelvin=1e5,
                           drift=4e5, B=(0.0, 0.0, 5e-9), seed=1)
    d = _lib.SimDerived()
    p = bad.cstruct()
    assert _lib.lib().picard_sim_derive(ct.byref(p), ct.byref(d)) == _lib.PICARD_ERR_ARG
