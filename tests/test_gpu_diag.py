"""Device diagnostics (picard.f90:649-672) against the oracle's sequential
sums on the same particles and fields: both forms, after whole steps.  The
device sums in a fixed tree order, so the bar is 1e-12 relative."""
import numpy as np
import pytest

import oracle
from picard_amd import _lib, host
from tests.pic_driver import oracle_grid, small_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [_lib.PICARD_FP64, _lib.PICARD_FP32_SPLIT])
def test_diagnostics_match_oracle(precision):
    print(f"[gpu-diag] precision={precision}", flush=True)
    cfg = small_config(n=12, ppc=6)
    d = host.derive(cfg)
    dom = host.Domain(cfg, max_particles=200000, d=d, precision=precision)
    dom.generate_initial(21)
    dom.calc_distance()
    for it in (1, 2):
        dom.generate_inflow(21, it)
        dom.step()
    got = dom.diagnostics()
    rp, sp = dom.particles()
    E = dom.field(_lib.FIELD_E)
    uP, uE, uB = oracle.diagnostics(oracle_grid(cfg), rp, sp, E, d.n2p, d.mass, d.E0, d.B0)
    print(f"[gpu-diag] {got} oracle {(uP, uE, uB)}", flush=True)
    assert got["num_local"] == got["num_global"] == rp.shape[0]
    assert abs(got["particle_energy"] - uP) <= 1e-12 * uP
    assert abs(got["e_energy"] - uE) <= 1e-12 * uE
    assert abs(got["b_energy"] - uB) <= 1e-12 * uB
    dom.close()
