"""The FP64 form with the order-free deposit and face sums
(PICARD_DEPOSIT_FIXED_POINT: int64 fixed-point CIC weights and FFT
convolutions for BCpotentials) against the reference-order FP64 form, which is
bit-exact against the oracle.  Stated tolerance: the deposit and the face
potentials within 1e-12 relative of the reference order; whole steps keep the
particle count and the sweep count within one, phi within 1e-9."""
import numpy as np
import pytest

from picard_amd import _lib, host
from tests.pic_driver import small_config

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("flow", [(True, False, False), (True, True, True)])
def test_order_free_fp64_against_reference_order(flow):
    cfg = small_config(14, ppc=6, flow=flow, steps_seed=41)
    d = host.derive(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    # move the electrons off their ions (by under half a cell) so C != 0
    e = sp == 1
    rp[e, :3] += rng.uniform(-0.4, 0.4, (int(e.sum()), 3)) * d.dxyz[None, :]
    doms = [host.Domain(cfg, d=d), host.Domain(cfg, d=d, order_free=True)]
    for dom in doms:
        dom.append(rp, sp)
        dom.calc_distance()
    # one deposit and face sum on identical particles
    for dom in doms:
        dom.particles_gridify()
        dom.bc_potentials()
    ref, new = doms
    eN = float(np.abs(new.field(_lib.FIELD_N) - ref.field(_lib.FIELD_N)).max() /
               np.abs(ref.field(_lib.FIELD_N)).max())
    eC = _rel(new.field(_lib.FIELD_C), ref.field(_lib.FIELD_C))
    eU = float(np.abs(new.field(_lib.FIELD_U) - ref.field(_lib.FIELD_U)).max() /
               np.abs(ref.field(_lib.FIELD_U)).max())
    print(f"[gpu-order-free] flow={flow}: N {eN:.2e} C {eC:.2e} faces {eU:.2e}", flush=True)
    assert eN <= 1e-12 and eU <= 1e-12
    assert eC <= 1e-9            # C is a difference of the two species' densities
    # whole steps from the same state
    for it in range(1, 4):
        inflow = host.bc_inflow(cfg, d, rng)
        st = []
        for dom in doms:
            dom.append(*inflow)
            st.append(dom.step())
        e = _rel(new.field(_lib.FIELD_U), ref.field(_lib.FIELD_U))
        print(f"[gpu-order-free] flow={flow} step {it}: n {st[0].num_local}/{st[1].num_local} "
              f"sweeps {st[0].sweeps}/{st[1].sweeps} phi {e:.2e}", flush=True)
        assert st[0].num_local == st[1].num_local
        assert abs(st[0].sweeps - st[1].sweeps) <= 1
        assert e <= 1e-9
    for dom in doms:
        dom.close()
