"""The 32^3 workload (H+ and e- at 8 per cell, 400 km/s drift, 10 eV,
B = (0, 0, 5 nT), x in/outflow, dt = 0.1/w_pe, Jacobi tolerance 1e-6) for all
100 steps of the main loop (picard.f90:554-633) through picard_step, bit for
bit against the oracle's per-step digests (tests/golden/config1_digests.json,
written by tests/golden/make_config1_digests.py): num_local, BCoutflow
removals, Jacobi sweep counts, and SHA-256 of the particle, N, C, U and E
byte images after every step.  The inflow is the seeded host source, so the
particles appended each step are the oracle's."""
import json
import os

import numpy as np
import pytest

from picard_amd import _lib, host
from tests.pic_driver import step_digest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config1_digests.json")


def _say(msg):
    print(f"[gpu-config1] {msg}", flush=True)


def test_config1_100_steps_bit_exact():
    with open(FIXTURE) as f:
        want = json.load(f)["steps"]
    assert len(want) == 100
    cfg = host.config1()
    d = host.derive(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    dom = host.Domain(cfg, d=d)
    dom.append(rp, sp)
    dom.calc_distance()
    for w in want:
        it = w["step"]
        dom.append(*host.bc_inflow(cfg, d, rng))                  # picard.f90:560
        st = dom.step()
        got_rp, got_sp = dom.particles()
        got = step_digest(it, dict(num_local=st.num_local, removed=st.removed, sweeps=st.sweeps),
                          got_rp, got_sp, dom.field(_lib.FIELD_N), dom.field(_lib.FIELD_C),
                          dom.field(_lib.FIELD_U), dom.field(_lib.FIELD_E))
        if it % 10 == 0 or it == 1:
            _say(f"step {it}: n={st.num_local} removed={st.removed} sweeps={st.sweeps}")
        # the hashes and counts are exact; U's 2-norm is a readable value whose
        # last bit depends on the host's BLAS summation order
        bad = [k for k in w if k != "U_2" and got[k] != w[k]]
        assert not bad, f"step {it}: {bad} differ from the oracle (got {[got[k] for k in bad]})"
        assert abs(got["U_2"] - w["U_2"]) <= 1e-12 * abs(w["U_2"]), (it, got["U_2"], w["U_2"])
    dom.close()
