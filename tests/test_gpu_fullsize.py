"""The benchmark workload at its full size (256^3 cells, H+ and e- at 32 per
cell, 1.07e9 particles, Haser source): one step of each form, checked through
properties that hold at any size, since the oracle cannot follow 1e9
particles in seconds.

- Deposit: the interior density sums give back the particle count,
  sum_s sum N(s) / n2p(s) = num_local (every weight stays in the interior
  after the guard folds, ParticlesGridify.f90:129-285), to FP64 rounding.
- Solve: the final potential meets the stop rule of CalcEpotential.f90:60-61,
  |b - A U| <= maxerr |b| with b = -C/(dV eps0), evaluated on the host in
  FP64 (with the final periodic images; the rule itself uses the previous
  iterate's, so the bound is taken with a factor 1.5).
- Field: E is the central difference of the final U with BCfields applied,
  bit for bit in the FP64 form and to FP32 rounding in the split form (the
  oracle's CalcEfield/BCfields on the device's U).
- The two forms: potential within 1e-5 and densities within 1e-4 relative
  (the precision form's tolerances), on the same device-generated plasma."""
import numpy as np
import pytest

import oracle
from picard_amd import _lib, host
from tests.pic_driver import oracle_grid

pytestmark = pytest.mark.gpu

SEED = 20261202
EPS0 = 8.854187817e-12


def _say(msg):
    print(f"[gpu-fullsize] {msg}", flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _one_step(cfg, precision):
    d = host.derive(cfg)
    ncell = cfg.nx * cfg.ny * cfg.nz
    per_cell = int(d.ppc[:cfg.species].sum())
    cap = int(1.12 * per_cell * ncell) + 64 * per_cell * (cfg.ny + 2) * (cfg.nz + 2) + (1 << 20)
    dom = host.Domain(cfg, max_particles=cap, d=d, precision=precision)
    dom.generate_initial(SEED)
    dom.calc_distance()
    dom.set_production(SEED)
    dom.generate_inflow(SEED, 1)
    st = dom.step()
    dom.synchronize()
    out = {k: dom.field(w) for k, w in (("U", _lib.FIELD_U), ("C", _lib.FIELD_C),
                                         ("E", _lib.FIELD_E), ("N", _lib.FIELD_N))}
    out["stats"] = (st.num_local, st.removed, st.sweeps)
    dom.close()
    return d, out


def _check_form(cfg, d, f, name, exact_e):
    g = oracle_grid(cfg)
    num_local, removed, sweeps = f["stats"]
    _say(f"{name}: num_local {num_local}, removed {removed}, sweeps {sweeps}")
    N = f["N"]
    back = sum(N[1:-1, 1:-1, 1:-1, s].sum() / d.n2p[s] for s in range(cfg.species))
    _say(f"{name}: density sums give {back:.6f} particles")
    assert abs(back - num_local) <= 1e-9 * num_local
    U, C = f["U"], f["C"]
    dx, dy, dz = (float(v) for v in d.dxyz[:3])
    b = -C[1:-1, 1:-1, 1:-1] / (d.dV * EPS0)
    c = U[1:-1, 1:-1, 1:-1]
    lap = ((U[1:-1, 1:-1, 2:] + U[1:-1, 1:-1, :-2] - 2 * c) / dx ** 2
           + (U[1:-1, 2:, 1:-1] + U[1:-1, :-2, 1:-1] - 2 * c) / dy ** 2
           + (U[2:, 1:-1, 1:-1] + U[:-2, 1:-1, 1:-1] - 2 * c) / dz ** 2)
    rr = float(np.linalg.norm(b - lap) / np.linalg.norm(b))
    _say(f"{name}: relative residual of the final potential {rr:.3e} (maxerr {cfg.maxerr:g})")
    assert rr <= 1.5 * cfg.maxerr
    E = oracle.calc_efield(g, U)
    oracle.bc_fields(g, E)
    if exact_e:
        assert np.array_equal(f["E"], E)
    else:
        er = _rel(f["E"], E)
        _say(f"{name}: E against the central difference of U: {er:.3e}")
        assert er <= 1e-7


def test_fullsize_step_properties():
    cfg = host.config2(256)
    _say("split form, one step at 256^3")
    d, fs = _one_step(cfg, _lib.PICARD_FP32_SPLIT)
    _check_form(cfg, d, fs, "split", exact_e=False)
    _say("FP64 form, one step at 256^3")
    d, f64 = _one_step(cfg, _lib.PICARD_FP64)
    _check_form(cfg, d, f64, "fp64", exact_e=True)
    assert fs["stats"][0] == f64["stats"][0]            # the same particles after the step
    phi = _rel(fs["U"], f64["U"])
    dens = _rel(fs["N"], f64["N"])
    _say(f"split vs FP64: phi {phi:.3e}, densities {dens:.3e}, sweeps {fs['stats'][2]} / "
         f"{f64['stats'][2]}")
    assert phi <= 1e-5
    assert dens <= 1e-4
