"""Writes the per-step digests of the 32^3 workload (H+ and e- at 8 per cell,
400 km/s drift, 10 eV, B = (0, 0, 5 nT), x in/outflow, dt = 0.1/w_pe, Jacobi
tolerance 1e-6) over 100 steps of the main loop (picard.f90:554-633), run by
the oracle from the seeded host sources of picard_amd.host.

    python tests/golden/make_config1_digests.py [--steps 100]

Output: tests/golden/config1_digests.json, one record per step: num_local,
BCoutflow removals, Jacobi sweeps, and SHA-256 of the particle
(real_particles(8, n), int_particles(n)), N, C, U and E byte images after
AdvanceVelocity, with the max-norm and 2-norm of U.  Data only."""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from picard_amd import host  # noqa: E402
from tests.pic_driver import OracleSim, step_digest  # noqa: E402

OUT = os.path.join(HERE, "config1_digests.json")


def run(steps, progress=None):
    cfg = host.config1()
    d = host.derive(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    orc = OracleSim(cfg, d, rp, sp)
    out = []
    for it in range(1, steps + 1):
        orc.append(*host.bc_inflow(cfg, d, rng))          # picard.f90:560
        st = orc.step()
        out.append(step_digest(it, st, orc.rp, orc.sp, orc.N, orc.C, orc.U, orc.E))
        if progress:
            progress(out[-1])
    return cfg, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100)
    args = ap.parse_args()
    t0 = time.perf_counter()
    cfg, recs = run(args.steps, lambda r: print(f"step {r['step']}: n={r['num_local']} "
                                                f"removed={r['removed']} sweeps={r['sweeps']} "
                                                f"({time.perf_counter() - t0:.0f} s)", flush=True))
    with open(OUT, "w") as f:
        json.dump({"workload": "32^3 cells dx=10 m, H+ and e- at 8 per cell, 400 km/s, 10 eV, "
                               "B=(0,0,5 nT), x in/outflow, y/z periodic, dt=0.1/w_pe, "
                               "Jacobi tol 1e-6, seed %d" % cfg.seed,
                   "steps": recs}, f, indent=0)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
