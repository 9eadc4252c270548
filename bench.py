#!/usr/bin/env python3
"""Benchmark of the PIC time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is one pass of the main loop (picard.f90:556-631, no output step):
BCinflow's fresh guard-cell plasma is appended, then AdvancePosition,
BCoutflow, ParticlesTransfer, ParticlesGridify, BCpotentials, CalcEpotential,
CalcEfield, BCfields and AdvanceVelocity run on the device.

metric: particle-steps/s = sum over timed steps of num_local at the velocity
push, divided by the wall time of the timed steps (inputs resident in HBM;
the inflow particles of every step are generated on the host beforehand).

Prints one JSON line (rank 0, N = 1) with the roofline of the dominant kernel
(the Jacobi sweep) and a CPU baseline of the oracle on a bounded sample."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s FP32-split vs FP64 at ≤1e-5 rel φ err; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gpu_run(cfg, steps, warmup):
    from picard_amd import host
    d = host.derive(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    inflow = [host.bc_inflow(cfg, d, rng) for _ in range(warmup + steps)]
    dom = host.Domain(cfg, d=d)
    dom.append(rp, sp)
    dom.calc_distance()
    dom.synchronize()
    for it in range(warmup):
        dom.append(*inflow[it])
        st = dom.step()
        log(f"warmup {it + 1}/{warmup}: n={st.num_local} sweeps={st.sweeps}")
    dom.synchronize()
    dom.timing(True)
    pushed = 0
    sweeps = []
    t0 = time.perf_counter()
    for it in range(steps):
        dom.append(*inflow[warmup + it])
        st = dom.step()
        pushed += st.pushed
        sweeps.append(st.sweeps)
    dom.synchronize()
    elapsed = time.perf_counter() - t0
    jac_ms, jac_launches = dom.timing_read()
    dom.timing(False)
    log(f"timed {steps} steps in {elapsed:.3f} s, sweeps/step {sweeps}")
    dom.close()
    return dict(pushed=pushed, elapsed=elapsed, sweeps=sweeps, jac_ms=jac_ms,
                jac_launches=jac_launches)


def cpu_baseline(cfg, budget_s=20.0):
    """The oracle (C restatement, 1 core) on the same workload: a bounded
    number of steps after one warmup step."""
    import oracle
    from picard_amd import host
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from tests.pic_driver import OracleSim
    d = host.derive(cfg)
    rng = np.random.default_rng(cfg.seed)
    rp, sp = host.particles_initialize(cfg, d, rng)
    sim = OracleSim(cfg, d, rp, sp)
    sim.append(*host.bc_inflow(cfg, d, rng))
    sim.step()
    pushed, steps = 0, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 50:
        sim.append(*host.bc_inflow(cfg, d, rng))
        sim.step()
        pushed += sim.n
        steps += 1
    el = time.perf_counter() - t0
    _ = oracle
    return dict(value=pushed / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of the {cfg.nx}x{cfg.ny}x{cfg.nz} workload after 1 warmup "
                       f"step, oracle C restatement (reference order, FP64), 1 thread, "
                       f"{el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the domain-decomposed design is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from picard_amd import host
    cfg = host.config1()
    r = gpu_run(cfg, args.steps, args.warmup)
    value = r["pushed"] / r["elapsed"]
    ncell = cfg.nx * cfg.ny * cfg.nz
    bytes_per_launch = 24.0 * ncell           # read U, read b, write U per interior cell (FP64)
    avg_s = (r["jac_ms"] / 1e3) / max(r["jac_launches"], 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"FP64 form, {cfg.nx}x{cfg.ny}x{cfg.nz} cells dx=10 m, H+ and e- at 8 "
                        "per cell (~524k particles), drifting Maxwellian 400 km/s, 10 eV, "
                        "B=(0,0,5 nT), x in/outflow, y/z periodic, dt=0.1/w_pe, Jacobi tol 1e-6",
            "form": "fp64-reference-order",
            "grid": [cfg.nx, cfg.ny, cfg.nz],
            "mean_jacobi_sweeps_per_step": float(np.mean(r["sweeps"])),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_jacobi_pass",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": avg_s * 1e6,
            "launches": r["jac_launches"],
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
