#!/usr/bin/env python3
"""Benchmark of the PIC time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 256]
                    [--form both|split|fp64] [--no-cpu-baseline]
    python bench.py --solve-only SWEEPS [--form split|fp64]   (profiling)
    python bench.py --copy-cal                                (profiling)
    python bench.py --error-steps N [--grid G]               (split vs FP64 errors)

Workload: 256^3 cells of 10 m, H+ and e- at 32 per cell (1.07e9 particles),
drifting Maxwellian 400 km/s at 10 eV, B = (0, 0, 5 nT), x in/outflow, y/z
periodic, dt = 0.1/w_pe, Jacobi tolerance 1e-6, and a Haser source of 18 amu
ions at the centre.  Plasma, inflow and production are generated on the device
(Philox streams, same seeds for both forms).

A step is one pass of the main loop (picard.f90:556-631, no output step):
BCinflow, AdvancePosition, ParticlesProduce, BCoutflow, ParticlesTransfer,
ParticlesGridify, BCpotentials, CalcEpotential, CalcEfield, BCfields,
AdvanceVelocity.

metric: particle-steps/s = sum over the timed steps of num_local at the
velocity push / wall time of the timed steps (state resident in HBM).  The
FP32-split form gives `value`; the FP64 form (reference order, bit-exact
against the oracle) runs the same steps after it, and the line reports both,
their ratio, and the split form's error against the FP64 form after the first
step (phi, E, densities: normwise relative).

Prints one JSON line with the roofline of the dominant kernel (the split
form's FP32 Jacobi sweep, timed live with per-launch HIP events) and a CPU
baseline of the oracle on a bounded sample."""
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
SEED = 20261204


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _rel(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def run_form(cfg, precision, steps, warmup, capture_after=1):
    """W warmup + K timed steps of one form; returns timings and the fields
    after step `capture_after` (read outside the timed region)."""
    from picard_amd import _lib, host
    d = host.derive(cfg)
    ncell = cfg.nx * cfg.ny * cfg.nz
    per_cell = int(d.ppc[:cfg.species].sum())
    cap = int(1.12 * per_cell * ncell) + 64 * per_cell * (cfg.ny + 2) * (cfg.nz + 2) + (1 << 20)
    name = "split" if precision == _lib.PICARD_FP32_SPLIT else "fp64"
    t0 = time.perf_counter()
    dom = host.Domain(cfg, max_particles=cap, d=d, precision=precision)
    dom.generate_initial(SEED)
    dom.calc_distance()
    dom.set_production(SEED)
    dom.synchronize()
    log(f"{name}: setup {time.perf_counter() - t0:.1f} s, {dom.num_local} particles, cap {cap}")
    fields = None
    it = 0
    for _ in range(warmup):
        it += 1
        dom.generate_inflow(SEED, it)
        st = dom.step()
        dom.synchronize()
        log(f"{name}: warmup step {it}: n={st.num_local} sweeps={st.sweeps}")
        if it == capture_after:
            fields = {k: dom.field(w) for k, w in (("U", _lib.FIELD_U), ("E", _lib.FIELD_E),
                                                    ("N", _lib.FIELD_N))}
    dom.synchronize()
    dom.timing(True)
    pushed, sweeps = 0, []
    t0 = time.perf_counter()
    for _ in range(steps):
        it += 1
        dom.generate_inflow(SEED, it)
        st = dom.step()
        pushed += st.pushed
        sweeps.append(st.sweeps)
        log(f"{name}: timed step {it}: n={st.num_local} sweeps={st.sweeps} "
            f"({time.perf_counter() - t0:.1f} s)")
    dom.synchronize()
    elapsed = time.perf_counter() - t0
    jac_ms, jac_launches = dom.timing_read()
    dom.timing(False)
    if fields is None:
        fields = {k: dom.field(w) for k, w in (("U", _lib.FIELD_U), ("E", _lib.FIELD_E),
                                                ("N", _lib.FIELD_N))}
    dom.close()
    log(f"{name}: {steps} steps in {elapsed:.2f} s, sweeps {sweeps}, "
        f"solve kernel {1e3 * jac_ms / max(jac_launches, 1):.2f} us per launch, avg over {jac_launches}")
    return dict(pushed=pushed, elapsed=elapsed, sweeps=sweeps, jac_ms=jac_ms,
                jac_launches=jac_launches, fields=fields, captured_at=min(capture_after, it))


def cpu_baseline(budget_s=25.0):
    """The oracle (C restatement of the reference order, FP64, 1 thread) on a
    bounded sample of the workload: whole steps of a 48^3 block of it (same
    cell size, plasma and 32 particles per cell per species, no source)."""
    from picard_amd import host
    from tests.pic_driver import OracleSim
    cfg = host.config2(48, species=2, Qn=(0.0,) * 8)
    d = host.derive(cfg)
    rng = np.random.default_rng(SEED)
    rp, sp = host.particles_initialize(cfg, d, rng)
    sim = OracleSim(cfg, d, rp, sp)
    pushed, steps = 0, 0
    t0 = time.perf_counter()
    while steps < 1 or (time.perf_counter() - t0 < budget_s * 0.5 and steps < 20):
        sim.append(*host.bc_inflow(cfg, d, rng))
        sim.step()
        pushed += sim.n
        steps += 1
        log(f"cpu baseline: step {steps} n={sim.n} ({time.perf_counter() - t0:.1f} s)")
    el = time.perf_counter() - t0
    return dict(value=pushed / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} step(s) of a 48^3 block of the workload (10 m cells, H+ and e- "
                       f"at 32 per cell, {sim.n} particles, no Haser source); oracle C "
                       f"restatement in reference order, FP64, 1 thread, {el:.1f} s")


def solve_only(cfg, form, sweeps, inner=0):
    """Profiling mode: one field solve of the workload's grid on a seeded
    random charge, capped at `sweeps` sweeps (no particles), so that counter
    passes over every dispatch stay short; `inner` > 0 sets the split form's
    FP32 sweeps per FP64 refinement (else the library default)."""
    from picard_amd import _lib, host
    d = host.derive(cfg)
    prec = _lib.PICARD_FP32_SPLIT if form == "split" else _lib.PICARD_FP64
    dom = host.Domain(cfg, max_particles=1024, d=d, precision=prec)
    if inner > 0 and form == "split":
        dom.split_solver(inner)
    dom.calc_distance()
    rng = np.random.default_rng(SEED)
    dom.set_field(_lib.FIELD_C, rng.normal(size=(cfg.nz + 2, cfg.ny + 2, cfg.nx + 2)) * 1e-16)
    dom.bc_potentials()
    dom.synchronize()
    t0 = time.perf_counter()
    sw = dom.calc_epotential(maxerr=1e-30, maxit=max(sweeps // 2, 1))
    dom.synchronize()
    el = time.perf_counter() - t0
    dom.close()
    print(json.dumps({"mode": "solve-only", "form": form, "grid": [cfg.nx, cfg.ny, cfg.nz],
                      "sweeps": sw, "seconds": el}), flush=True)


def error_report(cfg, steps):
    """Error mode: N steps of both forms side by side from the same
    device-generated plasma (the workload at --grid, its Haser source
    included).  Per step, the normwise relative error of the split form's
    phi, E and densities against the FP64 form's; after the last step, the
    particle trajectories matched by id (slot 8 of the FP64 particle, the id
    of the split one): position and velocity differences over the particles
    both forms still hold."""
    from picard_amd import _lib, host
    d = host.derive(cfg)
    ncell = cfg.nx * cfg.ny * cfg.nz
    per_cell = int(d.ppc[:cfg.species].sum())
    cap = int(1.12 * per_cell * ncell) + 64 * per_cell * (cfg.ny + 2) * (cfg.nz + 2) + (1 << 20)
    doms = {}
    for name, prec in (("fp64", _lib.PICARD_FP64), ("split", _lib.PICARD_FP32_SPLIT)):
        dom = host.Domain(cfg, max_particles=cap, d=d, precision=prec)
        dom.generate_initial(SEED)
        dom.calc_distance()
        dom.set_production(SEED)
        doms[name] = dom
    per_step = []
    for it in range(1, steps + 1):
        st = {}
        for name, dom in doms.items():
            dom.generate_inflow(SEED, it)
            st[name] = dom.step()
        f = {name: {k: dom.field(w) for k, w in (("U", _lib.FIELD_U), ("E", _lib.FIELD_E),
                                                  ("N", _lib.FIELD_N))}
             for name, dom in doms.items()}
        a, b = f["split"], f["fp64"]
        row = dict(step=it, n=[st["split"].num_local, st["fp64"].num_local],
                   sweeps=[st["split"].sweeps, st["fp64"].sweeps],
                   phi_rel=_rel(a["U"], b["U"]), E_rel=_rel(a["E"], b["E"]), N_rel=_rel(a["N"], b["N"]))
        per_step.append(row)
        log(f"error report: step {it}: " + ", ".join(f"{k} {v}" for k, v in row.items() if k != "step"))
    ra, _ = doms["fp64"].particles()
    rb, _ = doms["split"].particles()
    for dom in doms.values():
        dom.close()
    ia, ib = np.argsort(ra[:, 7], kind="stable"), np.argsort(rb[:, 7], kind="stable")
    ra, rb = ra[ia], rb[ib]
    common, ka, kb = np.intersect1d(ra[:, 7], rb[:, 7], return_indices=True)
    pa, pb = ra[ka], rb[kb]
    dxv = np.linalg.norm(pa[:, :3] - pb[:, :3], axis=1)
    dvv = np.linalg.norm(pa[:, 3:6] - pb[:, 3:6], axis=1)
    vrms = float(np.sqrt(np.mean(np.sum(pa[:, 3:6] ** 2, axis=1))))
    dx = float(d.dxyz[0])
    traj = dict(matched=int(len(common)), only_fp64=int(len(ra) - len(common)),
                only_split=int(len(rb) - len(common)),
                dx_max_m=float(dxv.max()), dx_rms_m=float(np.sqrt(np.mean(dxv ** 2))),
                dx_max_cells=float(dxv.max() / dx),
                dv_max_rel=float(dvv.max() / vrms), dv_rms_rel=float(np.sqrt(np.mean(dvv ** 2)) / vrms))
    print(json.dumps({"mode": "error-report", "grid": [cfg.nx, cfg.ny, cfg.nz], "steps": steps,
                      "per_step": per_step, "trajectories": traj}), flush=True)


def copy_calibration():
    """Counter calibration: three device copies of 1 GiB (read 1 GiB, write
    1 GiB each), the plain copy kernel the HBM byte counters are scaled by."""
    import torch
    a = torch.zeros(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "copy-cal", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                      "copies": 3}), flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/hbm_traffic_256cube.json: FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by what it reads for a 1 GiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_256cube.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--form", choices=("both", "split", "fp64"), default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--solve-only", type=int, default=0, metavar="SWEEPS",
                    help="profiling: one capped field solve of --form (split or fp64) only")
    ap.add_argument("--inner", type=int, default=0, metavar="M",
                    help="with --solve-only: FP32 sweeps per FP64 refinement of the split form")
    ap.add_argument("--error-steps", type=int, default=0, metavar="N",
                    help="error mode: N steps of both forms from the same plasma, split vs FP64 "
                         "errors per step and particle trajectories after the last")
    ap.add_argument("--copy-cal", action="store_true",
                    help="profiling: the 1 GiB copies the byte counters are scaled by")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the domain "
              "decomposition it would use is described with the design); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from picard_amd import _lib, host
    cfg = host.config2(args.grid)
    if args.copy_cal:
        copy_calibration()
        return
    if args.solve_only:
        solve_only(cfg, "fp64" if args.form == "fp64" else "split", args.solve_only, args.inner)
        return
    if args.error_steps:
        error_report(cfg, args.error_steps)
        return
    runs = {}
    if args.form in ("both", "split"):
        runs["split"] = run_form(cfg, _lib.PICARD_FP32_SPLIT, args.steps, args.warmup)
    if args.form in ("both", "fp64"):
        runs["fp64"] = run_form(cfg, _lib.PICARD_FP64, args.steps, args.warmup)
    head = runs.get("split") or runs["fp64"]
    headname = "split" if "split" in runs else "fp64"
    value = head["pushed"] / head["elapsed"]
    ncell = cfg.nx * cfg.ny * cfg.nz
    # Sweeps per launch of each form's solve kernel: the split form runs two
    # FP32 sweeps per launch on rows of at most 256 cells, or wider rows with
    # x not periodic (the default inner length, 512, is even), the FP64 form
    # one pass per launch.
    spl = {"split": 2 if cfg.nx <= 256 or cfg.flow[0] else 1, "fp64": 1}
    # Algorithmic bytes per launch of the dominant kernel, per cell-sweep
    # times the cell-sweeps of a launch: an FP32 sweep reads the correction e
    # and the FP32 residual and writes e (12 B per interior cell); the FP64
    # pass reads U and b and writes U (24 B per cell).
    bpc = (12.0 if headname == "split" else 24.0) * spl[headname]
    bytes_per_launch = bpc * ncell
    avg_s = (head["jac_ms"] / 1e3) / max(head["jac_launches"], 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    per_form = {}
    for k, r in runs.items():
        launch_us = 1e3 * r["jac_ms"] / max(r["jac_launches"], 1)
        per_form[k] = dict(value=r["pushed"] / r["elapsed"], ms_per_step=1e3 * r["elapsed"] / args.steps,
                           sweeps_per_step=r["sweeps"], sweeps_per_launch=spl[k],
                           launch_kernel_us=launch_us, sweep_kernel_us=launch_us / spl[k])
    errors = None
    if "split" in runs and "fp64" in runs:
        a, b = runs["split"]["fields"], runs["fp64"]["fields"]
        errors = {"after_step": runs["split"]["captured_at"],
                  "phi_rel": _rel(a["U"], b["U"]), "E_rel": _rel(a["E"], b["E"]),
                  "N_rel": _rel(a["N"], b["N"]),
                  "phi_rel_inf": float(np.abs(a["U"] - b["U"]).max() / np.abs(b["U"]).max())}
    # the FP64 form runs the two-cells-per-lane pass when nx is even
    kname = (("k_mx_pair" if spl["split"] == 2 else "k_mx_sweep") if headname == "split"
             else ("k_jacobi_pass2" if cfg.nx % 2 == 0 else "k_jacobi_pass"))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if headname == "split" else "f64",
        "data": "synthetic (device Philox plasma, inflow and Haser production)",
        "config": {
            "workload": f"{cfg.nx}^3 cells dx=10 m, H+ and e- at 32 per cell, 400 km/s, 10 eV, "
                        f"B=(0,0,5 nT), x in/outflow, y/z periodic, dt=0.1/w_pe, Jacobi tol 1e-6, "
                        f"Haser 18 amu source Q={host.HASER_Q_USED:.2g}/s (stated "
                        f"{host.HASER_Q_STATED:.0e}/s scaled to ~2^18 ions/step), "
                        f"value = {headname} form",
            "grid": [cfg.nx, cfg.ny, cfg.nz],
            "forms": per_form,
            "split_over_fp64": (per_form["split"]["value"] / per_form["fp64"]["value"]
                                if len(per_form) == 2 else None),
            "split_error_vs_fp64": errors,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(kname),
            "bytes_per_launch": bytes_per_launch,
            "sweeps_per_launch": spl[headname],
            "avg_launch_us": avg_s * 1e6,
            "timed_launches": head["jac_launches"],
            "timed_every": 16,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
