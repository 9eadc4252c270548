#!/usr/bin/env python3
"""Benchmark of the PIC time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 256]
                    [--fp64-steps K64] [--fp64-warmup W64]
                    [--form both|split|fp64] [--no-cpu-baseline] [--no-server]
    python bench.py --solve-only SWEEPS [--form split|fp64]   (profiling)
    python bench.py --copy-cal                                (profiling)
    python bench.py --stream-rates                            (stream reference)
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
FP32-split form gives `value` (K timed steps after W warmup steps); the FP64
form (reference order, bit-exact against the oracle) runs after it from the
same plasma, W64 warmup and K64 timed steps (an FP64 step costs about six
split steps, so its count is its own and the whole run stays within a few
minutes at K = 20).  The line reports both forms, their ratio over the same
step indices (whole step, and the solve phase alone per sweep), and the split
form's error against the FP64 form after the first step (phi, E, densities:
normwise relative).  The metric's second half, p50/p99 job completion latency,
comes from the ensemble server on its standard stream (40 Poisson jobs at 2
jobs/s, 32^3-160^3 grids, step counts scaled by 1/100), run after the forms
and reported as `server` in the same line (`--no-server` skips it; `--server`
runs it alone with its own options).

Prints one JSON line with the roofline of the dominant kernel (the split
form's FP32 Jacobi sweep, timed live over one launch in 16 by the kernel's
own clock, with HIP events around the same launches beside it), the rate a
plain stream with the same byte mix and array size reaches beside it, and a
CPU baseline of the oracle on a bounded sample."""
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


def run_form(cfg, precision, steps, warmup, capture_after=1, order_free=False,
             sweeps_per_launch=0):
    """W warmup + K timed steps of one form; returns timings (whole and per
    step, with the solve phase's wall time per step from HIP events) and the
    fields after step `capture_after` (read outside the timed region).
    sweeps_per_launch > 0 caps the split solve's sweeps per launch (else the
    library's default, 2)."""
    from picard_amd import _lib, host
    d = host.derive(cfg)
    ncell = cfg.nx * cfg.ny * cfg.nz
    per_cell = int(d.ppc[:cfg.species].sum())
    cap = int(1.12 * per_cell * ncell) + 64 * per_cell * (cfg.ny + 2) * (cfg.nz + 2) + (1 << 20)
    name = ("split" if precision == _lib.PICARD_FP32_SPLIT else
            "fp64_order_free" if order_free else "fp64")
    t0 = time.perf_counter()
    dom = host.Domain(cfg, max_particles=cap, d=d, precision=precision, order_free=order_free)
    if sweeps_per_launch and precision == _lib.PICARD_FP32_SPLIT:
        dom.split_sweeps_per_launch(sweeps_per_launch)
    dom.generate_initial(SEED)
    dom.calc_distance()
    dom.set_production(SEED)
    dom.synchronize()
    log(f"{name}: setup {time.perf_counter() - t0:.1f} s, {dom.num_local} particles, cap {cap}")
    fields = None
    it = 0
    warm_steps = []
    for _ in range(warmup):
        it += 1
        ts = time.perf_counter()
        dom.generate_inflow(SEED, it)
        st = dom.step()
        dom.synchronize()
        te = time.perf_counter()
        warm_steps.append(dict(step=it, seconds=te - ts, pushed=st.pushed, sweeps=st.sweeps))
        log(f"{name}: warmup step {it}: n={st.num_local} sweeps={st.sweeps} {te - ts:.2f} s")
        if it == capture_after:
            fields = {k: dom.field(w) for k, w in (("U", _lib.FIELD_U), ("E", _lib.FIELD_E),
                                                    ("N", _lib.FIELD_N))}
    dom.synchronize()
    dom.timing(True)
    pushed, sweeps, per_step = 0, [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        it += 1
        ts = time.perf_counter()
        dom.generate_inflow(SEED, it)
        st = dom.step()
        dom.synchronize()
        te = time.perf_counter()
        pushed += st.pushed
        sweeps.append(st.sweeps)
        per_step.append(dict(step=it, seconds=te - ts, pushed=st.pushed, sweeps=st.sweeps))
        log(f"{name}: timed step {it}: n={st.num_local} sweeps={st.sweeps} "
            f"{te - ts:.2f} s ({te - t0:.1f} s)")
    dom.synchronize()
    elapsed = time.perf_counter() - t0
    jac_ms, jac_launches = dom.timing_read()
    jac_sweeps = dom.timing_sweeps()
    clk_ms, clk_launches = dom.timing_kernel_clock()
    solve_ms, solves = dom.timing_solve()
    dom.timing(False)
    if fields is None:
        fields = {k: dom.field(w) for k, w in (("U", _lib.FIELD_U), ("E", _lib.FIELD_E),
                                                ("N", _lib.FIELD_N))}
    dom.close()
    log(f"{name}: {steps} steps in {elapsed:.2f} s, sweeps {sweeps}, solve phase "
        f"{solve_ms / max(solves, 1):.1f} ms per step, solve kernel "
        f"{1e3 * clk_ms / max(clk_launches, 1):.2f} us per launch by its clock "
        f"({1e3 * jac_ms / max(jac_launches, 1):.2f} by events), avg over {jac_launches}")
    return dict(pushed=pushed, elapsed=elapsed, sweeps=sweeps, jac_ms=jac_ms,
                jac_launches=jac_launches, jac_sweeps=jac_sweeps, clk_ms=clk_ms,
                clk_launches=clk_launches, solve_ms=solve_ms,
                solves=solves, per_step=per_step,
                warm_steps=warm_steps, fields=fields, captured_at=min(capture_after, it))


def _host_cores():
    """Logical CPUs the process may run on, and the machine's count."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return usable, os.cpu_count()


def cpu_baseline(budget_s=20.0, workload=None):
    """The oracle (C restatement of the reference order, FP64, 1 thread) on a
    bounded sample of the workload: whole steps of a 48^3 block of it (same
    cell size, plasma and 32 particles per cell per species, no source), each
    phase timed.  `workload` (cells, particles, sweeps per step of the full
    grid, from the GPU run) scales the sample's phase times to the whole
    workload: particle phases by the particle count, the solve by cells x
    sweeps, the face sum by face nodes x cells; that estimate is reported
    beside the measured rate and labelled as such."""
    from picard_amd import host
    from tests.pic_driver import OracleSim
    n = 48
    cfg = host.config2(n, species=2, Qn=(0.0,) * 8)
    d = host.derive(cfg)
    rng = np.random.default_rng(SEED)
    rp, sp = host.particles_initialize(cfg, d, rng)
    sim = OracleSim(cfg, d, rp, sp)
    sim.times = {}
    pushed, steps, sweeps = 0, 0, 0
    t0 = time.perf_counter()
    while steps < 1 or (time.perf_counter() - t0 < budget_s * 0.5 and steps < 20):
        sim.append(*host.bc_inflow(cfg, d, rng))
        st = sim.step()
        pushed += sim.n
        sweeps += st["sweeps"]
        steps += 1
        log(f"cpu baseline: step {steps} n={sim.n} sweeps={st['sweeps']} "
            f"({time.perf_counter() - t0:.1f} s)")
    el = time.perf_counter() - t0
    usable, total = _host_cores()
    out = dict(value=pushed / el, unit="particle-steps/s", cores=1, kind="port",
               host_cores=usable, host_cores_total=total,
               sample=f"{steps} whole step(s) of a 48^3 block of the workload (10 m cells, H+ and "
                      f"e- at 32 per cell, {sim.n} particles, x in/outflow, no Haser source; "
                      f"{sweeps // steps} Jacobi sweeps per step); oracle C restatement in "
                      f"reference order, FP64, 1 thread of the {usable} usable host cores, "
                      f"{el:.1f} s")
    if workload:
        t = {k: v / steps for k, v in sim.times.items()}
        cells_s = n ** 3
        face_s = 2 * n * n * cells_s
        part = sum(v for k, v in t.items() if k not in ("solve", "bcpot", "efield"))
        P = workload["particles"]
        est = (part * P / (pushed / steps) + t.get("efield", 0.0) * workload["cells"] / cells_s +
               t.get("solve", 0.0) * (workload["cells"] * workload["sweeps"]) /
               (cells_s * sweeps / steps) +
               t.get("bcpot", 0.0) * (2 * workload["nx"] ** 2 * workload["cells"]) / face_s)
        out["workload_estimate"] = dict(
            value=P / est, seconds_per_step=est,
            method="the sample's phase times per step scaled to the full workload: particle "
                   "phases x particles, the E field x cells, the solve x cells x sweeps (sweeps "
                   "per step from the "
                   "GPU run), the face sum x face nodes x cells; one thread; an estimate, "
                   "not a run")
    return out


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
    if max(doms["fp64"].num_local, doms["split"].num_local) > 300_000_000:
        # matching 1e9 particles by id on the host takes minutes: per-step
        # fields only at that size
        for dom in doms.values():
            dom.close()
        print(json.dumps({"mode": "error-report", "grid": [cfg.nx, cfg.ny, cfg.nz],
                          "steps": steps, "per_step": per_step,
                          "trajectories": "not matched (more than 3e8 particles)"}), flush=True)
        return
    log("error report: matching particles by id")
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


def ensemble_bench(count, steps, warmup, tune_sweeps=32, baseline=True, seed=20261210,
                   graphs=True):
    """The 64-simulation ensemble: grids log-uniform in 16^3-96^3, 4-32
    particles per cell per species, plasma drawn per simulation, stepped
    together in one set of launches per phase (FP32-split form), the solve
    tile shape from the autotuner; W warmup and K timed steps.  Baseline: the
    same simulations one at a time in the single-domain FP32-split form (each
    one's second step timed after its first), summed."""
    from picard_amd import _lib, ensemble, host
    specs = ensemble.draw_ensemble(count, seed=seed)
    t0 = time.perf_counter()
    ens = ensemble.Ensemble.sized_for(specs, steps_room=steps + warmup + 4, graphs=graphs)
    slots = [ens.add(s) for s in specs]
    log(f"ensemble: {count} simulations, {sum(s.cells for s in specs)} cells, "
        f"{sum(ens.status(q).num_local for q in slots)} particles, setup "
        f"{time.perf_counter() - t0:.1f} s")
    tile, times = ens.tune(tune_sweeps)
    log(f"ensemble: tuner per-sweep us {np.round(times, 2).tolist()} -> {tile.cells_per_thread} "
        f"cells/thread, {tile.blocks_per_cu} blocks/CU")
    for w in range(warmup):
        st = ens.step()
        log(f"ensemble: warmup step {w + 1}: {st.particles} particles, sweeps max {st.max_sweeps} "
            f"sum {st.sum_sweeps}, {st.step_ms:.0f} ms (solve {st.solve_ms:.0f})")
    pushed, cell_sweeps, solve_ms, per = 0, 0, 0.0, []
    t0 = time.perf_counter()
    for k in range(steps):
        st = ens.step()
        pushed += st.pushed
        cell_sweeps += st.cell_sweeps
        solve_ms += st.solve_ms
        per.append(dict(step_ms=st.step_ms, solve_ms=st.solve_ms, max_sweeps=st.max_sweeps,
                        sum_sweeps=st.sum_sweeps, chunks=st.solve_chunks, noconv=st.noconv))
        log(f"ensemble: timed step {warmup + k + 1}: {st.particles} particles, sweeps max "
            f"{st.max_sweeps} sum {st.sum_sweeps}, {st.step_ms:.0f} ms (solve {st.solve_ms:.0f})")
    el = time.perf_counter() - t0
    ens.close()
    out = dict(value=pushed / el, seconds=el, steps=steps, warmup=warmup, per_step=per,
               tile=dict(cells_per_thread=tile.cells_per_thread, blocks_per_cu=tile.blocks_per_cu),
               tuner_us_per_sweep=[float(v) for v in times],
               solve_fraction=solve_ms / 1e3 / el,
               solve_sweep_bytes_per_s=12.0 * cell_sweeps / (solve_ms / 1e3) if solve_ms else 0.0,
               sims=count, cells=int(sum(s.cells for s in specs)))
    if baseline:
        tot_ps, tot_s = 0, 0.0
        for q, s in enumerate(specs):
            hx = [0.5 * s.dx * n for n in (s.nx, s.ny, s.nz)]
            cfg = host.Config(nx=s.nx, ny=s.ny, nz=s.nz, species=2, flow=(True, False, False),
                              gmin=tuple(-h for h in hx), gmax=tuple(hx),
                              density=(0.0, s.density, 0, 0, 0, 0, 0, 0),
                              ppc=(0, s.ppc, 0, 0, 0, 0, 0, 0),
                              kelvin=(s.kelvin,) * 2 + (0.0,) * 6, v0=(s.drift, 0.0, 0.0),
                              B0=tuple(s.B))
            d = host.derive(cfg)
            dom = host.Domain(cfg, d=d, precision=_lib.PICARD_FP32_SPLIT)
            dom.generate_initial(s.seed)
            dom.calc_distance()
            dom.generate_inflow(s.seed, 1)
            dom.step()
            dom.synchronize()
            ts = time.perf_counter()
            dom.generate_inflow(s.seed, 2)
            st = dom.step()
            dom.synchronize()
            tot_s += time.perf_counter() - ts
            tot_ps += st.pushed
            dom.close()
            if q % 16 == 15:
                log(f"ensemble baseline: {q + 1} simulations one at a time, {tot_s:.2f} s")
        out["one_at_a_time"] = dict(value=tot_ps / tot_s, seconds=tot_s,
                                    how="each simulation alone in the single-domain FP32-split "
                                        "form, its step 2 timed after its step 1, summed")
        out["batched_over_one_at_a_time"] = out["value"] / out["one_at_a_time"]["value"]
    return out


def server_bench(rate, jobs, scale_steps, nmin, nmax, time_limit, seed=20261211,
                 lanes_split=80):
    """The ensemble server on a seeded Poisson stream of jobs (rate jobs/s),
    grids log-uniform in nmin^3-nmax^3, 8-32 particles per cell, 200-2000
    steps scaled by `scale_steps`, admitted by a 250 GB HBM budget; jobs up
    to lanes_split^3 in one lane and the rest in another (0: one lane)."""
    from picard_amd import server
    js = server.draw_jobs(rate, jobs, seed=seed, nmin=nmin, nmax=nmax,
                          steps=(max(1, int(200 * scale_steps)), max(1, int(2000 * scale_steps))))
    srv = server.Server(tune_sweeps=32,
                        lanes=server.two_lanes(lanes_split) if lanes_split else None)
    rep = srv.run(js, time_limit=time_limit, log=log)
    srv.close()
    done = rep.pop("done")
    rep["offered_gpu_s_per_s_estimate"] = server.offered_load(js, sweeps_per_n2=6.0)
    rep["jobs"] = jobs
    rep["rate_jobs_per_s"] = rate
    rep["latencies_s"] = sorted(round(j.latency, 3) for j in done)
    return rep


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


def stream_rates():
    """The streaming rates the solve kernels are held against: torch's copy
    (read 1, write 1) and add (read 2, write 1, the byte mix of an FP32 sweep)
    on float32 arrays of the 256^3 solve's size (67 MB each, all three fit
    the 256 MB last-level cache) and of 1 GiB each (they do not), timed with
    events over 20 launches after 3 warmups."""
    import torch
    out = {"mode": "stream-rates"}
    rates = {}
    for label, n in (("67MB", 1 << 24), ("1GiB", 1 << 28)):
        x = torch.rand(n, device="cuda")
        y = torch.rand(n, device="cuda")
        z = torch.empty_like(x)
        for op, nbytes, fn in (("copy", 8 * n, lambda: z.copy_(x)),
                               ("add", 12 * n, lambda: torch.add(x, y, out=z))):
            for _ in range(3):
                fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                fn()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / 20
            rates[f"{op}_{label}"] = {"us": us, "GBs": nbytes / us / 1e3}
            log(f"stream rates: {op} {label}: {us:.1f} us, {nbytes / us / 1e3:.0f} GB/s")
        del x, y, z
    out["stream_rates"] = rates
    print(json.dumps(out), flush=True)


def stream_reference(n):
    """The rate a plain stream with an FP32 sweep's byte mix reaches on this
    GPU: torch's add of two float32 arrays of n elements into a third (read
    8 B, write 4 B per element, as the solve's arrays: they fit the 256 MB
    last-level cache together), timed with events over 20 launches after 3
    warmups.  It runs in a child process: torch brings its own HIP runtime,
    which does not start in a process where the library's has."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-ref", str(n)],
                           capture_output=True, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError, OSError) as exc:
        log(f"stream reference: not measured ({exc!r})")
        return None


def stream_reference_child(n):
    """--stream-ref N: the measurement of stream_reference, in this process."""
    import torch
    x = torch.rand(n, device="cuda")
    y = torch.rand(n, device="cuda")
    z = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, y, out=z)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        torch.add(x, y, out=z)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / 20
    del x, y, z
    torch.cuda.empty_cache()
    print(json.dumps({"op": f"torch.add, float32, {n} elements (read 8 B, write 4 B each)",
                      "us": us, "GBs": 12.0 * n / us / 1e3}), flush=True)


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
    ap.add_argument("--steps", type=int, default=2,
                    help="timed steps of the FP32-split form (the value)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed steps of the FP64 form, after --fp64-warmup untimed ones (an "
                         "FP64 step at 256^3 takes ~23 s, a split one ~4 s)")
    ap.add_argument("--fp64-warmup", type=int, default=1)
    ap.add_argument("--no-order-free", action="store_true",
                    help="skip the order-free FP64 form (fixed-point deposit, FFT face sums)")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--form", choices=("both", "split", "fp64"), default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-server", action="store_true",
                    help="skip the job-server leg of the default run (the metric's p50/p99 job "
                         "latency)")
    ap.add_argument("--solve-only", type=int, default=0, metavar="SWEEPS",
                    help="profiling: one capped field solve of --form (split or fp64) only")
    ap.add_argument("--inner", type=int, default=0, metavar="M",
                    help="with --solve-only: FP32 sweeps per FP64 refinement of the split form")
    ap.add_argument("--error-steps", type=int, default=0, metavar="N",
                    help="error mode: N steps of both forms from the same plasma, split vs FP64 "
                         "errors per step and particle trajectories after the last")
    ap.add_argument("--sweeps-per-launch", type=int, default=0, choices=(0, 1, 2, 3),
                    help="at most this many FP32 sweeps per launch of the split solve "
                         "(0: the library's default, 2)")
    ap.add_argument("--stream-ref", type=int, default=0, metavar="N",
                    help="print the stream reference rate for N float32 elements and exit")
    ap.add_argument("--stream-rates", action="store_true",
                    help="time torch copy and add at the solve's array size and at 1 GiB")
    ap.add_argument("--copy-cal", action="store_true",
                    help="profiling: the 1 GiB copies the byte counters are scaled by")
    ap.add_argument("--ensemble", type=int, default=0, metavar="SIMS",
                    help="the batched ensemble of SIMS simulations (64 in the workload)")
    ap.add_argument("--no-baseline", action="store_true",
                    help="with --ensemble: skip the one-at-a-time baseline")
    ap.add_argument("--no-graphs", action="store_true",
                    help="with --ensemble: launch the phases directly instead of replaying graphs")
    ap.add_argument("--server", action="store_true", help="the ensemble server")
    ap.add_argument("--rate", type=float, default=2.0, help="--server: arrivals per second")
    ap.add_argument("--jobs", type=int, default=40, help="--server: jobs in the stream")
    ap.add_argument("--scale-steps", type=float, default=0.01,
                    help="--server: scale of the 200-2000 step counts")
    ap.add_argument("--nmax", type=int, default=160, help="--server: largest grid edge")
    ap.add_argument("--nmin", type=int, default=32, help="--server: smallest grid edge")
    ap.add_argument("--time-limit", type=float, default=240.0, help="--server: seconds")
    ap.add_argument("--lanes-split", type=int, default=80,
                    help="--server: jobs with grid edges up to this in one lane, the rest in "
                         "another, each stepped by its own host thread (0: one lane)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the domain "
              "decomposition it would use is described with the design); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from picard_amd import _lib, host
    if args.ensemble:
        r = ensemble_bench(args.ensemble, args.steps, args.warmup, baseline=not args.no_baseline,
                           graphs=not args.no_graphs)
        print(json.dumps({"metric": "particle-steps/s, batched FP32-split ensemble",
                          "value": r.pop("value"), "unit": "particle-steps/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f32+f64",
                          "data": "synthetic (per-simulation Philox plasma and inflow)",
                          "config": {"workload": f"{args.ensemble} simulations, grids log-uniform "
                                                 "16^3-96^3, 4-32 ppc per species, n 1-20 cm^-3, "
                                                 "T 1-50 eV, drift 300-700 km/s, |B| 1-10 nT "
                                                 "random direction, dx = Debye length, x "
                                                 "in/outflow, y/z periodic", **r}}), flush=True)
        return
    if args.server:
        r = server_bench(args.rate, args.jobs, args.scale_steps, args.nmin, args.nmax,
                         args.time_limit, lanes_split=args.lanes_split)
        print(json.dumps({"metric": "p50/p99 job completion latency and particle-steps/s, "
                                    "ensemble server",
                          "value": r["particle_steps_per_s"], "unit": "particle-steps/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32+f64",
                          "data": "synthetic (seeded Poisson arrivals, per-job Philox plasma)",
                          "config": {"workload": f"Poisson arrivals at {args.rate} jobs/s, "
                                                 f"{args.jobs} jobs, grids log-uniform "
                                                 f"{args.nmin}^3-{args.nmax}^3, 8-32 ppc, "
                                                 f"200-2000 steps x {args.scale_steps}, 250 GB "
                                                 "HBM admission budget, "
                                                 + (f"lanes: edges up to {args.lanes_split} and "
                                                    "above" if args.lanes_split else "one lane"),
                                     **r}}), flush=True)
        return
    cfg = host.config2(args.grid)
    if args.stream_ref:
        stream_reference_child(args.stream_ref)
        return
    if args.copy_cal:
        copy_calibration()
        return
    if args.stream_rates:
        stream_rates()
        return
    if args.solve_only:
        solve_only(cfg, "fp64" if args.form == "fp64" else "split", args.solve_only, args.inner)
        return
    if args.error_steps:
        error_report(cfg, args.error_steps)
        return
    runs = {}
    if args.form in ("both", "split"):
        runs["split"] = run_form(cfg, _lib.PICARD_FP32_SPLIT, args.steps, args.warmup,
                                 sweeps_per_launch=args.sweeps_per_launch)
    if args.form in ("both", "fp64"):
        runs["fp64"] = run_form(cfg, _lib.PICARD_FP64, args.fp64_steps if "split" in runs else
                                args.steps, args.fp64_warmup if "split" in runs else args.warmup)
    if args.form == "both" and not args.no_order_free:
        runs["fp64_order_free"] = run_form(cfg, _lib.PICARD_FP64, args.fp64_steps,
                                           args.fp64_warmup, order_free=True)
    head = runs.get("split") or runs["fp64"]
    headname = "split" if "split" in runs else "fp64"
    head_steps = len(head["per_step"])
    value = head["pushed"] / head["elapsed"]
    ncell = cfg.nx * cfg.ny * cfg.nz
    # Sweeps per launch of each form's solve kernel, over the timed launches:
    # the split form runs two FP32 sweeps per launch on rows of at most 256
    # cells or wider rows with x not periodic (three where asked for, on rows
    # of at most 256 cells with x not periodic); the FP64 form one pass per
    # launch.
    spl = {k: (r["jac_sweeps"] / r["jac_launches"] if r["jac_launches"] and r["jac_sweeps"]
               else 1.0) for k, r in runs.items()}
    # Bytes a launch of the dominant kernel must move: an FP32 sweep reads the
    # correction e and the FP32 residual and writes e (12 B per interior
    # cell), and a launch of several sweeps keeps the intermediate iterates on
    # chip, so it moves those 12 B once; the FP64 pass reads U and b and
    # writes U (24 B per cell).  The sweep-byte rate (12 B per cell and sweep)
    # is reported beside it.
    bytes_per_launch = (12.0 if headname == "split" else 24.0) * ncell
    sweep_bytes_per_launch = (12.0 if headname == "split" else 24.0) * spl[headname] * ncell
    # the dominant kernel's average launch duration: by the kernels' own clock
    # over the sampled launches (first block start to last block end, what
    # rocprof's kernel trace reports), the HIP dispatch-packet events around
    # the same launches reported beside it (they add the packet processing)
    avg_ev = (head["jac_ms"] / 1e3) / max(head["jac_launches"], 1)
    avg_s = ((head["clk_ms"] / 1e3) / head["clk_launches"]) if head["clk_launches"] else avg_ev
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    per_form = {}
    for k, r in runs.items():
        launch_ev_us = 1e3 * r["jac_ms"] / max(r["jac_launches"], 1)
        launch_us = 1e3 * r["clk_ms"] / r["clk_launches"] if r["clk_launches"] else launch_ev_us
        ns = len(r["per_step"])
        per_form[k] = dict(value=r["pushed"] / r["elapsed"], timed_steps=ns,
                           ms_per_step=1e3 * r["elapsed"] / ns,
                           sweeps_per_step=r["sweeps"], sweeps_per_launch=spl[k],
                           solve_ms_per_step=r["solve_ms"] / max(r["solves"], 1),
                           solve_us_per_sweep=1e3 * r["solve_ms"] / max(sum(r["sweeps"]), 1),
                           launch_kernel_us=launch_us, sweep_kernel_us=launch_us / spl[k],
                           launch_events_us=launch_ev_us,
                           per_step=r["per_step"], warmup_steps=r["warm_steps"])
    ratio = None
    if "split" in runs and "fp64" in runs:
        # like for like: the split form's rate over the same step indices the
        # FP64 form timed (both run the same plasma, inflow and source from
        # the same seeds; the sweep count per step changes from step to step,
        # 1.8e5 to 2.5e5 over the first 25 steps, so other steps are not
        # comparable), taken from its warmup steps where the indices fall
        # there; and the solve phase alone, per sweep
        def same_steps(other):
            idx = {p["step"] for p in runs[other]["per_step"]}
            sp_same = [p for p in runs["split"]["warm_steps"] + runs["split"]["per_step"]
                       if p["step"] in idx]
            fp = runs[other]["per_step"]
            if len(sp_same) != len(fp):
                return None
            return ((sum(p["pushed"] for p in sp_same) / sum(p["seconds"] for p in sp_same)) /
                    (sum(p["pushed"] for p in fp) / sum(p["seconds"] for p in fp)))

        ratio = {"whole_step_timed": per_form["split"]["value"] / per_form["fp64"]["value"],
                 "solve_per_sweep": (per_form["fp64"]["solve_us_per_sweep"] /
                                     max(per_form["split"]["solve_us_per_sweep"], 1e-30)),
                 "whole_step_same_steps": same_steps("fp64"),
                 "same_steps": sorted(p["step"] for p in runs["fp64"]["per_step"])}
        if "fp64_order_free" in runs:
            ratio["whole_step_same_steps_vs_order_free_fp64"] = same_steps("fp64_order_free")
    errors = None
    if "split" in runs and "fp64" in runs:
        a, b = runs["split"]["fields"], runs["fp64"]["fields"]
        errors = {"after_step": runs["split"]["captured_at"],
                  "phi_rel": _rel(a["U"], b["U"]), "E_rel": _rel(a["E"], b["E"]),
                  "N_rel": _rel(a["N"], b["N"]),
                  "phi_rel_inf": float(np.abs(a["U"] - b["U"]).max() / np.abs(b["U"]).max())}
    # the FP64 form runs the two-cells-per-lane pass when nx is even; the
    # split form's two-sweep launches share rows through LDS on rows of at
    # most 256 cells with x not periodic, except the launches marked for the
    # stop test (k_mx_pair; one in 16), and the timed launches sample both
    lds = (cfg.nx <= 256 and cfg.flow[0] and os.environ.get("PICARD_PAIR_KERNEL") != "regs")
    kname = (((("k_mx_triple_lds" if args.sweeps_per_launch == 3 else "k_mx_pair_lds") if lds
               else "k_mx_pair") if spl["split"] > 1.5 else "k_mx_sweep")
             if headname == "split"
             else ("k_jacobi_pass2" if cfg.nx % 2 == 0 else "k_jacobi_pass"))
    stream = stream_reference(ncell) if headname == "split" else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": head_steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["elapsed"] / head_steps,
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
            "split_over_fp64": ratio,
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
            "sweep_bytes_per_launch": sweep_bytes_per_launch,
            "sweep_bytes_GBs": sweep_bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0,
            "avg_launch_us": avg_s * 1e6,
            "duration_from": "kernel clock (100 MHz real-time counter, first block start to "
                             "last block end) of one launch in 16",
            "avg_launch_us_events": avg_ev * 1e6,
            "frac_events": (bytes_per_launch / avg_ev / 1e9 / HBM_PEAK_GBS) if avg_ev > 0 else None,
            "timed_launches": head["jac_launches"],
            "timed_every": 16,
            "launch_mix": {"k_mx_pair_lds": "k_mx_pair_lds, with k_mx_pair on the launches "
                                            "that apply the stop test",
                           "k_mx_triple_lds": "k_mx_triple_lds (three sweeps), with k_mx_pair "
                                              "(two) on the launches that apply the stop test "
                                              "and k_mx_pair_lds (two) where a refinement "
                                              "step's end calls for it"}.get(kname, kname),
            "stream_reference": stream,
            "frac_of_stream": (achieved / stream["GBs"]) if stream else None,
        },
        "cpu_baseline": None,
    }
    if not args.no_server and args.form == "both":
        # the metric's second half: the ensemble server's p50/p99 job
        # completion latency on its standard stream (40 Poisson jobs at 2
        # jobs/s, 32^3-160^3 grids, step counts scaled by 1/100), run after the
        # single-domain forms have released their device memory
        try:
            r = server_bench(2.0, 40, 0.01, 32, 160, 240.0, lanes_split=80)
            out["server"] = {
                "workload": "Poisson arrivals at 2 jobs/s, 40 jobs, grids log-uniform "
                            "32^3-160^3, 8-32 ppc, 200-2000 steps x 0.01, 250 GB HBM admission "
                            "budget, two lanes (edges up to 80 and above)",
                **{k: r[k] for k in ("p50_latency_s", "p99_latency_s", "particle_steps_per_s",
                                     "wall_s", "jobs_done", "p50_wait_s", "p99_wait_s",
                                     "peak_jobs", "peak_bytes")},
                "lanes": r.get("lanes"),
            }
        except Exception as exc:          # the single-domain line still prints
            log(f"server leg failed: {exc!r}")
            out["server"] = {"error": repr(exc)}
    if not args.no_cpu_baseline:
        P = head["pushed"] / head_steps
        out["cpu_baseline"] = cpu_baseline(workload=dict(
            particles=P, cells=ncell, nx=cfg.nx,
            sweeps=float(np.mean(head["sweeps"]))))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
