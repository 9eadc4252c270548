"""Ensemble server: simulation jobs arriving as a Poisson process, admitted by
an HBM budget into the running ensemble as slots and memory free up, all
running jobs advanced together one step per iteration (continuous batching:
a job that arrives while others are on the GPU joins the next step's
launches), and each job retired when it has run its steps.  Reports
completion latency (finish - arrival) percentiles beside total
particle-steps/s.

No reference counterpart: the reference runs one simulation per process
(picard.f90:12-684)."""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib, ensemble

GB = 1e9


@dataclass
class Job:
    jid: int
    arrival: float                 # seconds after the server starts
    spec: ensemble.SimSpec
    steps: int
    admitted: float | None = None
    finished: float | None = None
    slot: int = -1
    done_steps: int = 0
    particle_steps: int = 0
    bytes: int = 0
    result: dict | None = None     # final fields and particles, when kept

    @property
    def latency(self) -> float | None:
        return None if self.finished is None else self.finished - self.arrival


def draw_jobs(rate: float, count: int, seed: int = 20261211, nmin: int = 32, nmax: int = 160,
              ppc=(8, 32), steps=(200, 2000)) -> list[Job]:
    """`count` jobs of a Poisson process at `rate` jobs/s: grids log-uniform in
    [nmin, nmax]^3, ppc uniform in 8..32, step counts uniform in 200..2000,
    plasma parameters as the 64-simulation ensemble draws them."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = ensemble.draw_ensemble(count, seed=seed + 1, nmin=nmin, nmax=nmax, ppc=ppc)
    jobs = []
    for q in range(count):
        t += rng.exponential(1.0 / rate)
        jobs.append(Job(jid=q, arrival=t, spec=specs[q],
                        steps=int(rng.integers(steps[0], steps[1] + 1))))
    return jobs


def job_bytes(spec: ensemble.SimSpec) -> int:
    """HBM a job holds: its grids and its particle slots (initial load plus one
    step's inflow, both particle arenas and the sort key)."""
    g, parts = ensemble.footprint(spec)
    return g + parts * ensemble.PARTICLE_BYTES


class Server:
    """Continuous-batching server over one ensemble."""

    def __init__(self, hbm_budget: float = 250 * GB, max_sims: int = 64,
                 max_cells: int = 160_000_000, max_particles: int = 1_500_000_000,
                 device: int = 0, graphs: bool = True, tune_sweeps: int = 0):
        self.budget = hbm_budget
        self.ens = ensemble.Ensemble(max_sims=max_sims, max_cells=max_cells,
                                     max_particles=max_particles, device=device, graphs=graphs)
        self.max_sims = max_sims
        self.max_cells = max_cells
        self.max_particles = max_particles
        self.tune_sweeps = tune_sweeps
        self.tuned = False

    def close(self):
        self.ens.close()

    def _fits(self, job: Job, running: dict, cells: int, particles: int, held: int) -> bool:
        _, parts = ensemble.footprint(job.spec)
        return (len(running) < self.max_sims and held + job.bytes <= self.budget and
                cells + job.spec.cells <= self.max_cells and
                particles + parts <= self.max_particles)

    def run(self, jobs: list[Job], keep_results: bool = False, time_limit: float | None = None,
            log=None) -> dict:
        """Serve `jobs` (arrival times relative to the start) until all have
        finished or `time_limit` seconds have passed."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        for j in pending:
            j.bytes = job_bytes(j.spec)
        running: dict[int, Job] = {}
        done: list[Job] = []
        t0 = time.perf_counter()
        steps = 0
        total_ps = 0
        busy = 0.0
        held = cells = particles = 0
        peak = dict(jobs=0, bytes=0)
        while pending or running:
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            # admit, in arrival order, every job that has arrived and fits
            while pending and pending[0].arrival <= now:
                j = pending[0]
                if not self._fits(j, running, cells, particles, held):
                    if not running:
                        raise RuntimeError(f"job {j.jid} does not fit an empty server")
                    break
                pending.popleft()
                j.slot = self.ens.add(j.spec)
                j.admitted = time.perf_counter() - t0
                running[j.slot] = j
                held += j.bytes
                cells += j.spec.cells
                particles += ensemble.footprint(j.spec)[1]
            if not running:
                # idle until the next arrival
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.5)))
                continue
            if self.tune_sweeps and not self.tuned:
                self.ens.tune(self.tune_sweeps)
                self.tuned = True
            ts = time.perf_counter()
            st = self.ens.step()
            busy += time.perf_counter() - ts
            steps += 1
            total_ps += st.pushed
            peak["jobs"] = max(peak["jobs"], len(running))
            peak["bytes"] = max(peak["bytes"], held)
            finished = []
            for slot, j in running.items():
                s = self.ens.status(slot)
                j.done_steps = s.steps
                j.particle_steps += s.pushed
                if s.steps >= j.steps:
                    finished.append(slot)
            tf = time.perf_counter() - t0
            for slot in finished:
                j = running.pop(slot)
                j.finished = tf
                if keep_results:
                    j.result = dict(U=self.ens.field(slot, _lib.FIELD_U),
                                    N=self.ens.field(slot, _lib.FIELD_N),
                                    particles=self.ens.particles(slot))
                self.ens.remove(slot)
                held -= j.bytes
                cells -= j.spec.cells
                particles -= ensemble.footprint(j.spec)[1]
                done.append(j)
            if log and steps % 10 == 0:
                log(f"server: t={tf:.1f} s step {steps}: running {len(running)}, done {len(done)}, "
                    f"pending {len(pending)}")
        wall = time.perf_counter() - t0
        lat = np.array([j.latency for j in done]) if done else np.zeros(0)
        wait = np.array([j.admitted - j.arrival for j in done]) if done else np.zeros(0)
        return dict(
            jobs_done=len(done), jobs_running=len(running), jobs_pending=len(pending),
            wall_s=wall, busy_s=busy, steps=steps,
            particle_steps=total_ps, particle_steps_per_s=total_ps / wall if wall > 0 else 0.0,
            p50_latency_s=float(np.percentile(lat, 50)) if len(lat) else None,
            p99_latency_s=float(np.percentile(lat, 99)) if len(lat) else None,
            p50_wait_s=float(np.percentile(wait, 50)) if len(wait) else None,
            p99_wait_s=float(np.percentile(wait, 99)) if len(wait) else None,
            peak_jobs=peak["jobs"], peak_bytes=peak["bytes"], done=done)


def offered_load(jobs: list[Job], sweeps_per_n2: float, sweep_bytes_per_cell: float = 12.0,
                 hbm_bytes_per_s: float = 5.0e12) -> float:
    """Estimate of the GPU seconds one second of arrivals asks for: per job,
    steps x (Jacobi sweeps per step, taken as sweeps_per_n2 * N^2, times the
    FP32 sweep bytes per cell) at a sustained HBM rate.  The solve dominates
    a step; particles and fields are left out, so this is a lower bound."""
    if not jobs:
        return 0.0
    span = max(j.arrival for j in jobs)
    work = 0.0
    for j in jobs:
        n = j.spec.nx
        work += j.steps * sweeps_per_n2 * n * n * j.spec.cells * sweep_bytes_per_cell
    return work / hbm_bytes_per_s / max(span, 1e-9)
