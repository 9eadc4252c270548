"""Ensemble server: simulation jobs arriving as a Poisson process, admitted by
an HBM budget into a running ensemble as slots and memory free up, all
running jobs of an ensemble advanced together one step per iteration
(continuous batching: a job that arrives while others are on the GPU joins
the next step's launches), and each job retired when it has run its steps.
Reports completion latency (finish - arrival) percentiles beside total
particle-steps/s.

Lanes: jobs can be routed by grid size to separate ensembles ("lanes"), each
served by its own host thread on its own HIP stream, so that the kernels of
all lanes share the GPU concurrently.  An ensemble step lasts as long as its
largest simulation's solve, so with one lane the small jobs wait on the
largest; in a lane of their own they step at their own pace.  Every job's
result is the same in any lane and any mix (the ensemble's determinism).

No reference counterpart: the reference runs one simulation per process
(picard.f90:12-684)."""
from __future__ import annotations

import math
import threading
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


@dataclass
class LaneSpec:
    """One ensemble of the server: the jobs whose largest grid edge is at
    most max_edge (and above the previous lane's), and its capacities."""
    max_edge: int
    max_sims: int
    max_cells: int
    max_particles: int


def two_lanes(split_edge: int = 80) -> list[LaneSpec]:
    """Jobs up to split_edge^3 and the rest, in about 150 GB of arenas."""
    return [LaneSpec(split_edge, 48, 40_000_000, 400_000_000),
            LaneSpec(1 << 30, 24, 120_000_000, 1_100_000_000)]


def lane_index(lanes: list[LaneSpec], job: Job) -> int:
    """The lane of `job` among lanes sorted by max_edge: the first whose
    max_edge holds the job's largest grid edge (the last lane otherwise)."""
    edge = max(job.spec.nx, job.spec.ny, job.spec.nz)
    for q, ls in enumerate(lanes):
        if edge <= ls.max_edge:
            return q
    return len(lanes) - 1


class _Lane:
    def __init__(self, spec: LaneSpec, device: int, graphs: bool):
        self.spec = spec
        self.ens = ensemble.Ensemble(max_sims=spec.max_sims, max_cells=spec.max_cells,
                                     max_particles=spec.max_particles, device=device,
                                     graphs=graphs)
        self.tuned = False


class Server:
    """Continuous-batching server over one ensemble per lane (one lane unless
    `lanes` says otherwise)."""

    def __init__(self, hbm_budget: float = 250 * GB, max_sims: int = 64,
                 max_cells: int = 160_000_000, max_particles: int = 1_500_000_000,
                 device: int = 0, graphs: bool = True, tune_sweeps: int = 0,
                 lanes: list[LaneSpec] | None = None):
        self.budget = hbm_budget
        if lanes is None:
            lanes = [LaneSpec(1 << 30, max_sims, max_cells, max_particles)]
        self.lanes = [_Lane(ls, device, graphs) for ls in
                      sorted(lanes, key=lambda ls: ls.max_edge)]
        self.tune_sweeps = tune_sweeps
        self._lock = threading.Lock()

    def close(self):
        for ln in self.lanes:
            ln.ens.close()

    def lane_of(self, job: Job) -> int:
        return lane_index([ln.spec for ln in self.lanes], job)

    def run(self, jobs: list[Job], keep_results: bool = False, time_limit: float | None = None,
            log=None) -> dict:
        """Serve `jobs` (arrival times relative to the start) until all have
        finished or `time_limit` seconds have passed; one host thread per
        lane."""
        for j in jobs:
            j.bytes = job_bytes(j.spec)
            spec = self.lanes[self.lane_of(j)].spec
            _, parts = ensemble.footprint(j.spec)
            if j.spec.cells > spec.max_cells or parts > spec.max_particles or j.bytes > self.budget:
                raise RuntimeError(f"job {j.jid} does not fit an empty server")
        queues = [deque() for _ in self.lanes]
        for j in sorted(jobs, key=lambda j: j.arrival):
            queues[self.lane_of(j)].append(j)
        sh = dict(held=0, running=0, steps=0, total_ps=0, busy=0.0, peak_jobs=0, peak_bytes=0)
        done: list[Job] = []
        left = {q: 0 for q in range(len(self.lanes))}
        errors: list[BaseException] = []
        t0 = time.perf_counter()

        def serve(q: int):
            ln, pending = self.lanes[q], queues[q]
            ens, cap = ln.ens, ln.spec
            running: dict[int, Job] = {}
            cells = particles = 0
            while pending or running:
                now = time.perf_counter() - t0
                if time_limit is not None and now > time_limit:
                    break
                # admit, in arrival order, every job of this lane that has
                # arrived and fits the lane and the shared budget
                while pending and pending[0].arrival <= now:
                    j = pending[0]
                    _, parts = ensemble.footprint(j.spec)
                    with self._lock:
                        fits = (len(running) < cap.max_sims and
                                cells + j.spec.cells <= cap.max_cells and
                                particles + parts <= cap.max_particles and
                                sh["held"] + j.bytes <= self.budget)
                        if fits:
                            sh["held"] += j.bytes
                            sh["running"] += 1
                            sh["peak_jobs"] = max(sh["peak_jobs"], sh["running"])
                            sh["peak_bytes"] = max(sh["peak_bytes"], sh["held"])
                    if not fits:
                        break
                    pending.popleft()
                    j.slot = ens.add(j.spec)
                    j.admitted = time.perf_counter() - t0
                    running[j.slot] = j
                    cells += j.spec.cells
                    particles += parts
                if not running:
                    # idle until this lane's next arrival, or until another
                    # lane frees budget
                    wait = pending[0].arrival - now if pending else 0.0
                    time.sleep(min(max(wait, 0.001), 0.05))
                    continue
                if self.tune_sweeps and not ln.tuned:
                    ens.tune(self.tune_sweeps)
                    ln.tuned = True
                ts = time.perf_counter()
                st = ens.step()
                dt = time.perf_counter() - ts
                finished = []
                for slot, j in running.items():
                    s = ens.status(slot)
                    j.done_steps = s.steps
                    j.particle_steps += s.pushed
                    if s.steps >= j.steps:
                        finished.append(slot)
                tf = time.perf_counter() - t0
                retired = []
                for slot in finished:
                    j = running.pop(slot)
                    j.finished = tf
                    if keep_results:
                        j.result = dict(U=ens.field(slot, _lib.FIELD_U),
                                        N=ens.field(slot, _lib.FIELD_N),
                                        particles=ens.particles(slot))
                    ens.remove(slot)
                    cells -= j.spec.cells
                    particles -= ensemble.footprint(j.spec)[1]
                    retired.append(j)
                with self._lock:
                    sh["busy"] += dt
                    sh["steps"] += 1
                    sh["total_ps"] += st.pushed
                    sh["held"] -= sum(j.bytes for j in retired)
                    sh["running"] -= len(retired)
                    done.extend(retired)
                    nsteps = sh["steps"]
                if log and nsteps % 10 == 0:
                    log(f"server: t={tf:.1f} s step {nsteps}: lane {q} running {len(running)}, "
                        f"pending {len(pending)}, done {len(done)}")
            left[q] = len(running)

        def guarded(q: int):
            try:
                serve(q)
            except BaseException as exc:        # reported by run()
                errors.append(exc)

        threads = [threading.Thread(target=guarded, args=(q,), daemon=True)
                   for q in range(len(self.lanes))]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        wall = time.perf_counter() - t0
        lat = np.array([j.latency for j in done]) if done else np.zeros(0)
        wait = np.array([j.admitted - j.arrival for j in done]) if done else np.zeros(0)
        pend = sum(len(qq) for qq in queues)
        per_lane = []
        for q, ln in enumerate(self.lanes):
            lq = np.array([j.latency for j in done if self.lane_of(j) == q])
            per_lane.append(dict(max_edge=ln.spec.max_edge, jobs_done=int(len(lq)),
                                 p50_latency_s=float(np.percentile(lq, 50)) if len(lq) else None,
                                 p99_latency_s=float(np.percentile(lq, 99)) if len(lq) else None))
        return dict(
            jobs_done=len(done), jobs_running=sum(left.values()), jobs_pending=pend,
            wall_s=wall, busy_s=sh["busy"], steps=sh["steps"],
            particle_steps=sh["total_ps"],
            particle_steps_per_s=sh["total_ps"] / wall if wall > 0 else 0.0,
            p50_latency_s=float(np.percentile(lat, 50)) if len(lat) else None,
            p99_latency_s=float(np.percentile(lat, 99)) if len(lat) else None,
            p50_wait_s=float(np.percentile(wait, 50)) if len(wait) else None,
            p99_wait_s=float(np.percentile(wait, 99)) if len(wait) else None,
            peak_jobs=sh["peak_jobs"], peak_bytes=sh["peak_bytes"], lanes=per_lane, done=done)


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
