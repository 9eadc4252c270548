"""Host side of the ensemble server (no device work): the job stream's
arrivals are a seeded Poisson process at the stated rate, its draws stay in
their ranges, and a job's HBM footprint counts its grids and particle slots."""
import numpy as np

from picard_amd import ensemble, server


def test_poisson_arrivals_and_draws():
    jobs = server.draw_jobs(rate=5.0, count=400, seed=3)
    t = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.2) < 0.03                     # mean gap 1 / rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.15        # exponential: cv = 1
    for j in jobs:
        assert 32 <= j.spec.nx <= 160 and 8 <= j.spec.ppc <= 32 and 200 <= j.steps <= 2000
    again = server.draw_jobs(rate=5.0, count=400, seed=3)
    assert [(j.arrival, j.spec.seed) for j in again] == [(j.arrival, j.spec.seed) for j in jobs]


def test_job_bytes():
    spec = ensemble.draw_ensemble(1, seed=9)[0]
    g, parts = ensemble.footprint(spec)
    assert server.job_bytes(spec) == g + parts * ensemble.PARTICLE_BYTES
    big = ensemble.SimSpec(nx=160, ny=160, nz=160, ppc=32, dx=10.0, density=5e6, kelvin=1e5,
                           drift=4e5, B=(0, 0, 5e-9), seed=1)
    # a 160^3 job at 32 per cell: ~2.6e8 particles, about 20 GB
    assert 15e9 < server.job_bytes(big) < 30e9


def test_offered_load_estimate():
    jobs = server.draw_jobs(rate=2.0, count=50, seed=4)
    est = server.offered_load(jobs, sweeps_per_n2=6.0)
    # the stated stream asks for far more than one GPU second per second
    assert est > 10.0


def test_lane_routing():
    lanes = server.two_lanes(80)
    assert [ls.max_edge for ls in lanes] == sorted(ls.max_edge for ls in lanes)
    jobs = server.draw_jobs(rate=2.0, count=200, seed=5)
    q = [server.lane_index(lanes, j) for j in jobs]
    for j, k in zip(jobs, q):
        assert (k == 0) == (max(j.spec.nx, j.spec.ny, j.spec.nz) <= 80)
    # both lanes get jobs from the stated stream (edges log-uniform in 32..160)
    assert 0.3 < np.mean(np.array(q) == 0) < 0.8
    # the lanes' arenas fit in the budget's share of HBM
    total = sum(ls.max_particles * ensemble.PARTICLE_BYTES for ls in lanes)
    assert total < 160e9
