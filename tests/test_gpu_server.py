"""The ensemble server on the GPU: a small seeded Poisson stream of jobs is
admitted, batched and retired; every job finishes with its step count, the
latency report is complete, the HBM budget holds back jobs that do not fit,
and each job's final state is the same as the same simulation run alone,
also with jobs split by grid size over two lanes stepped concurrently."""
import numpy as np
import pytest

from picard_amd import _lib, ensemble, server

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-server] {msg}", flush=True)


def test_server_stream_and_results_alone():
    jobs = server.draw_jobs(rate=20.0, count=6, seed=77, nmin=12, nmax=20, ppc=(4, 6),
                            steps=(2, 4))
    srv = server.Server(max_sims=8, max_cells=60_000, max_particles=4_000_000)
    rep = srv.run(jobs, keep_results=True, time_limit=120.0, log=_say)
    srv.close()
    done = rep["done"]
    _say(f"server: {rep['jobs_done']} done in {rep['wall_s']:.2f} s over {rep['steps']} steps, "
         f"p50 {rep['p50_latency_s']:.3f} s p99 {rep['p99_latency_s']:.3f} s, peak "
         f"{rep['peak_jobs']} jobs")
    assert rep["jobs_done"] == 6 and rep["jobs_pending"] == 0 and rep["jobs_running"] == 0
    assert rep["peak_jobs"] >= 2                 # jobs overlapped on the device
    for j in done:
        assert j.done_steps == j.steps and j.latency > 0 and j.admitted >= j.arrival
    # each job alone, for its step count: the same fields and particles
    for j in done[:3]:
        ens = ensemble.Ensemble.sized_for([j.spec], steps_room=j.steps + 2)
        slot = ens.add(j.spec)
        for _ in range(j.steps):
            ens.step()
        U = ens.field(slot, _lib.FIELD_U)
        N = ens.field(slot, _lib.FIELD_N)
        rp, sp = ens.particles(slot)
        ens.close()
        assert np.array_equal(U, j.result["U"]) and np.array_equal(N, j.result["N"])
        assert np.array_equal(rp, j.result["particles"][0])
        assert np.array_equal(sp, j.result["particles"][1])


def test_server_budget_holds_jobs_back():
    """A budget that fits one job at a time: jobs run one after another, and
    each waits for the previous one's retirement."""
    jobs = server.draw_jobs(rate=50.0, count=3, seed=78, nmin=12, nmax=14, ppc=(4, 4),
                            steps=(2, 2))
    biggest = max(server.job_bytes(j.spec) for j in jobs)
    srv = server.Server(hbm_budget=1.5 * biggest, max_sims=8, max_cells=20_000,
                        max_particles=2_000_000)
    rep = srv.run(jobs, time_limit=60.0)
    srv.close()
    done = sorted(rep["done"], key=lambda j: j.admitted)
    assert rep["jobs_done"] == 3 and rep["peak_jobs"] == 1
    for a, b in zip(done, done[1:]):
        assert b.admitted >= a.finished


def _alone(j):
    ens = ensemble.Ensemble.sized_for([j.spec], steps_room=j.steps + 2)
    slot = ens.add(j.spec)
    for _ in range(j.steps):
        ens.step()
    out = (ens.field(slot, _lib.FIELD_U), ens.field(slot, _lib.FIELD_N), ens.particles(slot))
    ens.close()
    return out


def test_server_two_lanes_results_alone():
    """Jobs up to 17^3 in one lane, larger ones in another, each lane stepped
    by its own host thread on its own stream: every job ends with its step
    count and the state it has alone."""
    jobs = server.draw_jobs(rate=30.0, count=8, seed=79, nmin=12, nmax=26, ppc=(4, 6),
                            steps=(2, 4))
    lanes = [server.LaneSpec(17, 8, 40_000, 2_000_000), server.LaneSpec(1 << 30, 8, 100_000,
                                                                        4_000_000)]
    srv = server.Server(lanes=lanes)
    used = {srv.lane_of(j) for j in jobs}
    rep = srv.run(jobs, keep_results=True, time_limit=120.0, log=_say)
    srv.close()
    _say(f"two lanes: {rep['jobs_done']} done in {rep['wall_s']:.2f} s, lanes {rep['lanes']}")
    assert used == {0, 1}
    assert rep["jobs_done"] == 8 and rep["jobs_pending"] == 0 and rep["jobs_running"] == 0
    assert [ln["jobs_done"] for ln in rep["lanes"]] == [
        sum(1 for j in jobs if srv.lane_of(j) == q) for q in range(2)]
    small = [j for j in rep["done"] if srv.lane_of(j) == 0][:1]
    large = [j for j in rep["done"] if srv.lane_of(j) == 1][:1]
    for j in small + large:
        assert j.done_steps == j.steps
        U, N, (rp, sp) = _alone(j)
        assert np.array_equal(U, j.result["U"]) and np.array_equal(N, j.result["N"])
        assert np.array_equal(rp, j.result["particles"][0])
        assert np.array_equal(sp, j.result["particles"][1])
