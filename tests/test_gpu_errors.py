"""Failure statuses of the C ABI and the split form's cell sort.

Every failure here is detected on the host or counted by a kernel's own
bounds checks (no kernel is driven out of bounds): appending past
max_particles returns PICARD_ERR_CAPACITY, a particle whose CIC stencil
leaves the grid arrays returns PICARD_ERR_RANGE from ParticlesGridify and
AdvanceVelocity, a solve that runs to 2*maxit+1 sweeps returns
PICARD_WARN_NOCONV (the reference only prints, CalcEpotential.f90:67-69),
and picard_split_solver rejects refinement lengths outside 2..512."""
import ctypes as ct

import numpy as np
import pytest

from picard_amd import _lib, host
from picard_amd._lib import lib

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-errors] {msg}", flush=True)


def _cfg(n=8, flow=(True, False, False), gmin=None, gmax=None):
    half = 5.0 * n
    return host.Config(nx=n, ny=n, nz=n, flow=flow,
                       gmin=gmin or (-half,) * 3, gmax=gmax or (half,) * 3,
                       ppc=(0, 4, 0, 0, 0, 0, 0, 0))


def _particles(cfg, n, rng):
    rp = np.zeros((n, 8))
    for c in range(3):
        rp[:, c] = cfg.gmin[c] + (cfg.gmax[c] - cfg.gmin[c]) * rng.random(n)
    rp[:, 3:6] = rng.normal(0.0, 1e5, (n, 3))
    rp[:, 7] = np.arange(n)
    return rp, rng.integers(1, 3, n).astype(np.int32)


@pytest.mark.parametrize("precision", [_lib.PICARD_FP64, _lib.PICARD_FP32_SPLIT])
def test_append_past_capacity(precision):
    _say(f"capacity precision={precision}")
    cfg = _cfg()
    dom = host.Domain(cfg, max_particles=1000, precision=precision)
    rng = np.random.default_rng(1)
    dom.append(*_particles(cfg, 600, rng))
    rp, sp = _particles(cfg, 401, rng)
    rc = lib().picard_particles_append(dom.h, host._dp(rp), host._ip(sp), rp.shape[0])
    assert rc == _lib.PICARD_ERR_CAPACITY
    assert "capacity" in lib().picard_last_error().decode()
    assert dom.num_local == 600                     # nothing of the refused batch landed
    rp, sp = _particles(cfg, 400, rng)              # exactly to capacity is allowed
    dom.append(rp, sp)
    assert dom.num_local == 1000
    dom.close()


def test_particle_outside_grid_is_range_error():
    """FP64 form, no BCoutflow in between: a particle two cells past the
    x face (CIC stencil beyond the guard layer) is not deposited and not
    pushed; both calls report it."""
    _say("range")
    cfg = _cfg()
    d = host.derive(cfg)
    dom = host.Domain(cfg, max_particles=4096, d=d)
    rng = np.random.default_rng(2)
    rp, sp = _particles(cfg, 500, rng)
    rp[17, 0] = cfg.gmax[0] + 2.5 * d.dxyz[0]
    dom.append(rp, sp)
    rc = lib().picard_particles_gridify(dom.h)
    assert rc == _lib.PICARD_ERR_RANGE
    assert "1 particle(s)" in lib().picard_last_error().decode()
    rc = lib().picard_advance_velocity(dom.h, ct.c_double(d.dt))
    assert rc == _lib.PICARD_ERR_RANGE
    got, _ = dom.particles()
    assert np.array_equal(got[17], rp[17])          # left unchanged
    dom.close()


@pytest.mark.parametrize("precision", [_lib.PICARD_FP64, _lib.PICARD_FP32_SPLIT])
def test_solve_past_maxit_warns(precision):
    """maxit = 3 on a random charge with tolerance 1e-12: every stop test
    fails, the solve runs 2*maxit+1 = 7 sweeps and returns the warning."""
    _say(f"noconv precision={precision}")
    cfg = _cfg(10)
    dom = host.Domain(cfg, max_particles=1024, precision=precision)
    dom.calc_distance()
    rng = np.random.default_rng(3)
    dom.set_field(_lib.FIELD_C, rng.normal(size=dom.shape) * 1e-16)
    dom.bc_potentials()
    sw = ct.c_int64()
    rc = lib().picard_calc_epotential(dom.h, ct.c_double(1e-12), 3, ct.byref(sw))
    assert rc == _lib.PICARD_WARN_NOCONV
    assert sw.value == 7
    rc = lib().picard_calc_epotential(dom.h, ct.c_double(1e-1), 1000, ct.byref(sw))
    assert rc == _lib.PICARD_OK and sw.value < 2001
    dom.close()


@pytest.mark.parametrize("inner", [0, 1, 513, 4096])
def test_split_solver_rejects_untested_lengths(inner):
    cfg = _cfg()
    dom = host.Domain(cfg, max_particles=1024, precision=_lib.PICARD_FP32_SPLIT)
    assert lib().picard_split_solver(dom.h, inner) == _lib.PICARD_ERR_ARG
    assert lib().picard_split_solver(dom.h, 512) == _lib.PICARD_OK
    dom.close()


def _split_key(c, sp, nx, ny, S):
    return (c[:, 0] + (nx + 1) * (c[:, 1] + (ny + 1) * c[:, 2])) * S + (sp - 1)


@pytest.mark.parametrize("n", [10_001, 64 * 37, 5])
def test_split_sort_permutation(n):
    """The split form's counting sort (one atomic per distinct key and wave):
    shuffled particles of many keys per wave, deleted ones mixed in, n not a
    multiple of 64.  After BCoutflow the survivors are exactly the particles
    the deletion rule keeps (BCoutflow.f90:28-66), each once and unchanged,
    in non-decreasing (cell, species) key with the out-of-box keepers last;
    after ParticlesTransfer every key is inside the grid, still sorted.

    Unit cells from the origin and offsets k/1024 make the FP64 -> split
    conversion exact, so the host computes each particle's cell and offset."""
    nx, ny, nz = 9, 7, 5
    cfg = host.Config(nx=nx, ny=ny, nz=nz, flow=(True, False, False), gmin=(0.0, 0.0, 0.0),
                      gmax=(float(nx), float(ny), float(nz)), ppc=(0, 4, 0, 0, 0, 0, 0, 0))
    dims = np.array([nx, ny, nz])
    rng = np.random.default_rng(n)
    c = np.stack([rng.integers(-dims[d] - 2, 2 * dims[d] + 3, n) for d in range(3)], axis=1)
    inside = rng.random(n) < 0.7                      # most particles in the box
    for d in range(3):
        c[inside, d] = rng.integers(0, dims[d] + 1, int(inside.sum()))
    f = rng.integers(0, 1024, (n, 3)) / 1024.0
    sp = rng.integers(1, 3, n).astype(np.int32)
    rp = np.zeros((n, 8))
    rp[:, :3] = c + f - 0.5                           # xi = x + 1/2 = c + f
    rp[:, 3:6] = rng.normal(0.0, 1e3, (n, 3)).astype(np.float32)
    # the deletion rule on (c, f): below is xi < 1/2, above is xi > n + 1/2
    below = (c < 0) | ((c == 0) & (f < 0.5))
    above = (c > dims) | ((c == dims) & (f > 0.5))
    band_lo = ((c + dims) < 0) | (((c + dims) == 0) & (f < 0.5))
    band_hi = ((c - dims) > dims) | (((c - dims) == dims) & (f > 0.5))
    delete = (below[:, 0] | above[:, 0]) | np.any(band_lo | band_hi, axis=1)
    keep = ~delete
    _say(f"sort n={n} kept={int(keep.sum())}")
    dom = host.Domain(cfg, max_particles=n + 64, precision=_lib.PICARD_FP32_SPLIT)
    dom.append(rp, sp)
    num, removed = dom.bc_outflow()
    assert (num, removed) == (int(keep.sum()), int(delete.sum()))
    got, gsp = dom.particles()
    ids = got[:, 7].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.flatnonzero(keep))      # each survivor once
    assert np.array_equal(got[:, :6], rp[ids, :6]) and np.array_equal(gsp, sp[ids])
    gc = c[ids]
    in_grid = np.all((gc >= 0) & (gc <= dims), axis=1)
    key = np.where(in_grid, _split_key(gc, gsp, nx, ny, 2), (nx + 1) * (ny + 1) * (nz + 1) * 2)
    assert np.all(np.diff(key) >= 0)
    dom.particles_transfer()                                      # wrap y, z; sort again
    got2, gsp2 = dom.particles()
    assert dom.num_local == num
    ids2 = got2[:, 7].astype(np.int64)
    assert np.array_equal(np.sort(ids2), np.flatnonzero(keep))
    xi = got2[:, :3] + 0.5
    c2 = np.floor(xi).astype(np.int64)
    assert np.all((c2 >= 0) & (c2 <= dims))
    assert np.all(np.diff(_split_key(c2, gsp2, nx, ny, 2)) >= 0)
    dom.close()
