"""CPU tests of the product's host side: the C-ABI library loads and exports
every symbol include/picard_hip.h declares (no compute calls without a GPU),
derived constants, the host sources, and the parallel BCoutflow rule that the
GPU compaction implements."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import oracle
from picard_amd import _lib, host
from tests.helpers import random_particles

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols(header="picard_hip.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(picard_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    declared = _declared_symbols()
    assert len(declared) >= 20
    L = ct.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


def test_every_header_is_covered():
    """Every header under include/ is one of the two bound ones, and every
    symbol each declares is exported and bound."""
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["picard_ensemble.h", "picard_hip.h"]
    declared = _declared_symbols("picard_ensemble.h")
    assert len(declared) >= 12
    L = ct.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.ENSEMBLE_EXPORTS) == declared
    _lib.lib()                                          # every binding resolves


def test_library_missing_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libpicard_hip.so")
    with pytest.raises(RuntimeError):
        _lib.lib()


def test_derived_constants_config1():
    d = host.derive(host.config1())
    assert np.allclose(d.dxyz, 10.0)
    assert d.ppc[0] == 8 and d.ppc[1] == 8 and d.ipp[1] == 8 and d.epp[1] == 1
    assert d.density[0] == d.density[1] == 5.0e6
    assert abs(d.omega_pe - 1.2615e5) / 1.2615e5 < 1e-3
    assert abs(d.dt - 7.927e-7) / 7.927e-7 < 1e-3
    assert abs(abs(d.tB[0, 3]) - 3.49e-4) / 3.49e-4 < 2e-3
    assert d.n2p[0] == d.n2p[1] == 1000.0 * 5e6 / 8
    assert np.allclose(d.ve[:3], [4e5, 0, 0])
    # E0 = -(ve x B0), the motional field (picard.f90:424-426)
    assert np.allclose(d.E0[:3], -np.cross(d.ve[:3], d.B0[:3]))
    # Boris: |sB| = 2|tB|/(1+|tB|^2) gives a rotation by 2 atan|tB|
    t = d.tB[0, 3]
    assert t < 0.0                             # electrons: eta < 0
    assert abs(d.sB[0, 3] - 2 * t / (1 + t * t)) <= 1e-18


def test_sources_config1_shapes_and_statistics():
    cfg = host.config1()
    d = host.derive(cfg)
    rng = np.random.default_rng(0)
    rp, sp = host.particles_initialize(cfg, d, rng)
    assert rp.shape == (32 ** 3 * 16, 8)
    assert np.array_equal(rp[0::2, :3], rp[1::2, :3])          # co-located ion/electron
    assert np.all(sp[0::2] == 2) and np.all(sp[1::2] == 1)
    for c in range(3):
        assert rp[:, c].min() >= cfg.gmin[c] and rp[:, c].max() <= cfg.gmax[c]
    # drifting Maxwellian: mean 400 km/s along x, thermal width sqrt(kT/m)
    vi = rp[sp == 2, 3:6]
    assert abs(vi[:, 0].mean() - 4e5) < 200.0
    sig = np.sqrt(host.KB * 116045.0 / host.MP)
    assert abs(vi[:, 1].std() / sig - 1.0) < 0.01
    ri, si = host.bc_inflow(cfg, d, rng)
    assert ri.shape[0] == 2 * 32 * 32 * 16
    x = ri[:, 0]
    assert np.all((x >= cfg.gmax[0]) & (x <= cfg.gmax[0] + 10.0) |
                  (x >= cfg.gmin[0] - 10.0) & (x <= cfg.gmin[0]))


def _parallel_outflow(del_flags):
    """The compaction the GPU kernels implement (particles.hip), restated with
    numpy.  Slots >= n' are visited in descending order as the sequential loop
    does: a kept slot enqueues its particle; a deleted slot with a kept slot
    above it dequeues the oldest pending particle and enqueues it again (it
    moved into the hole and will move once more).  The enqueue stream e[t]
    is either a particle or a pointer to an earlier entry, resolved by pointer
    jumping; deleted slot < n' of descending rank r receives e[R + r].
    Returns the final slot -> original index map."""
    f = del_flags.astype(np.int64)
    n = f.size
    scan = np.concatenate([[0], np.cumsum(f)])
    dtot = scan[n]
    nn = n - dtot
    m = scan[nn]
    q = np.arange(nn, n)
    kept_above = (n - 1 - q) - (dtot - scan[q + 1])
    rot = (f[q] == 1) & (kept_above > 0)
    scan_r = np.concatenate([[0], np.cumsum(rot)])
    rtot = scan_r[-1]
    rot_above = rtot - scan_r[1:]
    T = m + rtot
    val = np.full(T, -1, np.int64)
    ptr = np.full(T, -1, np.int64)
    for a, qq in enumerate(q):
        t = kept_above[a] + rot_above[a]
        if not f[qq]:
            val[t] = qq
        elif rot[a]:
            ptr[t] = rot_above[a]
    while np.any(val < 0):                         # pointer jumping
        nv, nptr = val.copy(), ptr.copy()
        todo = val < 0
        tgt = ptr[todo]
        done_t = val[tgt] >= 0
        idx = np.nonzero(todo)[0]
        nv[idx[done_t]] = val[tgt[done_t]]
        nptr[idx[~done_t]] = ptr[tgt[~done_t]]
        val, ptr = nv, nptr
    out = np.arange(n)
    for i in range(nn):
        if f[i]:
            r = m - 1 - scan[i]
            out[i] = val[rtot + r]
    return out[:nn]


@pytest.mark.parametrize("seed", range(40))
def test_parallel_outflow_rule_matches_sequential(orc, seed):
    rng = np.random.default_rng(seed)
    g = oracle.Grid(5, 5, 5, 2, flow=(True, bool(seed & 1), bool(seed & 2)))
    n = int(rng.integers(1, 400))
    spread = [0.0, 0.5, 3.0][seed % 3]
    rp, sp = random_particles(rng, g, n, spread=spread)
    rp[:, 7] = np.arange(n)                    # identity in slot 8
    flags = np.zeros(n, np.int64)
    for i in range(n):
        x = rp[i]
        dele = False
        for c in range(3):
            if g.flow[c] and (x[c] < g.gmin[c] or x[c] > g.gmax[c]):
                dele = True
            L = g.gmax[c] - g.gmin[c]
            if x[c] < g.gmin[c] - L or x[c] > g.gmax[c] + L:
                dele = True
        flags[i] = dele
    seq = rp.copy()
    nn, _ = oracle.bc_outflow(g, seq, sp.copy(), n)
    par = _parallel_outflow(flags)
    assert nn == par.size
    assert np.array_equal(seq[:nn, 7].astype(np.int64), par)


@pytest.mark.parametrize("pattern", ["counterexample", "inflow_tail", "all", "none", "alternate"])
def test_parallel_outflow_rule_patterns(orc, pattern):
    rng = np.random.default_rng(5)
    if pattern == "counterexample":       # 1-based n = 8, deleted {1, 2, 6}
        flags = np.array([1, 1, 0, 0, 0, 1, 0, 0])
    elif pattern == "inflow_tail":        # sparse losses below, dense losses in the appended tail
        flags = np.concatenate([(rng.random(3000) < 0.02), (rng.random(1500) < 0.5)]).astype(int)
    elif pattern == "all":
        flags = np.ones(50, int)
    elif pattern == "none":
        flags = np.zeros(50, int)
    else:
        flags = (np.arange(101) % 2).astype(int)
    n = flags.size
    g = oracle.Grid(4, 4, 4, 2, flow=(True, False, False))
    rp = np.zeros((n, 8))
    rp[:, 0] = np.where(flags == 1, g.gmax[0] + 1.0, 0.0)   # outside x only when flagged
    rp[:, 7] = np.arange(n)
    sp = np.ones(n, np.int32)
    nn, _ = oracle.bc_outflow(g, rp, sp, n)
    par = _parallel_outflow(flags)
    assert nn == par.size
    assert np.array_equal(rp[:nn, 7].astype(np.int64), par)


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123 (the
    generator behind the device sources' streams)."""
    from tests.philox import philox4x32
    cases = [
        ((0, 0, 0, 0), 0x0000000000000000, (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, 0xffffffffffffffff, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), 0x299f31d0a4093822,
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        got = tuple(int(w) for w in philox4x32(*ctr, key))
        assert got == want, [hex(v) for v in got]


def test_poisson_charge_haser_term_matches_oracle_distance():
    """The Poisson case's Haser term per cell equals e Q/(4 pi vn r^2) exp(-r/L) dV
    with r from the oracle's restatement of ParticlesProduce.f90:41-59 (itself
    pinned against the reference build); the box is symmetric about the source,
    so the charge is too; the blobs add a bounded part."""
    from oracle import haser
    cfg = host.config3(10)
    d = host.derive(cfg)
    dx, dy, dz = (float(v) for v in d.dxyz[:3])
    C = host.poisson_charge(cfg, blobs=0)
    Q, vn, L = host.HASER_Q_STATED, 1.0e3, 1.0e6
    ref = np.zeros_like(C)
    for k in range(cfg.nz):
        for j in range(cfg.ny):
            for i in range(cfg.nx):
                r1, r2 = haser.cell_distance(cfg.gmin[0] + dx * i, cfg.gmin[1] + dy * j,
                                             cfg.gmin[2] + dz * k, dx, dy, dz)
                ref[k + 1, j + 1, i + 1] = (host.EC * Q / (2.0 * host.TAU * r2 * vn) *
                                            np.exp(-r1 / L) * d.dV)
    np.testing.assert_allclose(C, ref, rtol=1e-14, atol=0)
    inner = C[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(inner, inner[::-1, ::-1, ::-1], rtol=1e-14)
    Cb = host.poisson_charge(cfg)
    amp = 1.0e-3 * host.EC * cfg.density[1] * d.dV
    assert np.all(np.abs(Cb - C) <= 16 * amp * (1 + 1e-12))
    assert np.any(Cb != C)


class _HalfRng:
    """A stand-in generator whose every uniform is 1/2."""

    def random(self, shape=None):
        return np.full(shape, 0.5) if shape is not None else 0.5


@pytest.mark.parametrize("flow", [(True, False, False), (True, True, True), (False, True, True),
                                  (True, False, True)])
def test_host_sources_follow_the_reference_layout(flow):
    """The host sources (seeded numpy draws) put their particles in the slots,
    species and cells of the oracle's restatement of BCinflow and
    ParticlesInitialize (pinned bit for bit against the reference build):
    with every uniform 1/2 both place each particle at its cell's centre, so
    positions agree exactly, face by face, including the z faces' y guard
    rows (BCinflow.f90:337, :407) and the + faces' L_max edges."""
    import oracle
    from oracle import sources
    cfg = host.Config(nx=5, ny=4, nz=3, species=4, flow=flow, gmin=(-25.0, -24.0, -12.0),
                      gmax=(25.0, 24.0, 12.0), ppc=(0, 4, 2, 1, 0, 0, 0, 0),
                      density=(0.0, 5.0e6, 1.0e6, 5.0e5, 0, 0, 0, 0),
                      charge3=(2 * host.EC, -host.EC) + (host.EC,) * 4)
    d = host.derive(cfg)
    g = oracle.Grid(5, 4, 3, 4, flow=flow, gmin=cfg.gmin, gmax=cfg.gmax)
    pl = dict(species=4, ipp=d.ipp, epp=d.epp, hpp=d.hpp, kelvin=d.kelvin, mass=d.mass,
              v0=d.v0[:, :3], ve=d.ve[:3], tB=d.tB[:, :3])
    half = sources.PerIonDraws(lambda face, local: np.full(64, 0.5))
    for hfn, ofn in ((host.particles_initialize, sources.particles_initialize),
                     (host.bc_inflow, sources.bc_inflow)):
        rh, sh = hfn(cfg, d, _HalfRng())
        ro, so = ofn(g, pl, half)
        assert np.array_equal(sh, so)
        assert np.array_equal(rh[:, :3], ro[:, :3])
        np.testing.assert_allclose(rh[:, 3:6], ro[:, 3:6], rtol=1e-13, atol=1e-9)
