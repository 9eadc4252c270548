"""The mixed-precision Poisson case (Gaussian blobs plus a Haser 1/r^2
profile, every boundary open, relative residual 1e-10) at reduced sizes:
the FP32-split solve against the FP64 solve (bit-exact to the reference's
CalcEpotential), and both against what the stop rule fixes -- the final
potential's FP64 residual b - lap(U), evaluated here on the host, is within
the tolerance of |b| (CalcEpotential.f90:50-62)."""
import numpy as np
import pytest

from picard_amd import _lib, host

pytestmark = pytest.mark.gpu

EPS0 = 8.854187817e-12


def _say(msg):
    print(f"[gpu-poisson] {msg}", flush=True)


def _residual_ratio(U, C, d):
    """|b - lap U| / |b| over interior nodes, b = -C / (dV eps0)."""
    dx, dy, dz = (float(v) for v in d.dxyz[:3])
    b = -C[1:-1, 1:-1, 1:-1] / (d.dV * EPS0)
    c = U[1:-1, 1:-1, 1:-1]
    lap = ((U[1:-1, 1:-1, 2:] + U[1:-1, 1:-1, :-2] - 2 * c) / (dx * dx) +
           (U[1:-1, 2:, 1:-1] + U[1:-1, :-2, 1:-1] - 2 * c) / (dy * dy) +
           (U[2:, 1:-1, 1:-1] + U[:-2, 1:-1, 1:-1] - 2 * c) / (dz * dz))
    return float(np.linalg.norm(b - lap) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [40, 42])
def test_poisson_case_split_vs_fp64(n):
    cfg = host.config3(n)
    d = host.derive(cfg)
    C = host.poisson_charge(cfg)
    _say(f"n={n}: |C| {np.abs(C).max():.3e}")
    out = []
    for prec in (_lib.PICARD_FP64, _lib.PICARD_FP32_SPLIT):
        dom = host.Domain(cfg, max_particles=1024, d=d, precision=prec)
        dom.calc_distance()
        dom.set_field(_lib.FIELD_C, C)
        dom.bc_potentials()
        sw = dom.calc_epotential()
        U = dom.field(_lib.FIELD_U)
        out.append((sw, U))
        dom.close()
    (sa, Ua), (sb, Ub) = out
    ra, rb = _residual_ratio(Ua, C, d), _residual_ratio(Ub, C, d)
    rel = float(np.linalg.norm(Ub - Ua) / np.linalg.norm(Ua))
    _say(f"n={n}: sweeps fp64={sa} split={sb}  residual fp64 {ra:.3e} split {rb:.3e}  "
         f"U rel {rel:.3e}")
    assert sa < cfg.maxit and sb < cfg.maxit
    assert ra <= 1.0e-10 * (1 + 1e-6)
    assert rb <= 1.2e-10
    # FP32 rounding moves the split stop by a sweep or two; with two sweeps
    # per launch every second iterate is tested near the threshold (every
    # 32nd far from it), bounded here by seven sweeps more
    assert -2 <= sb - sa <= 9
    assert rel <= 1e-8


@pytest.mark.timeout(1500)
def test_poisson_512_split_to_1e10(capsys):
    """The case at its stated size: 512^3 cells, 16 Gaussian blobs plus the
    Haser profile, every boundary open (the FFT BCpotentials on all six
    faces), FP32 sweeps with FP64 refinement run to relative residual 1e-10
    in one solve (about 1.2e6 sweeps, so maxit = 2^21: the reference's
    default 2^20 would print its warning past that count, CalcEpotential.f90:
    67-69).  Checked against what the stop rule fixes: the final potential's
    FP64 residual b - lap(U), evaluated here on the host, within 1.2e-10 of
    |b|, and the solve stopping below 2*maxit+1.  The solve runs in a worker
    thread (the C call releases the interpreter) so progress prints while it
    runs."""
    import threading
    import time
    n = 512
    cfg = host.config3(n, maxit=2 ** 21)
    d = host.derive(cfg)
    C = host.poisson_charge(cfg)
    dom = host.Domain(cfg, max_particles=1024, d=d, precision=_lib.PICARD_FP32_SPLIT)
    t0 = time.perf_counter()
    dom.calc_distance()
    dom.set_field(_lib.FIELD_C, C)
    dom.bc_potentials()
    dom.synchronize()
    with capsys.disabled():
        _say(f"512^3: distance table and open-boundary potential in {time.perf_counter() - t0:.1f} s")
    out = {}

    def solve():
        try:
            out["sweeps"] = dom.calc_epotential()
        except Exception as e:          # reported by the main thread
            out["error"] = e

    th = threading.Thread(target=solve)
    t0 = time.perf_counter()
    th.start()
    while th.is_alive():
        th.join(20.0)
        if th.is_alive():
            with capsys.disabled():
                _say(f"512^3 split solve running, {time.perf_counter() - t0:.0f} s")
    el = time.perf_counter() - t0
    assert "error" not in out, out.get("error")
    sw = out["sweeps"]
    U = dom.field(_lib.FIELD_U)
    dom.close()
    r = _residual_ratio(U, C, d)
    with capsys.disabled():
        _say(f"512^3 split: {sw} sweeps in {el:.1f} s ({1e6 * el / sw:.1f} us per sweep), "
             f"host FP64 residual {r:.3e}")
    assert 1.0e5 < sw < 2 * cfg.maxit + 1
    assert r <= 1.2e-10
