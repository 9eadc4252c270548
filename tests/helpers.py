"""Shared test helpers: seeded synthetic inputs on small grids."""
import numpy as np

ELEC = 1.60217646e-19  # picard.f90:57
EPS0 = 8.854187817e-12  # picard.f90:65


def random_particles(rng, g, n, species=2, spread=0.0, ids=True):
    """Particles uniform over the global box widened by `spread` cells each side."""
    rp = np.zeros((n, 8), np.float64)
    for c in range(3):
        lo = g.gmin[c] - spread * g.dxyz[c]
        hi = g.gmax[c] + spread * g.dxyz[c]
        rp[:, c] = lo + (hi - lo) * rng.random(n)
        rp[:, 3 + c] = rng.normal(0.0, 3.0e5, n)
    if ids:
        rp[:, 7] = np.arange(1, n + 1, dtype=np.float64)
        rp[:, 6] = rng.random(n)
    sp = rng.integers(1, species + 1, n).astype(np.int32)
    return rp, sp


def boris_tables(dt, B0, eta, gyration=(1,) * 8):
    """tB/sB exactly as picard.f90:452-466."""
    B0 = np.asarray(B0, np.float64)
    B4 = np.sqrt(B0[0] * B0[0] + B0[1] * B0[1] + B0[2] * B0[2])
    Bv = np.array([B0[0], B0[1], B0[2], B4])
    tB = np.zeros((8, 4))
    for s in range(8):
        if gyration[s]:
            tB[s, :] = 0.5 * dt * Bv * eta[s]
    sB = 2.0 * tB / (1.0 + tB[:, 3:4] ** 2)
    return tB, sB
