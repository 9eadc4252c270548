"""picard_particles_read_noalloc (staged through the library's fixed device
array, no allocation) against picard_particles_read on the same particles,
bit for bit, in both precision forms: two calls in a row, the second smaller
and from another offset, so the second reads what the first left in the
array; the per-call limit is refused with a status."""
import ctypes as ct

import numpy as np
import pytest

from picard_amd import _lib, host
from picard_amd._lib import lib

pytestmark = pytest.mark.gpu


def _noalloc(dom, first, count):
    rp = np.full((count, 8), np.nan)
    sp = np.full(count, -7, np.int32)
    rc = lib().picard_particles_read_noalloc(dom.h, first, count,
                                             rp.ctypes.data_as(ct.POINTER(ct.c_double)),
                                             sp.ctypes.data_as(ct.POINTER(ct.c_int32)))
    return rc, rp, sp


@pytest.mark.parametrize("precision", [_lib.PICARD_FP64, _lib.PICARD_FP32_SPLIT])
def test_read_noalloc_matches_read(precision):
    print(f"[gpu-noalloc] precision={precision}", flush=True)
    cfg = host.Config(nx=40, ny=40, nz=40, flow=(True, False, False), gmin=(-200.0,) * 3,
                      gmax=(200.0,) * 3, ppc=(0, 3, 0, 0, 0, 0, 0, 0))
    d = host.derive(cfg)
    dom = host.Domain(cfg, max_particles=500_000, d=d, precision=precision)
    dom.generate_initial(20261212)
    n = dom.num_local
    assert n == 40 ** 3 * 6 > _lib.READ_NOALLOC_MAX
    for first, count in ((n - _lib.READ_NOALLOC_MAX, _lib.READ_NOALLOC_MAX), (1234, 777)):
        rc, rp, sp = _noalloc(dom, first, count)
        assert rc == _lib.PICARD_OK, lib().picard_last_error()
        ref_rp, ref_sp = dom.particles(first, count)
        assert np.array_equal(rp, ref_rp) and np.array_equal(sp, ref_sp), (first, count)
    rc, _, _ = _noalloc(dom, 0, _lib.READ_NOALLOC_MAX + 1)
    assert rc == _lib.PICARD_ERR_ARG
    assert "at most 246723" in lib().picard_last_error().decode()
    rc, _, _ = _noalloc(dom, n - 10, 11)
    assert rc == _lib.PICARD_ERR_ARG
    dom.close()
