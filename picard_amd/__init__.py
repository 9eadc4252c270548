"""picard_amd -- MI355X (gfx950) implementation of Picard's PIC time step.

The product is libpicard_hip.so (HIP kernels behind the C ABI in
include/picard_hip.h).  This package holds its ctypes binding and the host
side: run configuration, the derived constants of picard.f90, the host
particle sources and a device-resident Domain."""
from . import _lib  # noqa: F401
from .host import (Config, Derived, Domain, bc_inflow, config1, derive,  # noqa: F401
                   particles_initialize)

__all__ = ["Config", "Derived", "Domain", "bc_inflow", "config1", "derive",
           "particles_initialize"]
