"""ctypes binding of libpicard_hip.so (include/picard_hip.h, include/picard_ensemble.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C picard_amd/csrc``).  There is no fallback: if the library is
missing or cannot be loaded, every entry point raises."""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpicard_hip.so")

PICARD_OK = 0
PICARD_ERR_ARG = -1
PICARD_ERR_HIP = -2
PICARD_ERR_CAPACITY = -3
PICARD_ERR_RANGE = -4
PICARD_ERR_STATE = -5
PICARD_WARN_NOCONV = 1

PICARD_FP64 = 0
PICARD_FP32_SPLIT = 1
PICARD_DEPOSIT_REFERENCE_ORDER = 0
PICARD_DEPOSIT_FIXED_POINT = 1

FIELD_U, FIELD_E, FIELD_N, FIELD_C, FIELD_D, FIELD_F = 0, 1, 2, 3, 4, 5

# every symbol include/picard_hip.h declares
EXPORTS = (
    "picard_last_error", "picard_create", "picard_destroy", "picard_set_species",
    "picard_synchronize", "picard_particles_append", "picard_particles_read",
    "picard_particles_count", "picard_advance_position", "picard_bc_outflow",
    "picard_particles_transfer", "picard_particles_gridify", "picard_calc_distance",
    "picard_bc_potentials", "picard_calc_epotential", "picard_calc_efield", "picard_bc_fields",
    "picard_advance_velocity", "picard_field_read", "picard_field_write", "picard_step",
    "picard_timing_enable", "picard_timing_read", "picard_timing_solve", "picard_split_solver",
    "picard_split_sweeps_per_launch", "picard_generate_initial",
    "picard_generate_inflow", "picard_particles_produce", "picard_set_production",
    "picard_diagnostics", "picard_advance_velocity_imp", "picard_advance_velocity_exp",
    "picard_particles_gridify_all", "picard_particles_read_noalloc", "picard_timing_sweeps",
    "picard_timing_kernel_clock",
)

READ_NOALLOC_MAX = 246723

# every symbol include/picard_ensemble.h declares
ENSEMBLE_EXPORTS = (
    "picard_sim_derive", "picard_sim_footprint", "picard_ensemble_create",
    "picard_ensemble_destroy", "picard_ensemble_add", "picard_ensemble_remove",
    "picard_ensemble_step", "picard_ensemble_status", "picard_ensemble_field_read",
    "picard_ensemble_particles_read", "picard_ensemble_tune", "picard_ensemble_set_tile",
    "picard_ensemble_get_tile",
)


class Geometry(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("species", ct.c_int32),
        ("flow", ct.c_int32 * 3),
        ("precision", ct.c_int32),
        ("deposit", ct.c_int32),
        ("device", ct.c_int32),
        ("max_particles", ct.c_int64),
        ("lmin", ct.c_double * 3), ("lmax", ct.c_double * 3),
        ("gmin", ct.c_double * 3), ("gmax", ct.c_double * 3),
        ("dxyz", ct.c_double * 3),
    ]


class Species(ct.Structure):
    _fields_ = [
        ("charge", ct.c_double * 8), ("n2p", ct.c_double * 8), ("eta", ct.c_double * 8),
        ("tB", ct.c_double * 32), ("sB", ct.c_double * 32),
        ("ve", ct.c_double * 4), ("B0", ct.c_double * 4), ("E0", ct.c_double * 4),
        ("fluid", ct.c_int32 * 8), ("gyration", ct.c_int32 * 8),
    ]


class Plasma(ct.Structure):
    _fields_ = [("ipp", ct.c_int32 * 8), ("epp", ct.c_int32 * 8), ("hpp", ct.c_int32 * 8),
                ("mass", ct.c_double * 8), ("kelvin", ct.c_double * 8), ("v0", ct.c_double * 32)]


class Haser(ct.Structure):
    _fields_ = [("Qn", ct.c_double * 8), ("vn", ct.c_double * 8), ("nu_i", ct.c_double * 8),
                ("nu_d", ct.c_double * 8), ("Ek", ct.c_double * 8), ("rate_scale", ct.c_double)]


class Diag(ct.Structure):
    _fields_ = [("num_local", ct.c_int64), ("num_global", ct.c_int64),
                ("particle_energy", ct.c_double), ("e_energy", ct.c_double),
                ("b_energy", ct.c_double)]


class StepStats(ct.Structure):
    _fields_ = [("num_local", ct.c_int64), ("removed", ct.c_int64), ("sweeps", ct.c_int64),
                ("pushed", ct.c_int64)]


class SimParams(ct.Structure):
    _fields_ = [("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32), ("ppc", ct.c_int32),
                ("dx", ct.c_double), ("density", ct.c_double), ("kelvin", ct.c_double),
                ("drift", ct.c_double), ("B", ct.c_double * 3), ("dt_wpe", ct.c_double),
                ("maxerr", ct.c_double), ("maxit", ct.c_int32), ("reserved", ct.c_int32),
                ("seed", ct.c_uint64)]


class SimDerived(ct.Structure):
    _fields_ = [("dxyz", ct.c_double), ("dV", ct.c_double), ("lmin", ct.c_double * 3),
                ("lmax", ct.c_double * 3), ("charge", ct.c_double * 2), ("n2p", ct.c_double * 2),
                ("eta", ct.c_double * 2), ("ve", ct.c_double * 3), ("dt", ct.c_double),
                ("omega_pe", ct.c_double), ("tB", (ct.c_double * 3) * 2), ("tB4", ct.c_double * 2),
                ("sB", (ct.c_double * 3) * 2)]


class EnsembleConfig(ct.Structure):
    _fields_ = [("device", ct.c_int32), ("max_sims", ct.c_int32), ("max_cells", ct.c_int64),
                ("max_particles", ct.c_int64), ("inner", ct.c_int32), ("use_graphs", ct.c_int32)]


class EnsembleTile(ct.Structure):
    _fields_ = [("cells_per_thread", ct.c_int32), ("threads", ct.c_int32),
                ("blocks_per_cu", ct.c_int32), ("reserved", ct.c_int32)]


class EnsembleStats(ct.Structure):
    _fields_ = [("sims", ct.c_int32), ("noconv", ct.c_int32), ("particles", ct.c_int64),
                ("pushed", ct.c_int64), ("max_sweeps", ct.c_int64), ("sum_sweeps", ct.c_int64),
                ("solve_chunks", ct.c_int64), ("cell_sweeps", ct.c_int64), ("step_ms", ct.c_double),
                ("solve_ms", ct.c_double)]


class SimStatus(ct.Structure):
    _fields_ = [("live", ct.c_int32), ("noconv", ct.c_int32), ("steps", ct.c_int64),
                ("num_local", ct.c_int64), ("pushed", ct.c_int64), ("sweeps", ct.c_int64)]


class PicardError(RuntimeError):
    def __init__(self, fn, rc, msg):
        super().__init__(f"{fn} failed with status {rc}: {msg}")
        self.rc = rc


_lib = None


def lib():
    """Load libpicard_hip.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                           "or make -C picard_amd/csrc")
    L = ct.CDLL(LIB_PATH)
    dp, ip, lp = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int64)
    vp = ct.c_void_p
    sig = {
        "picard_last_error": ([], ct.c_char_p),
        "picard_create": ([ct.POINTER(Geometry), ct.POINTER(Species), ct.POINTER(vp)], ct.c_int),
        "picard_destroy": ([vp], ct.c_int),
        "picard_set_species": ([vp, ct.POINTER(Species)], ct.c_int),
        "picard_synchronize": ([vp], ct.c_int),
        "picard_particles_append": ([vp, dp, ip, ct.c_int64], ct.c_int),
        "picard_particles_read": ([vp, ct.c_int64, ct.c_int64, dp, ip], ct.c_int),
        "picard_particles_count": ([vp, lp], ct.c_int),
        "picard_particles_read_noalloc": ([vp, ct.c_int64, ct.c_int64, dp, ip], ct.c_int),
        "picard_timing_sweeps": ([vp, lp], ct.c_int),
        "picard_timing_kernel_clock": ([vp, dp, lp], ct.c_int),
        "picard_advance_position": ([vp, ct.c_double], ct.c_int),
        "picard_bc_outflow": ([vp, lp, lp], ct.c_int),
        "picard_particles_transfer": ([vp], ct.c_int),
        "picard_particles_gridify": ([vp], ct.c_int),
        "picard_calc_distance": ([vp], ct.c_int),
        "picard_bc_potentials": ([vp], ct.c_int),
        "picard_calc_epotential": ([vp, ct.c_double, ct.c_int32, lp], ct.c_int),
        "picard_calc_efield": ([vp], ct.c_int),
        "picard_bc_fields": ([vp], ct.c_int),
        "picard_advance_velocity": ([vp, ct.c_double], ct.c_int),
        "picard_field_read": ([vp, ct.c_int32, dp, ct.c_int64], ct.c_int),
        "picard_field_write": ([vp, ct.c_int32, dp, ct.c_int64], ct.c_int),
        "picard_step": ([vp, ct.c_double, ct.c_double, ct.c_int32, dp, ip, ct.c_int64,
                         ct.POINTER(StepStats)], ct.c_int),
        "picard_timing_enable": ([vp, ct.c_int32], ct.c_int),
        "picard_timing_read": ([vp, dp, lp], ct.c_int),
        "picard_timing_solve": ([vp, dp, lp], ct.c_int),
        "picard_split_solver": ([vp, ct.c_int32], ct.c_int),
        "picard_split_sweeps_per_launch": ([vp, ct.c_int32], ct.c_int),
        "picard_generate_initial": ([vp, ct.POINTER(Plasma), ct.c_uint64], ct.c_int),
        "picard_generate_inflow": ([vp, ct.POINTER(Plasma), ct.c_uint64, ct.c_int64], ct.c_int),
        "picard_particles_produce": ([vp, ct.POINTER(Plasma), ct.POINTER(Haser), ct.c_uint64,
                                      ct.c_int64, ct.c_double, lp], ct.c_int),
        "picard_set_production": ([vp, ct.POINTER(Plasma), ct.POINTER(Haser), ct.c_uint64],
                                  ct.c_int),
        "picard_diagnostics": ([vp, dp, ct.POINTER(Diag)], ct.c_int),
        "picard_advance_velocity_imp": ([vp, ct.c_double, ct.c_double], ct.c_int),
        "picard_particles_gridify_all": ([vp], ct.c_int),
        "picard_advance_velocity_exp": ([vp, ct.c_double, ct.c_double], ct.c_int),
        # ensemble (include/picard_ensemble.h)
        "picard_sim_derive": ([ct.POINTER(SimParams), ct.POINTER(SimDerived)], ct.c_int),
        "picard_sim_footprint": ([ct.POINTER(SimParams), lp, lp], ct.c_int),
        "picard_ensemble_create": ([ct.POINTER(EnsembleConfig), ct.POINTER(vp)], ct.c_int),
        "picard_ensemble_destroy": ([vp], ct.c_int),
        "picard_ensemble_add": ([vp, ct.POINTER(SimParams), ip], ct.c_int),
        "picard_ensemble_remove": ([vp, ct.c_int32], ct.c_int),
        "picard_ensemble_step": ([vp, ct.POINTER(EnsembleStats)], ct.c_int),
        "picard_ensemble_status": ([vp, ct.c_int32, ct.POINTER(SimStatus)], ct.c_int),
        "picard_ensemble_field_read": ([vp, ct.c_int32, ct.c_int32, dp, ct.c_int64], ct.c_int),
        "picard_ensemble_particles_read": ([vp, ct.c_int32, dp, ip, ct.c_int64], ct.c_int),
        "picard_ensemble_tune": ([vp, ct.c_int32, ct.POINTER(EnsembleTile), dp, ct.c_int32, ip],
                                 ct.c_int),
        "picard_ensemble_set_tile": ([vp, ct.POINTER(EnsembleTile)], ct.c_int),
        "picard_ensemble_get_tile": ([vp, ct.POINTER(EnsembleTile)], ct.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(fn: str, rc: int, allow_warn: bool = True) -> int:
    if rc < 0 or (rc > 0 and not allow_warn):
        msg = lib().picard_last_error().decode(errors="replace")
        raise PicardError(fn, rc, msg)
    return rc
