"""ctypes bindings of libemu_amd.so (include/emu_amd.h, include/emu_init.h).

The library is the product: there is no fallback.  If libemu_amd.so is
missing or fails to load, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemu_amd.so")
MAX_ATTRIBS = 30

EMU_OK = 0
ERRORS = {1: "trace", 2: "negdiag", 3: "length", 4: "dt", 10: "arg", 11: "hip", 12: "unsupported",
          13: "nomem", 14: "move"}
PLAIN_FP64, FP32_BULK = 0, 1
SLOTS_PACKED, SLOTS_ALIGNED = 0x100, 0x200  # or-ed into emu_state_create's precision
DIRS_SPHERE, DIRS_CUBE26, DIRS_FIBONACCI = 0, 1, 2


class EmuParams(ctypes.Structure):
    _fields_ = [("nf", ctypes.c_int), ("shape_order", ctypes.c_int), ("ncell", ctypes.c_int * 3),
                ("plo", ctypes.c_double * 3), ("phi", ctypes.c_double * 3),
                ("rho", ctypes.c_double), ("Ye", ctypes.c_double), ("T", ctypes.c_double),
                ("mass", ctypes.c_double * 3),
                ("theta12", ctypes.c_double), ("theta13", ctypes.c_double), ("theta23", ctypes.c_double),
                ("alpha1", ctypes.c_double), ("alpha2", ctypes.c_double), ("deltaCP", ctypes.c_double),
                ("maxError", ctypes.c_double), ("cfl_factor", ctypes.c_double),
                ("flavor_cfl_factor", ctypes.c_double), ("max_adaptive_speedup", ctypes.c_double)]


class EmuParticles(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int64), ("pos", ctypes.c_void_p * 3), ("rdata", ctypes.c_void_p * MAX_ATTRIBS)]


class EmuGrid(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3), ("ncomp", ctypes.c_int)]


class EmuInitConfig(ctypes.Structure):
    _fields_ = [("simulation_type", ctypes.c_int), ("nppc", ctypes.c_int * 3), ("direction_set", ctypes.c_int),
                ("nphi_equator", ctypes.c_int), ("ndirections", ctypes.c_int),
                ("st3_wavelength_fraction_of_domain", ctypes.c_int), ("st3_amplitude", ctypes.c_double),
                ("st4_ndens", ctypes.c_double), ("st4_theta", ctypes.c_double), ("st4_phi", ctypes.c_double),
                ("st4_fluxfac", ctypes.c_double), ("st4_ndensbar", ctypes.c_double),
                ("st4_thetabar", ctypes.c_double), ("st4_phibar", ctypes.c_double),
                ("st4_fluxfacbar", ctypes.c_double), ("st4_amplitude", ctypes.c_double),
                ("st5", ctypes.c_double * 13), ("seed", ctypes.c_uint64)]


# (name, restype, argtypes)
_P = ctypes.c_void_p
_SIGS = [
    ("emu_create", ctypes.c_int, [ctypes.POINTER(EmuParams), ctypes.POINTER(_P)]),
    ("emu_destroy", ctypes.c_int, [_P]),
    ("emu_set_stream", ctypes.c_int, [_P, _P]),
    ("emu_sync", ctypes.c_int, [_P]),
    ("emu_last_error", ctypes.c_char_p, [_P]),
    ("emu_grid_ngrow", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    ("emu_deposit_to_mesh", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(EmuGrid)]),
    ("emu_deposit_to_mesh_static", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(EmuGrid)]),
    ("emu_fill_boundary", ctypes.c_int, [_P, ctypes.POINTER(EmuGrid)]),
    ("emu_interpolate_rhs_from_mesh", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(EmuGrid)]),
    ("emu_compute_dt", ctypes.c_int, [_P, ctypes.POINTER(EmuGrid), ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    ("emu_vvac_max", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(ctypes.c_double)]),
    ("emu_renormalize", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(ctypes.c_int64)]),
    ("emu_sync_location", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.c_int]),
    ("emu_redistribute_local", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles)]),
    ("emu_apply_rhs", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.c_double, ctypes.POINTER(EmuParticles)]),
    ("emu_copy_particles", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(EmuParticles)]),
    ("emu_rk_combine", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), ctypes.POINTER(EmuParticles), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(EmuParticles)]),
    ("emu_state_create", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_P)]),
    ("emu_state_destroy", ctypes.c_int, [_P]),
    ("emu_state_load", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), _P]),
    ("emu_state_set_vvac_max", ctypes.c_int, [_P, ctypes.c_double]),
    ("emu_state_store", ctypes.c_int, [_P, ctypes.POINTER(EmuParticles), _P]),
    ("emu_state_evolve", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double]),
    ("emu_state_info", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("emu_state_profile", ctypes.c_int, [_P, ctypes.c_int]),
    ("emu_state_kernel_ms", ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_int64)]),
    ("emu_state_flavor_amplitude", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    ("emu_adj_work_doubles", ctypes.c_int64, [_P, ctypes.c_int64]),
    ("emu_adj_rhs", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P]),
    ("emu_adj_rhs_vjp", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, _P]),
    ("emu_adj_objective", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, ctypes.POINTER(ctypes.c_double), _P, _P]),
    ("emu_adj_advance_positions", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, ctypes.c_double]),
    ("emu_adj_renormalize", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
    ("emu_adj_axpy", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_double, _P, _P, _P]),
    ("emu_adj_step_work_doubles", ctypes.c_int64, [_P, ctypes.c_int64]),
    ("emu_adj_step", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P]),
    ("emu_adj_step_vjp", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P]),
    ("emu_adj_step_work_doubles_f32", ctypes.c_int64, [_P, ctypes.c_int64]),
    ("emu_adj_step_f32", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, ctypes.c_double,
                                        ctypes.c_double, _P]),
    ("emu_adj_step_vjp_f32", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double,
                                            ctypes.c_double, _P]),
    ("emu_adj_step_keep_doubles_f32", ctypes.c_int64, [_P, ctypes.c_int64]),
    ("emu_adj_step_keep_f32", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, ctypes.c_double,
                                             ctypes.c_double, _P, _P]),
    ("emu_adj_step_vjp_kept_f32", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, _P,
                                                 ctypes.c_double, ctypes.c_double, _P]),
    ("emu_adj_objective_f32", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, ctypes.POINTER(ctypes.c_double), _P,
                                             _P]),
    ("emu_adj_permute_f32", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _P, _P, _P, ctypes.c_int]),
    ("emu_adj_step_keep_doubles", ctypes.c_int64, [_P, ctypes.c_int64]),
    ("emu_adj_step_keep", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P, _P]),
    ("emu_adj_step_vjp_kept", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double,
                                             _P]),
    ("emu_adj_status", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("emu_adj_cell_order", ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P]),
    ("emu_adj_permute", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _P, _P, _P, ctypes.c_int]),
    ("emu_init_count", ctypes.c_int64, [ctypes.POINTER(EmuParams), ctypes.POINTER(EmuInitConfig)]),
    ("emu_init_particles", ctypes.c_int, [_P, ctypes.POINTER(EmuInitConfig), ctypes.POINTER(EmuParticles), _P]),
]

EXPORTED = [name for name, _, _ in _SIGS]

_lib = None


def lib():
    """Load libemu_amd.so (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m emu_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


class EmuError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = "", bad_index: int = -1):
        self.code = code
        self.bad_index = bad_index
        super().__init__(f"{where}: emu error {code} ({ERRORS.get(code, '?')}) {msg}".strip())


def check(code: int, where: str, ctx=None, bad_index: int = -1):
    if code != EMU_OK:
        msg = lib().emu_last_error(ctx).decode() if ctx else ""
        raise EmuError(code, where, msg, bad_index)
