"""Loader and ctypes declarations of librrlbm.so (include/rrlbm.h).

The library is the only compute path: if librrlbm.so is missing or fails to
load, every entry point raises -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "librrlbm.so")


class RRLBMError(RuntimeError):
    """A non-zero rrlbm_status, with the library's message."""

    def __init__(self, status, message):
        super().__init__("rrlbm status %d: %s" % (status, message))
        self.status = status


class Params(ctypes.Structure):
    """rrlbm_params: run-time counterpart of parameters.h / globals.cu."""

    _fields_ = [
        ("lx", ctypes.c_int32), ("ly", ctypes.c_int32), ("lz", ctypes.c_int32),
        ("npop", ctypes.c_int32), ("use_les", ctypes.c_int32), ("les_lagged", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("wall_x", ctypes.c_int32), ("wall_y", ctypes.c_int32), ("wall_z", ctypes.c_int32),
        ("wall_rr", ctypes.c_int32), ("lid_face", ctypes.c_int32),
        ("tau", ctypes.c_double), ("force", ctypes.c_double * 3), ("cw", ctypes.c_double),
        ("lid_u", ctypes.c_double * 3), ("f16_scale", ctypes.c_double),
        ("xcd_remap", ctypes.c_int32), ("cells_per_thread", ctypes.c_int32), ("f16_rounding", ctypes.c_int32),
        ("pop_group_rows", ctypes.c_int32), ("pop_pad_bytes", ctypes.c_int32),
        ("persistent_wg_per_cu", ctypes.c_int32), ("collision_model", ctypes.c_int32),
        ("rr_order", ctypes.c_int32), ("steps_per_pass", ctypes.c_int32), ("pass_planes", ctypes.c_int32),
    ]


class Instance(ctypes.Structure):
    """rrlbm_instance: one periodic D3Q27 box of a batched ensemble."""

    _fields_ = [("lx", ctypes.c_int32), ("ly", ctypes.c_int32), ("lz", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("tau", ctypes.c_double), ("u0", ctypes.c_double)]


class Diag(ctypes.Structure):
    """rrlbm_diag: FP64 sums of diag_flow (output.cu:24-35) and conservation checks."""

    _fields_ = [
        ("mass_excess", ctypes.c_double), ("u_sum", ctypes.c_double * 3),
        ("momentum", ctypes.c_double * 3), ("kinetic_energy", ctypes.c_double),
        ("nonfinite_cells", ctypes.c_double), ("ncells", ctypes.c_double),
    ]

    def as_dict(self):
        return {
            "mass_excess": self.mass_excess, "u_sum": list(self.u_sum),
            "momentum": list(self.momentum), "kinetic_energy": self.kinetic_energy,
            "nonfinite_cells": self.nonfinite_cells, "ncells": self.ncells,
        }


class Tuning(ctypes.Structure):
    """rrlbm_tuning: the launch shape and population layout rrlbm_autotune chose."""

    _fields_ = [
        ("pop_group_rows", ctypes.c_int32), ("pop_pad_bytes", ctypes.c_int32),
        ("cells_per_thread", ctypes.c_int32), ("xcd_remap", ctypes.c_int32),
        ("persistent_wg_per_cu", ctypes.c_int32), ("candidates", ctypes.c_int32),
        ("ms_per_step", ctypes.c_double), ("ms_per_step_default", ctypes.c_double),
        ("steps_per_pass", ctypes.c_int32), ("pass_planes", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int

# name -> (argtypes); every function returns int status except where noted
SIGNATURES = {
    "rrlbm_params_default": [ctypes.POINTER(Params)],
    "rrlbm_lattice": [_I, ctypes.POINTER(ctypes.c_int32), _D, ctypes.POINTER(ctypes.c_int32)],
    "rrlbm_device_count": [ctypes.POINTER(ctypes.c_int)],
    "rrlbm_create": [ctypes.POINTER(Params), ctypes.POINTER(_P)],
    "rrlbm_destroy": [_P],
    "rrlbm_get_params": [_P, ctypes.POINTER(Params)],
    "rrlbm_synchronize": [_P],
    "rrlbm_memory_bytes": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "rrlbm_upload": [_P, _I, _D, ctypes.c_size_t],
    "rrlbm_download": [_P, _I, _D, ctypes.c_size_t],
    "rrlbm_init_arrays": [_P, ctypes.c_double, ctypes.c_double],
    "rrlbm_forcing": [_P],
    "rrlbm_forcing_gp": [_P, _I, ctypes.c_double],
    "rrlbm_les_wale": [_P],
    "rrlbm_collision_rr": [_P],
    "rrlbm_collision_bgk": [_P],
    "rrlbm_streaming": [_P],
    "rrlbm_bb": [_P],
    "rrlbm_swap_f": [_P],
    "rrlbm_macrovar": [_P],
    "rrlbm_rrbc": [_P],
    "rrlbm_macrovar_bn": [_P],
    "rrlbm_fdbc": [_P],
    "rrlbm_step_reference": [_P, _I],
    "rrlbm_init_rest": [_P],
    "rrlbm_init_taylor_green_2d": [_P, ctypes.c_double],
    "rrlbm_init_taylor_green_3d": [_P, ctypes.c_double],
    "rrlbm_init_channel": [_P, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                           ctypes.c_double],
    "rrlbm_init_equilibrium": [_P],
    "rrlbm_fused_pack": [_P],
    "rrlbm_fused_step": [_P, _I],
    "rrlbm_fused_step_gp": [_P, _I, ctypes.c_double],
    "rrlbm_fused_step_timed": [_P, _I, _D],
    "rrlbm_fused_unpack": [_P],
    "rrlbm_fused_diag": [_P, ctypes.POINTER(Diag)],
    "rrlbm_diag_flow": [_P, ctypes.POINTER(Diag)],
    "rrlbm_diag_flow_noalloc": [_D, _D, _D, _D, ctypes.c_int64, ctypes.POINTER(Diag), _P],
    "rrlbm_plane_profiles": [_P, _D, ctypes.c_size_t],
    "rrlbm_write_flow": [_P, ctypes.c_char_p, _I],
    "rrlbm_write_f": [_P, ctypes.c_char_p, _I],
    "rrlbm_read_restart": [_P, ctypes.c_char_p, _I],
    "rrlbm_release_reference": [_P],
    "rrlbm_autotune": [_P, _I, ctypes.POINTER(Tuning)],
    "rrlbm_graph_capture": [_P, _I],
    "rrlbm_graph_launch": [_P, _I],
    "rrlbm_graph_launch_timed": [_P, _I, _D],
    "rrlbm_batch_set_grid": [_P, _I],
    "rrlbm_batch_autotune": [_P, _I, ctypes.POINTER(ctypes.c_int), _D],
    "rrlbm_batch_create": [ctypes.POINTER(Instance), _I, _I, _I, ctypes.POINTER(_P)],
    "rrlbm_batch_destroy": [_P],
    "rrlbm_batch_cells": [_P, ctypes.POINTER(ctypes.c_int64)],
    "rrlbm_batch_init_taylor_green_3d": [_P],
    "rrlbm_batch_step": [_P, _I],
    "rrlbm_batch_step_sequential": [_P, _I],
    "rrlbm_batch_graph_capture": [_P, _I],
    "rrlbm_batch_graph_launch": [_P, _I],
    "rrlbm_batch_time": [_P, _I, _I, _D],
    "rrlbm_batch_diag": [_P, _I, ctypes.POINTER(Diag)],
    "rrlbm_batch_context": [_P, _I, ctypes.POINTER(_P)],
}

_lib = None


def load():
    """Load librrlbm.so once; raise loudly when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RRLBMError(-1, "librrlbm.so not found at %s: build it with "
                             "`python recursively-regularized-lbm_amd/build_lib.py`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.rrlbm_version.restype = ctypes.c_char_p
    lib.rrlbm_version.argtypes = []
    lib.rrlbm_last_error.restype = ctypes.c_char_p
    lib.rrlbm_last_error.argtypes = [_P]
    lib.rrlbm_batch_last_error.restype = ctypes.c_char_p
    lib.rrlbm_batch_last_error.argtypes = [_P]
    _lib = lib
    return lib


def check(status, ctx=None):
    if status != 0:
        msg = load().rrlbm_last_error(ctx)
        raise RRLBMError(status, msg.decode() if msg else "")


def check_batch(status, batch=None):
    if status != 0:
        msg = load().rrlbm_batch_last_error(batch)
        raise RRLBMError(status, msg.decode() if msg else "")
