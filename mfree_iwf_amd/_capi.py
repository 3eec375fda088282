"""ctypes binding of the C ABI declared in include/mfree_iwf_amd.h.

The shared library is built in-tree (``make -C mfree_iwf_amd/csrc``) and is
loaded from next to this file.  There is no fallback: if the library is
missing, importing the binding raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfree_iwf_amd.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NBH_OVERFLOW = -3
ERR_NO_NEIGHBOR = -4
ERR_NONFINITE = -5
ERR_CELLS = -6
ERR_PIPELINE = -7
ERR_CAPACITY = -9
ERR_UNSUPPORTED = -10
ERR_SCRATCH = -11
ERR_SKIN = -12

METHOD_UL_MONAGHAN = 0
METHOD_TL_RL = 1
METHOD_UL_GODUNOV = 2
METHOD_TL_COROTATIONAL = 3
METHOD_TL_POTENTIAL = 4
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PIPELINE_FUSED = 0
PIPELINE_REFERENCE = 1

dptr = C.POINTER(C.c_double)
u32ptr = C.POINTER(C.c_uint32)
fptr = C.POINTER(C.c_float)
i32ptr = C.POINTER(C.c_int32)


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "E", "nu", "rho0", "wdeltap", "corr_exp", "mghn_eps",
        "alpha", "beta", "eta", "xsph_eps", "dt", "time")]


class Options(C.Structure):
    _fields_ = [("method", C.c_int32), ("precision", C.c_int32), ("device", C.c_int32),
                ("max_nbh", C.c_uint32), ("use_graph", C.c_int32), ("deterministic", C.c_int32),
                ("tl_h_search", C.c_double), ("tile_cells", C.c_int32), ("pipeline", C.c_int32),
                ("xcd_chunk", C.c_uint32), ("cell_scale", C.c_double), ("neighbor_list", C.c_int32),
                ("list_lanes", C.c_int32), ("isolate_bodies", C.c_int32),
                ("list_staging", C.c_int32)]


PARTICLE_FIELDS = ("x", "y", "X", "Y", "m", "rho", "h", "vx", "vy", "Sxx", "Sxy", "Syy", "fcx", "fcy")


class Particles(C.Structure):
    _fields_ = [(n, dptr) for n in PARTICLE_FIELDS]


class RingsSpec(C.Structure):
    _fields_ = [("ri", C.c_double), ("ro", C.c_double), ("centre", C.c_double), ("nbox", C.c_uint32),
                ("hdx", C.c_double), ("rho0", C.c_double), ("v0", C.c_double), ("shuffle", C.c_int32),
                ("seed", C.c_uint64), ("jitter", C.c_double)]


class PlateSpec(C.Structure):
    _fields_ = [("L", C.c_double), ("W", C.c_double), ("r_hole", C.c_double), ("dx", C.c_double),
                ("hdx", C.c_double), ("rho0", C.c_double), ("strip", C.c_uint32)]


# name -> (restype, argtypes); every symbol include/mfree_iwf_amd.h declares.
class SurrogateOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("seed", C.c_uint64), ("max_batch_nodes", C.c_uint32),
                ("max_compression", C.c_float), ("base_order", C.c_int32)]


class JudgeStats(C.Structure):
    _fields_ = [("groups", C.c_uint32), ("fallbacks", C.c_uint32), ("learned_steps", C.c_uint32),
                ("plain_steps", C.c_uint32), ("max_deviation", C.c_double), ("last_deviation", C.c_double),
                ("ms_learned", C.c_double), ("ms_judge", C.c_double), ("ms_fallback", C.c_double),
                ("max_ratio_vel", C.c_double), ("max_ratio_rho", C.c_double), ("max_ratio_stress", C.c_double),
                ("worst_body", C.c_uint32), ("numeric_fallbacks", C.c_uint32)]


SIGNATURES = {
    "mfree_abi_version": (C.c_int32, []),
    "mfree_status_string": (C.c_char_p, [C.c_int32]),
    "mfree_cubic_spline": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dptr]),
    "mfree_device_count": (C.c_int32, []),
    "mfree_default_options": (None, [C.POINTER(Options)]),
    "mfree_create": (C.c_int32, [C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "mfree_destroy": (None, [C.c_void_p]),
    "mfree_last_error": (C.c_char_p, [C.c_void_p]),
    "mfree_add_body": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(Particles), C.c_uint32,
                                   C.POINTER(C.c_uint32)]),
    "mfree_set_strip_bc": (C.c_int32, [C.c_void_p, C.c_uint32, u32ptr, C.c_uint32, u32ptr, C.c_uint32,
                                       C.c_double]),
    "mfree_tl_precompute": (C.c_int32, [C.c_void_p]),
    "mfree_set_state": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(Particles)]),
    "mfree_get_state": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(Particles)]),
    "mfree_get_time": (C.c_int32, [C.c_void_p, C.c_uint32, dptr]),
    "mfree_body_status": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(C.c_int32)]),
    "mfree_vtk_write": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "mfree_vtk_write_async": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "mfree_vtk_wait": (C.c_int32, [C.c_void_p]),
    "mfree_vtk_write_host": (C.c_int32, [C.c_char_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)]),
    "mfree_energy": (C.c_int32, [C.c_void_p, C.c_uint32, dptr]),
    "mfree_energy_noalloc": (C.c_int32, [C.c_int32, C.c_void_p, C.POINTER(Params), C.POINTER(Particles),
                                         C.c_uint32, C.c_void_p]),
    "mfree_step": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "mfree_eval_derivatives": (C.c_int32, [C.c_void_p, dptr, dptr]),
    "mfree_get_pairs": (C.c_int32, [C.c_void_p, C.c_uint32, u32ptr, u32ptr, C.c_uint64,
                                    C.POINTER(C.c_uint64)]),
    "mfree_profile_enable": (C.c_int32, [C.c_void_p, C.c_int32]),
    "mfree_profile_read": (C.c_int32, [C.c_void_p, dptr, C.POINTER(C.c_uint64), dptr,
                                       C.POINTER(C.c_uint64)]),
    "mfree_calibrate_copy": (C.c_int32, [C.c_int32, C.c_uint64, C.c_uint32, dptr]),
    "mfree_tile_cells": (C.c_int32, [C.c_void_p]),
    "mfree_cell_scale": (C.c_double, [C.c_void_p]),
    "mfree_num_particles": (C.c_uint64, [C.c_void_p]),
    "mfree_num_bodies": (C.c_uint32, [C.c_void_p]),
    "mfree_scene_rings_count": (C.c_uint32, [C.POINTER(RingsSpec)]),
    "mfree_scene_rings": (C.c_int32, [C.POINTER(RingsSpec), C.POINTER(Particles), C.c_uint32,
                                      C.POINTER(C.c_uint32)]),
    "mfree_scene_plate_count": (C.c_uint32, [C.POINTER(PlateSpec), u32ptr, u32ptr]),
    "mfree_scene_plate": (C.c_int32, [C.POINTER(PlateSpec), C.POINTER(Particles), C.c_uint32, u32ptr,
                                      u32ptr, C.POINTER(C.c_uint32)]),
    "mfree_surrogate_default_options": (None, [C.POINTER(SurrogateOptions)]),
    "mfree_surrogate_create": (C.c_int32, [C.POINTER(SurrogateOptions), C.POINTER(C.c_void_p)]),
    "mfree_surrogate_destroy": (None, [C.c_void_p]),
    "mfree_surrogate_last_error": (C.c_char_p, [C.c_void_p]),
    "mfree_surrogate_collect": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "mfree_surrogate_snapshots": (C.c_uint64, [C.c_void_p]),
    "mfree_surrogate_dataset_nodes": (C.c_uint64, [C.c_void_p]),
    "mfree_surrogate_train": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_float, fptr]),
    "mfree_surrogate_record_velocity": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "mfree_surrogate_step": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32]),
    "mfree_surrogate_rollout": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_double,
                                            C.POINTER(JudgeStats)]),
    "mfree_mpnn_param_count": (C.c_uint64, []),
    "mfree_mpnn_layout": (C.c_uint32, [i32ptr]),
    "mfree_surrogate_get_params": (C.c_int32, [C.c_void_p, fptr, fptr]),
    "mfree_surrogate_set_params": (C.c_int32, [C.c_void_p, fptr, fptr]),
    "mfree_surrogate_graph": (C.c_int32, [C.c_void_p, C.c_void_p, u32ptr, u32ptr, fptr, C.c_uint64,
                                          C.POINTER(C.c_uint64)]),
    "mfree_mpnn_forward_backward": (C.c_int32, [C.c_int32, C.c_uint32, u32ptr, u32ptr, fptr, fptr, fptr, fptr,
                                                fptr, fptr, fptr]),
    "mfree_set_boundary": (C.c_int32, [C.c_void_p, C.c_uint32, u32ptr, C.c_uint32]),
    "mfree_update_boundaries": (C.c_int32, [C.c_void_p]),
    "mfree_get_boundary": (C.c_int32, [C.c_void_p, C.c_uint32, dptr, C.c_uint32, C.POINTER(C.c_uint32)]),
    "mfree_contact": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    "mfree_contact_all": (C.c_int32, [C.c_void_p]),
    "mfree_convex_hull": (C.c_uint32, [dptr, dptr, C.c_uint32, u32ptr]),
}

_lib = None


def lib():
    """Loads the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C mfree_iwf_amd/csrc` "
                "(the library has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MfreeError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def check(status, ctx=None, what=""):
    if status != OK:
        detail = ""
        if ctx:
            detail = lib().mfree_last_error(ctx).decode()
        if not detail:
            detail = lib().mfree_status_string(status).decode()
        raise MfreeError(status, f"{what}: {detail}" if what else detail)
