"""ctypes declarations of the C ABI in include/ariel.h.

The shared library ``ariel_amd/lib/libariel_amd.so`` is the product: hand-written
HIP kernels for gfx950 behind the C ABI. Loading fails loudly when it is missing;
there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libariel_amd.so")

ARIEL_OK = 0
(ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_CAPACITY, ERR_UNSUPPORTED,
 ERR_IO) = -1, -2, -3, -4, -5, -6, -7
AIR, FLUID, SOLID = 0, 1, 2
LIQUID_BOX, LIQUID_SPHERE = 0, 1
PRECOND_MIC0_TILED, PRECOND_MIC0, PRECOND_JACOBI, PRECOND_REFERENCE, PRECOND_MIC0_BLOCK, \
    PRECOND_LEARNED, PRECOND_MG = range(7)
STOP_RELATIVE, STOP_REFERENCE_ABS = 0, 1
SPLAT_SHARPEN, SPLAT_TRILINEAR = 0, 1
P2G_ORDERED, P2G_TILED = 0, 1
LEVELSET_SPHERES, LEVELSET_TRAILS = 0, 1
FIELD_UX, FIELD_UY, FIELD_UZ, FIELD_DUX, FIELD_DUY, FIELD_DUZ, FIELD_P, FIELD_D, FIELD_L = range(9)
(STAGE_SORT, STAGE_DENSITY, STAGE_MARK, STAGE_P2G, STAGE_DIVERGENCE, STAGE_SOLVE,
 STAGE_GRADIENT, STAGE_EXTRAPOLATE, STAGE_ADVECT, STAGE_RESAMPLE) = range(10)
NUM_STAGES = 10
STAGE_NAMES = ["sort", "density", "mark", "p2g", "divergence", "solve", "gradient",
               "extrapolate", "advect", "resample"]
SOLVE_CONVERGED, SOLVE_MAX_ITERS, SOLVE_BREAKDOWN, SOLVE_NONFINITE = range(4)
MAX_FORCES, MAX_LIQUIDS = 4, 8
(KERNEL_SORT, KERNEL_DENSITY, KERNEL_MARK, KERNEL_P2G, KERNEL_SPMV, KERNEL_PRECOND, KERNEL_AXPY,
 KERNEL_GRADIENT, KERNEL_EXTRAPOLATE, KERNEL_ADVECT, KERNEL_RESAMPLE) = range(11)
NUM_KERNELS = 11
KERNEL_NAMES = ["sort", "density", "mark", "p2g", "spmv", "precond", "axpy", "gradient",
                "extrapolate", "advect", "resample"]


class Liquid(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
                ("center", ctypes.c_float * 3), ("radius", ctypes.c_float),
                ("velocity", ctypes.c_float * 3)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("density", ctypes.c_float), ("dt", ctypes.c_float),
                ("picflip", ctypes.c_float), ("n_forces", ctypes.c_int32),
                ("forces", (ctypes.c_float * 3) * MAX_FORCES), ("n_liquids", ctypes.c_int32),
                ("liquids", Liquid * MAX_LIQUIDS), ("wall", ctypes.c_int32),
                ("seed", ctypes.c_uint32), ("jitter", ctypes.c_int32),
                ("precond", ctypes.c_int32), ("mic_tau", ctypes.c_float),
                ("mic_sigma", ctypes.c_float), ("tile", ctypes.c_int32), ("tol", ctypes.c_float),
                ("max_iters", ctypes.c_int32), ("refine_every", ctypes.c_int32),
                ("stop_rule", ctypes.c_int32), ("splat", ctypes.c_int32),
                ("resample", ctypes.c_int32), ("sharpen_re", ctypes.c_float),
                ("spring", ctypes.c_float), ("p2g", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 7)]


class SolverTrial(ctypes.Structure):
    _fields_ = [("precond", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("status", ctypes.c_int32), ("solve_ms", ctypes.c_float)]


class StepStats(ctypes.Structure):
    _fields_ = [("frame", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("solve_status", ctypes.c_int32), ("fluid_cells", ctypes.c_int32),
                ("particles", ctypes.c_int64), ("rhs_norm", ctypes.c_double),
                ("residual_norm", ctypes.c_double), ("true_residual_norm", ctypes.c_double),
                ("max_divergence", ctypes.c_double), ("max_density", ctypes.c_double),
                ("stage_ms", ctypes.c_double * NUM_STAGES), ("step_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("precond_fallback", ctypes.c_int32),
                ("fallback_iter", ctypes.c_int32), ("true_residual_norm64", ctypes.c_double)]


class KernelStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64), ("total_ms", ctypes.c_double),
                ("bytes_per_launch", ctypes.c_double), ("calls", ctypes.c_int64)]


_P = ctypes.c_void_p
_PF = ctypes.POINTER(ctypes.c_float)
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

# name -> (restype, argtypes); every entry point declared in include/ariel.h.
PROTOTYPES = {
    "ariel_desc_defaults": (_I32, [ctypes.POINTER(SceneDesc)]),
    "ariel_create": (_I32, [ctypes.POINTER(SceneDesc), ctypes.POINTER(_P)]),
    "ariel_destroy": (_I32, [_P]),
    "ariel_last_error": (ctypes.c_char_p, [_P]),
    "ariel_init": (_I32, [_P]),
    "ariel_step": (_I32, [_P, _I32]),
    "ariel_step_autotune": (_I32, [_P, _P, _I32, _P, _P]),
    "ariel_run_stage": (_I32, [_P, _I32]),
    "ariel_sync": (_I32, [_P]),
    "ariel_get_frame": (_I32, [_P, ctypes.POINTER(_I32)]),
    "ariel_get_dimensions": (_I32, [_P, _I32 * 3]),
    "ariel_particle_count": (_I32, [_P, ctypes.POINTER(_I64)]),
    "ariel_get_particles": (_I32, [_P, _P, _P, _P, _P, _I64, ctypes.POINTER(_I64)]),
    "ariel_set_particles": (_I32, [_P, _P, _P, _P, _I64]),
    "ariel_get_macgrid": (_I32, [_P, _I32, _P, _I64]),
    "ariel_set_macgrid": (_I32, [_P, _I32, _P, _I64]),
    "ariel_get_flags": (_I32, [_P, _P, _I64]),
    "ariel_set_flags": (_I32, [_P, _P, _I64]),
    "ariel_get_sort": (_I32, [_P, _P, _I64, _P, _I64]),
    "ariel_get_resample_accept": (_I32, [_P, _P, _I64]),
    "ariel_get_max_density": (_I32, [_P, _PF]),
    "ariel_set_max_density": (_I32, [_P, ctypes.c_float]),
    "ariel_get_stats": (_I32, [_P, ctypes.POINTER(StepStats)]),
    "ariel_apply_laplacian": (_I32, [_P, _P, _P]),
    "ariel_apply_preconditioner": (_I32, [_P, _P, _P]),
    "ariel_set_profiling": (_I32, [_P, _I32]),
    "ariel_reset_kernel_stats": (_I32, [_P]),
    "ariel_get_kernel_stats": (_I32, [_P, _I32, ctypes.POINTER(KernelStats)]),
    "ariel_export_levelset_noalloc": (_I32, [_P, _P, _I64]),
    "ariel_save_state": (_I32, [_P, ctypes.c_char_p]),
    "ariel_load_state": (_I32, [_P, ctypes.c_char_p]),
    "ariel_export_levelset": (_I32, [_P, ctypes.c_char_p]),
    "ariel_export_mesh": (_I32, [_P, ctypes.c_char_p]),
    "ariel_export_particles": (_I32, [_P, ctypes.c_char_p]),
    "ariel_set_levelset_mode": (_I32, [_P, _I32]),
    "ariel_create_from_json": (_I32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(SceneDesc),
                                      ctypes.POINTER(_P)]),
    "ariel_scene_emit": (_I32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(SceneDesc), _I32,
                                _PF, _PF, _I64, ctypes.POINTER(ctypes.c_int64)]),
    "ariel_get_desc": (_I32, [_P, ctypes.POINTER(SceneDesc)]),
    "ariel_write_vdb": (_I32, [_PF, _I32, ctypes.c_float, ctypes.c_char_p]),
    "ariel_write_particles": (_I32, [ctypes.c_char_p, _I64, _PF, _PF]),
    "ariel_mesh_sdf": (_I32, [_PF, _I32, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                              ctypes.POINTER(ctypes.c_int64)]),
    "ariel_selftest_bf16": (_I32, [ctypes.POINTER(ctypes.c_int64)]),
    "ariel_selftest_division": (_I32, [ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.POINTER(ctypes.c_int64)]),
    "ariel_set_learned_weights": (_I32, [_P, ctypes.c_void_p, ctypes.c_int64]),
    "ariel_solve_rhs": (_I32, [_P, _PF, _PF, ctypes.POINTER(StepStats)]),
    "ariel_batch_create": (_I32, [ctypes.POINTER(SceneDesc), _I32, ctypes.POINTER(_P)]),
    "ariel_batch_destroy": (_I32, [_P]),
    "ariel_batch_last_error": (ctypes.c_char_p, [_P]),
    "ariel_desc_from_json": (_I32, [ctypes.c_char_p, ctypes.POINTER(SceneDesc)]),
    "ariel_trainer_create": (_I32, [ctypes.POINTER(_P)]),
    "ariel_trainer_destroy": (_I32, [_P]),
    "ariel_trainer_last_error": (ctypes.c_char_p, [_P]),
    "ariel_trainer_set_weights": (_I32, [_P, _PF, ctypes.c_int64]),
    "ariel_trainer_get_weights": (_I32, [_P, _PF, ctypes.c_int64]),
    "ariel_trainer_step": (_I32, [_P, _I32, _I32, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float,
                                  _I32, ctypes.POINTER(ctypes.c_double), _PF]),
    "ariel_trainer_forward": (_I32, [_P, _I32, _I32, _P, _P, _P, _P]),
    "ariel_batch_size": (_I32, [_P, ctypes.POINTER(_I32)]),
    "ariel_batch_tank": (_I32, [_P, _I32, ctypes.POINTER(_P)]),
    "ariel_batch_init": (_I32, [_P]),
    "ariel_batch_autotune": (_I32, [_P]),
    "ariel_batch_get_tuning": (_I32, [_P, _I32, ctypes.POINTER(_I32)]),
    "ariel_batch_set_graphs": (_I32, [_P, _I32]),
    "ariel_batch_step": (_I32, [_P, _I32]),
    "ariel_batch_step_sequential": (_I32, [_P, _I32]),
}

_lib = None


def load():
    """Load libariel_amd.so and bind every prototype. Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m ariel_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
