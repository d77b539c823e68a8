"""ctypes binding of libhelm.so (the C ABI declared in include/helm.h).

The library is built in-tree by ``python -m nsrno_amd.build``.  There is no
CPU fallback: if the library is missing every entry point raises.
"""
import ctypes
from ctypes import POINTER, c_double, c_int32, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhelm.so"

HELM_CONVERGED, HELM_ITERATION_BUDGET, HELM_BREAKDOWN, HELM_STAGNATION, HELM_DIVERGED = 0, 1, 2, 3, 4
REASONS = {HELM_CONVERGED: "converged", HELM_ITERATION_BUDGET: "iteration_budget",
           HELM_BREAKDOWN: "breakdown", HELM_STAGNATION: "stagnation", HELM_DIVERGED: "diverged"}
HELM_GMRES, HELM_BICGSTAB = 0, 1
PC_NONE, PC_MG, PC_LEARNED, PC_SPECTRAL, PC_SPECTRAL_PLAIN = 0, 1, 2, 3, 4


class HelmDesc(ctypes.Structure):
    _fields_ = [("nz", c_int32), ("nx", c_int32), ("nrhs", c_int32), ("restart", c_int32),
                ("x0_mm", c_double), ("z0_mm", c_double), ("h_mm", c_double), ("pml_mm", c_double),
                ("frequency_mhz", c_double), ("c0_m_s", c_double), ("pml_strength", c_double),
                ("op_kind", c_int32)]


HELM_OP_FLUX5, HELM_OP_STAG8, HELM_OP_COND8 = 0, 1, 2


class HelmSpectralMedium(ctypes.Structure):
    _fields_ = [("h", c_int32), ("w", c_int32), ("pz", c_int32), ("px", c_int32),
                ("x0_m", c_double), ("z0_m", c_double), ("rho0", c_double), ("alpha_scale", c_double),
                ("relaxation", c_int32), ("sponge_cells", c_int32), ("sponge_amplitude", c_double),
                ("eta_per_unit_alpha", c_double * 2), ("taus_s", c_double * 2), ("relax_dt_s", c_double)]


class HelmConditionedMedium(ctypes.Structure):
    _fields_ = [("h", c_int32), ("w", c_int32), ("pz", c_int32), ("px", c_int32), ("pml", c_int32),
                ("dt_s", c_double), ("rho0", c_double), ("eta_per_unit_alpha", c_double * 2),
                ("taus_s", c_double * 2)]


CALLBACK = ctypes.CFUNCTYPE(None, c_void_p, c_int32, POINTER(c_double), c_int32)


class HelmSolveOpts(ctypes.Structure):
    _fields_ = [("method", c_int32), ("precond", c_int32), ("restart", c_int32), ("max_iterations", c_int32),
                ("tolerance", c_double), ("adjoint", c_int32), ("use_x0", c_int32),
                ("callback", CALLBACK), ("user", c_void_p), ("use_graph", c_int32), ("fallback", c_int32),
                ("stagnation_steps", c_int32), ("stagnation_ratio", c_double)]


class HelmSolveInfo(ctypes.Structure):
    _fields_ = [("reason", c_int32), ("iterations", c_int32), ("applications", c_int32), ("cycles", c_int32),
                ("relative_residual", c_double), ("fallback", c_int32), ("first_reason", c_int32),
                ("graph_fallbacks", c_int32)]


class HelmTuning(ctypes.Structure):
    _fields_ = [("stencil_rows", c_int32), ("stencil_rhs", c_int32), ("mg_workgroups", ctypes.c_int64),
                ("stencil_us", c_double), ("vcycle_us", c_double), ("stencil_block", c_int32)]


class HelmKernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64), ("total_ms", c_double),
                ("bytes", c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "helm_version": (ctypes.c_char_p, []),
    "helm_error_string": (ctypes.c_char_p, [c_int32]),
    "helm_last_error": (ctypes.c_char_p, []),
    "helm_plan_create": (c_int32, [POINTER(HelmDesc), POINTER(c_void_p)]),
    "helm_plan_destroy": (c_int32, [c_void_p]),
    "helm_set_medium": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "helm_set_medium_physical": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "helm_apply": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p]),
    "helm_inject": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_int32, c_void_p]),
    "helm_sample": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                              c_int32, c_void_p]),
    "helm_solve": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, POINTER(HelmSolveOpts),
                             POINTER(HelmSolveInfo), c_void_p]),
    "helm_batch_solve": (c_int32, [c_int32, POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p), c_int32,
                                   POINTER(HelmSolveOpts), POINTER(HelmSolveInfo), c_void_p]),
    "helm_krylov_create": (c_int32, [ctypes.c_int64, c_int32, c_int32, POINTER(c_void_p)]),
    "helm_krylov_destroy": (c_int32, [c_void_p]),
    "helm_krylov_begin": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int32, POINTER(c_double),
                                    POINTER(c_int32), c_void_p]),
    "helm_krylov_step": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_double, c_void_p]),
    "helm_krylov_end": (c_int32, [c_void_p, c_void_p, c_void_p, POINTER(c_int32), c_void_p]),
    "helm_born_jacobian": (c_int32, [ctypes.c_int64, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_int32,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "helm_capture_accumulate": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32, ctypes.c_int64,
                                          c_void_p, c_void_p]),
    "helm_band_rf": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    "helm_spectral_set_preconditioner": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "helm_fft2": (c_int32, [c_int32, c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "helm_mg_setup": (c_int32, [c_void_p, c_int32, c_double, c_int32, c_int32, c_int32, c_double, c_void_p]),
    "helm_precond_apply": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p]),
    "helm_precond_apply_ex": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_void_p]),
    "helm_material_sensitivity": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "helm_net_create": (c_int32, [c_int32, c_int32, c_int32, c_int32, POINTER(c_void_p)]),
    "helm_net_destroy": (c_int32, [c_void_p]),
    "helm_net_param_count": (c_int32, []),
    "helm_net_set_linear": (c_int32, [c_void_p, c_int32]),
    "helm_net_set_params": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "helm_net_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "helm_net_backward": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "helm_adam_step": (c_int32, [ctypes.c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double,
                                 c_double, c_double, c_double, c_double, c_void_p]),
    "helm_plan_set_net": (c_int32, [c_void_p, c_void_p]),
    "helm_net_correction": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "helm_net_features": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "helm_set_medium_spectral": (c_int32, [c_void_p, c_void_p, c_void_p, POINTER(HelmSpectralMedium), c_void_p]),
    "helm_spectral_incident": (c_int32, [c_void_p, c_double, c_double, c_void_p, c_void_p]),
    "helm_spectral_incident_noalloc": (c_int32, [c_void_p, c_double, c_double, c_void_p, c_void_p]),
    "helm_spectral_rhs": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "helm_autotune": (c_int32, [c_void_p, c_int32, POINTER(HelmTuning), c_void_p]),
    "helm_set_stencil_shape": (c_int32, [c_void_p, c_int32, c_int32, c_int32]),
    "helm_profile": (c_int32, [c_void_p, c_int32]),
    "helm_profile_read": (c_int32, [c_void_p, POINTER(HelmKernelStat), c_int32, POINTER(c_int32)]),
    "helm_warm_start": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p]),
    "helm_set_medium_conditioned": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p,
                                              POINTER(HelmConditionedMedium), c_void_p]),
    "helm_conditioned_source": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                          c_void_p, c_int32, c_void_p]),
    "helm_conditioned_gradient": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p]),
}

_LIB = None


def lib():
    """Load libhelm.so once; raise if it is missing (no fallback path)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m nsrno_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


class HelmError(RuntimeError):
    pass


def check(status, what):
    if status != 0:
        l = lib()
        raise HelmError(f"{what}: {l.helm_error_string(status).decode()} ({l.helm_last_error().decode()})")
