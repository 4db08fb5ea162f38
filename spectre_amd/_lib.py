"""Loader of the native library libspectre_amd.so (HIP kernels + C ABI).

The product path runs only through this library: if it is missing or fails to
load, every entry point raises; there is no CPU fallback.
"""
import ctypes
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libspectre_amd.so"

_lib = None


class SpectreAmdError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, status, message):
        super().__init__(f"spectre_amd status {status}: {message}")
        self.status = status


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64

_SIGNATURES = {
    "sxd_last_error": (ctypes.c_char_p, []),
    "sxd_version": (ctypes.c_char_p, []),
    "sxd_lgl_points_and_weights": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
    "sxd_lgl_differentiation_matrix": (ctypes.c_int, [ctypes.c_int, _dp]),
    "sxd_lg_points_and_weights": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
    "sxd_lg_differentiation_matrix": (ctypes.c_int, [ctypes.c_int, _dp]),
    "sxd_lg_boundary_terms": (ctypes.c_int, [ctypes.c_int, _dp]),
    "sxd_scalar_wave_rhs_gauss": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_logical_partial_derivatives": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _vp, _vp, _vp]),
    "sxd_partial_derivatives": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "sxd_partial_derivatives_adjoint": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _vp, _vp,
                                                       _vp, _vp]),
    "sxd_divergence": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "sxd_weak_divergence": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "sxd_flux_volume_terms": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _vp, _vp, _vp, _vp,
                                             ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "sxd_mesh_velocity_nonconservative": (ctypes.c_int, [_i64, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "sxd_scalar_wave_time_derivative": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_gh_time_derivative": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_scalar_wave_upwind_penalty": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_gh_upwind_penalty": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_lift_flux": (ctypes.c_int, [_i64, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]),
    "sxd_domain_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _ip, _ip]),
    "sxd_domain_destroy": (ctypes.c_int, [_vp]),
    "sxd_scalar_wave_rhs": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_scalar_wave_rhs_dirichlet": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_scalar_wave_rhs_spherical_radiation": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp,
                                                                ctypes.c_int, _vp, _vp]),
    "sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation": (
        ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "sxd_gh_rhs": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_gh_rhs_dirichlet": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_gh_rhs_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_gh_bjorhus_corrections": (ctypes.c_int, [ctypes.c_int, _i64] + [_vp] * 23),
    "sxd_gh_rhs_bjorhus": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_rk3_substep": (ctypes.c_int, [ctypes.c_int, _i64, ctypes.c_double, _vp, _vp, _vp, _vp, _vp]),
    "sxd_domain_create_mixed": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _ip, _ip]),
    "sxd_domain_create_mortars": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _ip,
                                                 ctypes.c_int, _ip]),
    "sxd_domain_create_mortars_partition": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _ip,
                                                           ctypes.c_int, _ip, _ip, ctypes.c_int]),
    "sxd_boundary_messages_layout": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip,
                                                    ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                                    ctypes.POINTER(_i64)]),
    "sxd_boundary_messages_expected": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "sxd_rhs_begin": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "sxd_boundary_messages_unpack": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _i64,
                                                    _vp]),
    "sxd_rhs_end": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "sxd_projection_matrix_parent_to_child": (ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                                             ctypes.c_int, _dp]),
    "sxd_projection_matrix_child_to_parent": (ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                                             ctypes.c_int, _dp]),
    "sxd_domain_num_points": (ctypes.c_int, [_vp, ctypes.POINTER(_i64)]),
    "sxd_domain_set_threads": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "sxd_domain_get_threads": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "sxd_domain_set_kernel": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "sxd_scalar_wave_rhs_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_gh_rhs_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sxd_rk3_adjoint_substep": (ctypes.c_int, [ctypes.c_int, _i64, ctypes.c_double, _vp, _vp, _vp,
                                               _vp, _vp]),
    "sxd_rk3_stepper_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, _vp, _vp,
                                              _vp]),
    "sxd_rk3_stepper_step": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _vp]),
    "sxd_rk3_stepper_destroy": (ctypes.c_int, [_vp]),
    "sxd_domain_get_kernel": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "sxd_domain_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp,
                                           _ip, ctypes.POINTER(ctypes.c_float)]),
}

# Kernel organisations of the whole-domain right-hand sides (SXD_KERNEL_*)
KERNELS = {"auto": 0, "composite": 1, "fused": 2, "fused_mfma": 3, "mixed_batched": 4,
           "mixed_per_extents": 5, "mortar": 6}
SYSTEMS = {"scalar_wave": 0, "gh": 1}


def exported_symbols():
    """Names of the C-ABI functions this binding expects the library to export."""
    return sorted(_SIGNATURES)


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SpectreAmdError(-2, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != 0:
        raise SpectreAmdError(status, lib().sxd_last_error().decode())
    return status
