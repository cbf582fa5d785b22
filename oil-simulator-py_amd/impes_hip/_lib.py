"""ctypes binding of libimpes_hip.so (declarations mirror include/impes_hip.h).

The library is loaded after ``import torch`` so that it binds to the HIP runtime
torch already loaded (both carry the SONAME ``libamdhip64.so.7``): one runtime,
one set of streams, device pointers shared freely.  There is no fallback: if the
shared library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime first)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libimpes_hip.so")

_lock = threading.Lock()
_lib = None


class ImpesError(RuntimeError):
    """Error reported by the HIP library (negative status code)."""


class AmgParams(C.Structure):
    _fields_ = [
        ("theta", C.c_double),
        ("max_levels", C.c_int),
        ("coarsest_size", C.c_int),
        ("pre_smooth", C.c_int),
        ("post_smooth", C.c_int),
        ("omega", C.c_double),
        ("equil_threshold", C.c_double),
        ("dense_max", C.c_int),
        ("coarse_sweeps", C.c_int),
        ("anchor", C.c_int),
        ("equil_mode", C.c_int),
        ("coarse_pin", C.c_int),
        ("trunc", C.c_double),
        ("pmax", C.c_int),
        ("coarsening", C.c_int),
        ("dominant", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("restarts", C.c_int),
        ("converged", C.c_int),
        ("precision", C.c_int),
        ("rel_residual", C.c_double),
        ("bnorm", C.c_double),
        ("r0_rel", C.c_double),
        ("checks", C.c_int),
        ("err_estimate", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class FluidStruct(C.Structure):
    _fields_ = [
        ("stone2", C.c_int),
        ("nw", C.c_double), ("no", C.c_double), ("ng", C.c_double), ("ko_end_w", C.c_double),
        ("ko_end_g", C.c_double), ("now_", C.c_double), ("nog", C.c_double),
        ("sw_cr", C.c_double), ("so_r", C.c_double), ("sg_cr", C.c_double),
        ("mu_o", C.c_double), ("mu_w", C.c_double), ("mu_g", C.c_double),
        ("rho_o_ref", C.c_double), ("rho_w_ref", C.c_double), ("rho_g_ref", C.c_double),
        ("c_o", C.c_double), ("c_w", C.c_double), ("c_g", C.c_double), ("c_r", C.c_double), ("cf", C.c_double),
        ("pc_scale", C.c_double), ("pc_exp", C.c_double), ("pcog_scale", C.c_double), ("pcog_exp", C.c_double),
        ("g", C.c_double),
        ("max_ds", C.c_double),
        ("max_substeps", C.c_int),
        ("rate_scale", C.c_double),
    ]


class WellStruct(C.Structure):
    _fields_ = [
        ("cell", C.c_int), ("producer", C.c_int), ("control", C.c_int), ("gas", C.c_int), ("surface", C.c_int),
        ("surface_phase", C.c_int), ("has_bhp_min", C.c_int), ("lim_mask", C.c_int),
        ("value", C.c_double), ("wi", C.c_double), ("bhp_min", C.c_double), ("lim", C.c_double * 4),
    ]


_P = C.c_void_p
_I = C.c_int
_L = C.c_longlong
_D = C.c_double
_F = C.c_float
_PI = C.POINTER(C.c_int)
_PD = C.POINTER(C.c_double)

# name -> argtypes ; every function returns int unless listed in _RESTYPE
SIGNATURES = {
    "impes_abi_version": [],
    "impes_last_error": [],
    "impes_spmv_csr_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P],
    "impes_spmv_csr_f32": [_I, _I, _I, _P, _P, _P, _P, _P, _P],
    "impes_spmv_st7_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P],
    "impes_l1jacobi_csr_f64": [_I, _I, _P, _P, _P, _P, _D, _P, _P, _P, _P],
    "impes_l1jacobi_csr_f32": [_I, _I, _P, _P, _P, _P, _F, _P, _P, _P, _P],
    "impes_dot_f64": [_L, _P, _P, _P, _P],
    "impes_copy_f64": [_L, _P, _P, _P],
    "impes_csr_diagonal": [_I, _P, _P, _P, _P, _P],
    "impes_amg_default_params": [C.POINTER(AmgParams)],
    "impes_amg_create": [C.POINTER(AmgParams), C.POINTER(_P)],
    "impes_amg_destroy": [_P],
    "impes_amg_setup_csr": [_P, _I, _I, _P, _P, _P, _P],
    "impes_amg_setup_st7": [_P, _I, _I, _I, _P, _P, _P, _P, _P],
    "impes_amg_update_csr": [_P, _P, _P],
    "impes_amg_update_st7": [_P, _P, _P, _P, _P, _P],
    "impes_amg_num_levels": [_P, C.POINTER(_I)],
    "impes_amg_level_info": [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)],
    "impes_amg_equilibrated": [_P, C.POINTER(_I)],
    "impes_amg_level_times": [_P, _I, _I, _PD, _PD, _I, _PD, _P],
    "impes_amg_setup_seconds": [_P, C.POINTER(_D)],
    "impes_amg_export_level": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "impes_amg_export_scaling": [_P, _P, _P],
    "impes_amg_vcycle": [_P, _I, _P, _P, _I, _P],
    "impes_amg_vcycle_transposed": [_P, _I, _P, _P, _I, _P],
    "impes_amg_precondition": [_P, _I, _P, _P, _P],
    "impes_amg_set_graphs": [_P, _I],
    "impes_solver_create": [C.POINTER(_P)],
    "impes_solver_destroy": [_P],
    "impes_solver_set_error_tol": [_P, _D],
    "impes_fgmres_csr": [_P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _D, _I, _I, C.POINTER(SolveStats), _P],
    "impes_fgmres_st7": [_P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _I, _I,
                         C.POINTER(SolveStats), _P],
    "impes_amg_cycles": [_P, _I, _P, _P, _I, _I, _D, _I, _PD, _PI, _P],
    "impes_amg_cycles_noalloc": [_P, _I, _P, _P, _I, _I, _D, _I, _PD, _PI, _P],
    "impes_model_create": [C.POINTER(Grid), C.POINTER(FluidStruct), _I, _PD, _I, C.POINTER(WellStruct),
                           C.POINTER(_P)],
    "impes_model_destroy": [_P],
    "impes_model_set_rock": [_P, _P, _P, _P, _P, _P],
    "impes_model_set_wells": [_P, _I, C.POINTER(WellStruct), _P],
    "impes_model_set_envs": [_P, _I, _P],
    "impes_model_set_amg": [_P, C.POINTER(AmgParams)],
    "impes_model_report_len": [_P, _PI],
    "impes_model_assemble": [_P, _D, _P, _P, _P, _P],
    "impes_model_solve_pressure": [_P, _I, _D, _I, _I, _P, _P, C.POINTER(SolveStats), _P],
    "impes_model_set_error_tol": [_P, _D],
    "impes_model_solve_cg": [_P, _D, _I, _P, _PI, _PI, _P],
    "impes_model_saturation": [_P, _D, _P, _P, _P, _P, _P, _P, _P, _PD, _P],
    "impes_model_export": [_P, _I, _P, _P],
    "impes_model_adjoint_reset": [_P, _P],
    "impes_model_transport_vjp": [_P, _D, _P, _P, _P, _P, _P, _P, _PD, _P, _P, _P, _P, _P],
    "impes_model_solve_adjoint": [_P, _I, _D, _I, _I, _P, _P, C.POINTER(SolveStats), _P],
    "impes_model_assemble_vjp": [_P, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "impes_model_rock_vjp": [_P, _P, _P, _P, _P, _P, _P, _P],
    "impes_model_adjoint_values": [_P, _PD, _P],
    "impes_model_eval_props": [_P, _L, _P, _P, _P, _P, _P],
    "impes_model_eval_pvt": [_P, _I, _L, _P, _P, _P],
    "impes_model_amg_levels": [_P, _PI, _PI, _PI, _I],
    "impes_model_amg_setup_seconds": [_P, _PD],
    "impes_model_amg_level_times": [_P, _I, _I, _PD, _PD, _I, _PD, _P],
    "impes_ktimer_config": [C.c_char_p, _I],
    "impes_ktimer_read": [_PD, _PD, C.POINTER(_L)],
}
_RESTYPE = {"impes_last_error": C.c_char_p, "impes_amg_default_params": None}


def load():
    """Load and declare the library (idempotent). Raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImpesError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python oil-simulator-py_amd/csrc/build.py`")
        lib = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = lib
        return lib


def exported_symbols() -> list[str]:
    return list(SIGNATURES)


def call(name: str, *args):
    """Call an int-returning entry point; raise ImpesError on a negative status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.impes_last_error()
        raise ImpesError(f"{name} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def ptr(t) -> C.c_void_p | None:
    """Device pointer of a CUDA tensor (None for None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ImpesError("IMPES HIP kernels need tensors on a ROCm (cuda) device")
    if not t.is_contiguous():
        raise ImpesError("IMPES HIP kernels need contiguous tensors")
    return C.c_void_p(t.data_ptr())


def stream_ptr(stream=None) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
