"""ctypes binding of libpascal_tcs.so (the C-ABI declared in include/pascal_tcs.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the shared library is missing or no HIP device is present the
calls raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpascal_tcs.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_f64 = ctypes.c_double


class TCSError(RuntimeError):
    pass


class tcs_params(ctypes.Structure):
    _fields_ = (
        [("n1m", _i32), ("n2m", _i32), ("n3m", _i32),
         ("pbc1", _i32), ("pbc2", _i32), ("pbc3", _i32),
         ("precision", _i32), ("tdma_split", _i32)]
        + [(n, _f64) for n in ("Ct", "Cmu", "Cmt", "Cmp", "DeltaT")]
        + [(f"{c}1{d}", _f64) for c in "abcd" for d in range(6)]
        + [(n, _f64) for n in ("KPP0", "Rho0", "Cp0", "Mu0", "a12pera11", "L1", "L2", "L3", "MaxCFL", "dtMax")]
    )


FIELD_IDS = {name: i for i, name in enumerate(
    ["T", "U", "V", "W", "P", "dU", "dV", "dW", "dP", "dPhat", "dP0", "invRho", "Mu",
     "KPP", "dKPP", "invRhoCp", "dinvRhoCp"])}
WALL_IDS = {name: i for i, name in enumerate(
    ["TBCup", "TBCbt", "UBCup", "UBCbt", "VBCup", "VBCbt", "WBCup", "WBCbt"])}

# name -> (restype, argtypes); every name is declared in include/pascal_tcs.h
SIGNATURES = {
    "tcs_last_error": (ctypes.c_char_p, []),
    "tcs_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "tcs_mesh_axis": (ctypes.c_int, [_i32, _f64, _i32, _f64, _i32, _dp, _dp, _dp]),
    "tcs_init": (ctypes.c_int, [ctypes.POINTER(tcs_params)] + [_dp] * 6 + [ctypes.POINTER(_vp)]),
    "tcs_finalize": (ctypes.c_int, [_vp]),
    "tcs_sync": (ctypes.c_int, [_vp]),
    "tcs_stream": (_vp, [_vp]),
    "tcs_field_upload": (ctypes.c_int, [_vp, _i32, _dp]),
    "tcs_field_download": (ctypes.c_int, [_vp, _i32, _dp]),
    "tcs_field_device_ptr": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_dp)]),
    "tcs_field_pitch": (ctypes.c_int, [_vp, ctypes.POINTER(_i32)]),
    "tcs_fields_touched": (ctypes.c_int, [_vp]),
    "tcs_field_copy_device": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "tcs_wall_upload": (ctypes.c_int, [_vp, _i32, _dp]),
    "tcs_wall_download": (ctypes.c_int, [_vp, _i32, _dp]),
    "tcs_thermal_boundary": (ctypes.c_int, [_vp, _f64, _f64]),
    "tcs_momentum_boundary": (ctypes.c_int, [_vp] + [_f64] * 6),
    "tcs_thermal_coeffi": (ctypes.c_int, [_vp]),
    "tcs_thermal_solver": (ctypes.c_int, [_vp, _f64]),
    "tcs_momentum_coeffi": (ctypes.c_int, [_vp]),
    "tcs_momentum_solvedU": (ctypes.c_int, [_vp, _f64]),
    "tcs_momentum_solvedV": (ctypes.c_int, [_vp, _f64]),
    "tcs_momentum_solvedW": (ctypes.c_int, [_vp, _f64]),
    "tcs_momentum_blockLdV": (ctypes.c_int, [_vp, _f64]),
    "tcs_momentum_blockLdU": (ctypes.c_int, [_vp, _f64]),
    "tcs_momentum_pseudoupdateUVW": (ctypes.c_int, [_vp]),
    "tcs_pressure_wave_number": (ctypes.c_int, [_vp]),
    "tcs_pressure_rhs": (ctypes.c_int, [_vp, _f64]),
    "tcs_pressure_poisson_fft2": (ctypes.c_int, [_vp]),
    "tcs_pressure_projection": (ctypes.c_int, [_vp, _f64]),
    "tcs_ghostcell_update": (ctypes.c_int, [_vp, _i32]),
    "tcs_post_div": (ctypes.c_int, [_vp, _dp]),
    "tcs_post_cfl": (ctypes.c_int, [_vp, _dp, _dp]),
    "tcs_step": (ctypes.c_int, [_vp, _f64]),
    "tcs_nusselt": (ctypes.c_int, [_vp, _f64, _f64, _dp, _dp]),
    "tcs_profile": (ctypes.c_int, [_vp, _i32]),
    "tcs_profile_count": (ctypes.c_int, [_vp, ctypes.POINTER(_i32)]),
    "tcs_profile_entry": (ctypes.c_int, [_vp, _i32, ctypes.c_char_p, _i32, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "tcs_adjoint_zero": (ctypes.c_int, [_vp]),
    "tcs_adjoint_step": (ctypes.c_int, [_vp, _f64]),
    "tcs_adjoint_copy": (ctypes.c_int, [_vp, _i32, _i32, _i32, _vp]),
    "tcs_adjoint_add_nusselt": (ctypes.c_int, [_vp, _f64, _f64, _f64]),
    "ptdma_many_solve_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    "ptdma_many_solve_cycle_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    "ptdma_many_solve_host": (ctypes.c_int, [_dp, _dp, _dp, _dp, _i32, _i32, _i32]),
    "ptdma_many_solve_host_noalloc": (ctypes.c_int, [_dp, _dp, _dp, _dp, _i32, _i32, _i32]),
    "ptdma_axis_solve_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp]),
    "ptdma_single_solve_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    "ptdma_single_solve_host": (ctypes.c_int, [_dp, _dp, _dp, _dp, _i32, _i32]),
    "ptdma_many_solve_part_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp]),
    "ptdma_axis_solve_part_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp]),
}

_LIB = None


def load():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise TCSError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no CPU fallback)")
        # torch ships its own HIP/HSA runtime, which its libraries reach under
        # the name libamdhip64.so.  Loaded after ours (soname libamdhip64.so.7
        # from /opt/rocm) it would be a second runtime in the process and find
        # no GPU.  Importing torch first makes our library bind to torch's
        # runtime (same soname), so both share one runtime and one device view.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = load().tcs_last_error()
        raise TCSError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")


def device_count():
    n = ctypes.c_int(0)
    rc = load().tcs_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0
