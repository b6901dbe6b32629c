"""ctypes binding of libpwdft_hip.so (the C-ABI declared in include/pwdft_hip.h).

The library is built in-tree by pwdft_develop_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
torch is imported before the library when it is available, so that the HIP
runtime torch ships (same SONAME) is the one both use in this process.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpwdft_hip.so")

PW_F64 = 0
PW_F32 = 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_LIB = None


class PwError(RuntimeError):
    """A non-zero status from the C-ABI (the reference raises std::runtime_error
    from ErrorHandling, src/global.cpp:71-91)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m pwdft_develop_amd.build` "
                          "(there is no CPU fallback)")
    try:
        import torch  # noqa: F401  (share torch's HIP runtime when present)
    except Exception:
        pass
    L = C.CDLL(LIB_PATH)
    L.pw_last_error.restype = C.c_char_p
    L.pw_last_error.argtypes = []
    sig = {
        "pw_sphere_build": [_ip, _dp, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp, _vp, _vp, _vp],
        "pw_fock_sphere_build": [_ip, _dp, C.c_double, C.POINTER(C.c_int), _vp, _vp],
        "pw_ekin_teter": [C.c_int, _dp, _dp, _dp],
        "pw_gaussian_projectors": [_ip, _dp, C.c_int, _ip, C.c_int, _dp, C.c_double, C.c_int, _dp, _dp, _dp],
        "pw_ctx_create": [C.POINTER(_vp), _ip, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _dp, _dp, C.c_int, _vp],
        "pw_ctx_destroy": [_vp],
        "pw_ctx_set_stream": [_vp, _vp],
        "pw_ctx_info": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long)],
        "pw_set_vtot": [_vp, _dp],
        "pw_set_nonlocal": [_vp, C.c_int, _dp, _dp],
        "pw_mult_spinor": [_vp, C.c_int, C.c_int, _vp, _vp],
        "pw_nonlocal_multx": [_vp, C.c_int, C.c_int, _vp, _vp],
        "pw_nonlocal_multx_noalloc": [_vp, C.c_int, C.c_int, _vp, _vp],
        "pw_apply_h": [_vp, C.c_int, C.c_int, _vp, _vp],
        "pw_teter_precond": [_vp, C.c_int, C.c_int, _vp, _vp],
        "pw_orthogonalize": [_vp, C.c_int, _vp],
        "pw_lanczos": [_vp, C.c_int, _vp, _dp, C.POINTER(C.c_double)],
        "pw_cheb_filter": [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _vp],
        "pw_rayleigh_ritz": [_vp, C.c_int, C.c_int, _vp, _dp, _dp],
        "pw_chebfsi_sweep": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _dp, _dp, _dp],
        "pw_free_workspace": [_vp],
        "pw_set_graphs": [_vp, C.c_int],
        "pw_syevd": [_vp, C.c_int, _vp, C.c_int, _vp],
        "pw_sygvd_batched": [_vp, C.c_int, _vp, C.c_int, C.c_long, _vp, C.c_int, C.c_long, _vp, C.c_long, _vp,
                             C.c_int],
        "pw_density": [_vp, C.c_int, C.c_int, _vp, _dp, _vp, C.c_int],
        "pw_band_energy": [_vp, C.c_int, C.c_int, _vp, _dp, C.POINTER(C.c_double)],
        "pw_harris_energy": [_vp, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, C.c_int, _dp, _dp,
                             C.c_int, C.c_int, _dp, C.POINTER(C.c_double)],
        "pw_calculate_density": [_vp, C.c_int, C.c_int, _vp, _dp, C.c_int, C.c_int, _vp,
                                 C.POINTER(C.c_double)],
        "pw_nonlocal_force": [_vp, C.c_int, C.c_int, _vp, _dp, C.c_int, C.c_int, _dp],
        "pw_set_exx": [_vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _vp, _dp],
        "pw_add_exx": [_vp, C.c_int, C.c_int, _vp, _vp],
        "pw_build_ace": [_vp, C.c_int, _vp],
        "pw_exx_divergence": [_ip, _dp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)],
        "pw_lobpcg": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _dp, _dp,
                      C.POINTER(C.c_int)],
        "pw_ppcg": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _vp, _dp, _dp,
                    C.POINTER(C.c_int)],
        "pw_set_timing": [_vp, C.c_int],
        "pw_get_timing": [_vp, _dp, C.POINTER(C.c_int)],
        "pw_batch_create": [C.POINTER(_vp), C.c_int, _ip, _dp, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _ip, _dp,
                            _dp, _ip, _vp],
        "pw_batch_destroy": [_vp],
        "pw_batch_set_stream": [_vp, _vp],
        "pw_batch_info": [_vp, C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "pw_batch_tune": [_vp, C.c_int, C.c_int, C.c_int],
        "pw_batch_set_graphs": [_vp, C.c_int],
        "pw_batch_set_timing": [_vp, C.c_int],
        "pw_batch_timing": [_vp, _dp],
        "pw_batch_autotune": [_vp, C.c_char_p, C.c_char_p, C.c_int],
        "pw_autotune": [_vp, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int],
        "pw_batch_apply_h": [_vp, _vp, _vp],
        "pw_batch_cheb_filter": [_vp, C.c_int, _dp, _vp],
        "pw_batch_cheb_filter_serial": [_vp, C.c_int, _dp, _vp],
        "pw_rs_create": [C.POINTER(_vp), _ip, _ip, C.c_double, _dp, C.c_int, _vp],
        "pw_rs_destroy": [_vp],
        "pw_rs_set_stream": [_vp, _vp],
        "pw_rs_info": [_vp, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_int)],
        "pw_rs_set_vtot": [_vp, _dp],
        "pw_rs_set_nonlocal": [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp],
        "pw_rs_mult_spinor": [_vp, C.c_int, _vp, _vp],
        "pw_rs_filter": [_vp, C.c_double, C.c_int, _vp],
        "pw_rs_teter_precond": [_vp, C.c_int, _vp, _vp],
        "pw_rs_ppcg": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _vp, _dp, _dp,
                       C.POINTER(C.c_int)],
        "pw_rs_calculate_density": [_vp, C.c_int, _vp, _dp, C.c_int, C.c_int, _vp, C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        raise PwError(rc, lib().pw_last_error().decode())
    return rc


def exported_symbols():
    """Names declared in include/pwdft_hip.h (used by the export test)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "pwdft_hip.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(pw_\w+)\s*\(", txt, re.M)))
