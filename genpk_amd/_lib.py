"""ctypes binding of libgenpk_amd.so (the C ABI declared in include/genpk_amd.h
and include/genpk_compat.h).

The shared library must have been built (``genpk_amd.build.build()`` or
``__graft_entry__.build()``); there is no fallback: every call below runs the
HIP kernels of that library.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgenpk_amd.so")

# Every symbol the public headers declare, with (restype, argtypes).
_c_int, _c_i64, _c_dbl, _c_vp, _c_sz = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "genpk_last_error": (ctypes.c_char_p, []),
    "genpk_plan_create": (_c_int, [ctypes.POINTER(_c_vp), _c_i64, _c_int, _c_vp]),
    "genpk_plan_destroy": (_c_int, [_c_vp]),
    "genpk_plan_set_stream": (_c_int, [_c_vp, _c_vp]),
    "genpk_plan_field_bytes": (_c_int, [_c_vp, ctypes.POINTER(_c_sz)]),
    "genpk_plan_field_ptr": (_c_int, [_c_vp, ctypes.POINTER(_c_vp)]),
    "genpk_plan_attach_field": (_c_int, [_c_vp, _c_vp]),
    "genpk_plan_set_options": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int]),
    "genpk_plan_get_options": (_c_int, [_c_vp, _ip, _ip, _ip, _ip]),
    "genpk_radix_plan": (_c_int, [_c_i64, _c_int, _ip, _c_int, _ip]),
    "genpk_clear": (_c_int, [_c_vp]),
    "genpk_deposit": (_c_int, [_c_vp, _c_dbl, _c_i64, _c_vp, _c_vp, _c_dbl]),
    "genpk_deposit_retain": (_c_int, [_c_vp, _c_dbl, _c_i64, _c_vp, _c_vp, _c_dbl]),
    "genpk_redeposit": (_c_int, [_c_vp]),
    "genpk_drop_order": (_c_int, [_c_vp]),
    "genpk_export_density": (_c_int, [_c_vp, _c_vp, _c_i64, _c_int]),
    "genpk_import_density": (_c_int, [_c_vp, _c_vp]),
    "genpk_import_spectrum": (_c_int, [_c_vp, _c_vp]),
    "genpk_r2c": (_c_int, [_c_vp]),
    "genpk_c2r": (_c_int, [_c_vp]),
    "genpk_bin": (_c_int, [_c_vp, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl]),
    "genpk_power": (_c_int, [_c_vp, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl]),
    "genpk_power_noalloc": (_c_int, [_c_vp, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl]),
    "genpk_adjoint_density": (_c_int, [_c_vp, _c_int, _c_vp, _c_dbl]),
    "genpk_adjoint_gather": (_c_int, [_c_vp, _c_dbl, _c_i64, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp]),
    "genpk_batch_create": (_c_int, [ctypes.POINTER(_c_vp), _c_int, _c_vp]),
    "genpk_batch_add": (_c_int, [_c_vp, _c_i64, _c_int, _c_dbl, _c_i64, _c_vp, _c_vp, _c_dbl, _c_int, _c_dbl, _c_vp,
                                 _c_vp, _c_vp]),
    "genpk_batch_plan_options": (_c_int, [_c_vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "genpk_batch_build": (_c_int, [_c_vp, _c_vp]),
    "genpk_batch_launch": (_c_int, [_c_vp]),
    "genpk_batch_branch_of": (_c_int, [_c_vp, _c_i64, _ip]),
    "genpk_batch_destroy": (_c_int, [_c_vp]),
    "genpk_timing_enable": (_c_int, [_c_vp, _c_int]),
    "genpk_timing_read": (_c_int, [_c_vp, ctypes.c_char_p, _c_sz, _dp, _ip, _c_int, _ip]),
    "genpk_nexttwo": (_c_int, [_c_int]),
    "genpk_type_str": (_c_int, [_c_int, ctypes.c_char_p, _c_sz]),
    "genpk_print_pk": (_c_int, [ctypes.c_char_p, _c_int, _dp, _dp, _ip]),
    # genpk_compat.h
    "invwindow": (_c_dbl, [_c_i64, _c_i64, _c_i64, _c_i64]),
    "powerspectrum": (_c_int, [_c_i64, _dp, _dp, _c_int, _dp, _ip, _dp, _c_dbl, _c_dbl]),
    "genpk_fft_r2c_3d_inplace": (_c_int, [_c_i64, _dp]),
    "genpk_compat_resident": (_c_int, [_c_int]),
    "genpk_compat_flush": (_c_int, []),
    # C++ linkage, gen-pk.h:93
    "_Z8fieldizediPdlPfS0_di": (_c_int, [_c_dbl, _c_int, _dp, _c_i64, _fp, _fp, _c_dbl, _c_int]),
}

_lib = None


class GenPKError(RuntimeError):
    pass


def lib():
    """Load libgenpk_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GenPKError(f"{LIB_PATH} is missing: build it with genpk_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise GenPKError(lib().genpk_last_error().decode())
    return rc
