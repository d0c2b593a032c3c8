"""ctypes binding of libdiscodj_hip.so (the C ABI of include/discodj_hip.h).

The library is built in-tree by `make -C discodj.jl_amd/csrc` (the repository's build step runs the
same make).  There is no fallback: if the library is missing, loading fails with an error that says how to build it.
"""
import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libdiscodj_hip.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "discodj_hip.h")

DJ_OK = 0
ERRORS = {-1: "DJ_ERR_INVALID_CONFIG", -2: "DJ_ERR_INVALID_ARGUMENT", -3: "DJ_ERR_UNSUPPORTED",
          -4: "DJ_ERR_OUT_OF_MEMORY", -5: "DJ_ERR_DEVICE", -6: "DJ_ERR_NONFINITE"}
DEPOSIT = {"float": 0, "fixed64": 1}


class DJError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class DJConfig(C.Structure):
    _fields_ = [("res_pm", C.c_int32), ("n_part", C.c_int32), ("boxsize", C.c_double),
                ("worder", C.c_int32), ("deconvolve", C.c_int32), ("antialias", C.c_int32),
                ("grad_order", C.c_int32), ("lap_order", C.c_int32), ("n_resample", C.c_int32),
                ("resampling", C.c_int32), ("precision", C.c_int32), ("deposit", C.c_int32),
                ("fixed_frac_bits", C.c_int32), ("lean", C.c_int32), ("device", C.c_int32)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("calls", C.c_int32), ("ms_total", C.c_double),
                ("bytes_per_call", C.c_double)]


STEP_CB = C.CFUNCTYPE(None, C.c_int32, C.c_void_p)

_SIGS = {
    "dj_create": (C.c_int32, [C.POINTER(DJConfig), C.POINTER(C.c_void_p)]),
    "dj_destroy": (C.c_int32, [C.c_void_p]),
    "dj_last_error": (C.c_char_p, [C.c_void_p]),
    "dj_set_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_get_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_set_state_device": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_get_state_device": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_acceleration": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_step": (C.c_int32, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]),
    "dj_run": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p, STEP_CB, C.c_void_p]),
    "dj_deposit": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_check_finite": (C.c_int32, [C.c_void_p]),
    "dj_set_kgain": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double]),
    "dj_synchronize": (C.c_int32, [C.c_void_p]),
    "dj_set_profiling": (C.c_int32, [C.c_void_p, C.c_int32]),
    "dj_reset_stats": (C.c_int32, [C.c_void_p]),
    "dj_kernel_stats": (C.c_int32, [C.c_void_p, C.POINTER(KernelStat), C.c_int32, C.POINTER(C.c_int32)]),
    "dj_fft3d": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]),
    "dj_version": (C.c_char_p, []),
    "dj_ics_lpt": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                               C.c_double, C.c_int32]),
    "dj_power_spectrum": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "dj_field_power_spectrum": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                            C.c_void_p, C.c_void_p]),
    "dj_field_cross_power_spectrum": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                                  C.c_void_p, C.c_void_p]),
    "dj_power_spectra_binned": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "dj_field_power_spectrum_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                                    C.c_void_p, C.c_void_p]),
    "dj_ngenic_field": (C.c_int32, [C.c_uint32, C.c_int32, C.c_void_p]),
    "dj_ngenic_field_on_device": (C.c_int32, [C.c_uint32, C.c_int32, C.c_int32, C.c_void_p]),
    "dj_ranlxd1_uniforms": (C.c_int32, [C.c_uint32, C.c_int32, C.c_void_p]),
    "dj_vjp_acceleration": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_vjp_deposit": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_run_checkpointed": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_vjp_run": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_vjp_ics_lpt": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                   C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "dj_ics_lpt_device": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                      C.c_double, C.c_int32]),
    "dj_vjp_ics_lpt_device": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                          C.c_double, C.c_int32, C.c_void_p]),
    "dj_deposit_device": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "dj_vjp_deposit_device": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "dj_density_loss_device": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    "dj_nn_set_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "dj_nn_get_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "dj_nn_enable": (C.c_int32, [C.c_void_p, C.c_int32, C.c_double]),
    "dj_nn_set_schedule": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
    "dj_nn_forward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "dj_nn_backward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    "dj_nn_train_step": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                     C.c_int32, C.POINTER(C.c_double)]),
}
NN_PARAMS = 4 * (27 * 16 * 16 + 16)

_lib = None


def header_exports():
    """Names of the functions declared by include/discodj_hip.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:dj_status|const char \*|void)\s*\*?\s*(dj_\w+)\s*\(", txt, re.M)))


def lib():
    """The loaded library; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C discodj.jl_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_hip = None


def hip():
    """The HIP runtime the library was loaded with (the copy of libamdhip64 mapped into this process for it):
    device memory for the _device entry points is allocated with it, so that library and caller share one
    runtime and one view of the device."""
    global _hip
    if _hip is None:
        lib()
        paths = [ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64.so" in ln]
        if not paths:
            raise RuntimeError("libamdhip64 is not mapped into this process")
        H = C.CDLL(paths[0])
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipSetDevice.argtypes = [C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


def check(rc, ctx=None):
    if rc != DJ_OK:
        msg = lib().dj_last_error(ctx)
        raise DJError(rc, msg.decode() if msg else "")
    return rc
