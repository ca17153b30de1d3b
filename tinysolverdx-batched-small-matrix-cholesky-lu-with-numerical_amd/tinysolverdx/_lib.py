"""ctypes binding of libtsdx_hip.so, the C ABI declared in include/tsdx_hip.h.

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
import ctypes as C
import os
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libtsdx_hip.so"

OK, EINVAL, ERUNTIME = 0, -1, -2

_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double

# name -> argument types (every function returns int unless listed in _RES)
SIGNATURES = {
    "tsdx_hip_available": [],
    "tsdx_hip_last_error": [],
    "tsdx_hip_version": [],
}
for _t in ("f32", "f64"):
    SIGNATURES.update({
        f"tsdx_hip_potrf_{_t}": [_vp, _i, _i64, _vp, _vp],
        f"tsdx_hip_getrf_{_t}": [_vp, _i, _i64, _vp, _vp, _vp],
        f"tsdx_hip_potrs_{_t}": [_vp, _i, _i64, _vp, _i, _vp],
        f"tsdx_hip_getrs_{_t}": [_vp, _vp, _i, _i64, _vp, _i, _vp],
        f"tsdx_hip_posv_{_t}": [_vp, _i, _i64, _vp, _i, _vp, _vp],
        f"tsdx_hip_gesv_{_t}": [_vp, _i, _i64, _vp, _i, _vp, _vp, _vp],
        f"tsdx_hip_posv_solve_{_t}": [_vp, _i, _i64, _vp, _vp, _i, _vp, _vp],
        f"tsdx_hip_gesv_solve_{_t}": [_vp, _i, _i64, _vp, _vp, _i, _vp, _vp, _vp],
        f"tsdx_hip_cholesky_batched_{_t}": [_vp, _i, _i64, _vp],
        f"tsdx_hip_lu_batched_{_t}": [_vp, _i, _i64, _vp, _vp],
        f"tsdx_hip_cholesky_solve_batched_{_t}": [_vp, _i, _i64, _vp, _i],
        f"tsdx_hip_lu_solve_batched_{_t}": [_vp, _vp, _i, _i64, _vp, _i],
        f"tsdx_hip_posv_batched_{_t}": [_vp, _i, _i64, _vp, _i, _vp],
        f"tsdx_hip_gesv_batched_{_t}": [_vp, _i, _i64, _vp, _i, _vp, _vp],
    })
for _t in ("f32", "f64"):
    SIGNATURES.update({
        f"tsdx_hip_potrf_emu_{_t}": [_vp, _i, _i64, _vp, _i, _vp],
        f"tsdx_hip_getrf_emu_{_t}": [_vp, _i, _i64, _vp, _vp, _i, _vp],
        f"tsdx_hip_potrs_emu_{_t}": [_vp, _i, _i64, _vp, _i, _i, _vp],
        f"tsdx_hip_getrs_emu_{_t}": [_vp, _vp, _i, _i64, _vp, _i, _i, _vp],
        f"tsdx_hip_cholesky_emu_batched_{_t}": [_vp, _i, _i64, _vp, _i],
        f"tsdx_hip_lu_emu_batched_{_t}": [_vp, _i, _i64, _vp, _vp, _i],
        f"tsdx_hip_cholesky_solve_emu_batched_{_t}": [_vp, _i, _i64, _vp, _i, _i],
        f"tsdx_hip_lu_solve_emu_batched_{_t}": [_vp, _vp, _i, _i64, _vp, _i, _i],
        f"tsdx_hip_posv_emu_batched_{_t}": [_vp, _i, _i64, _vp, _i, _vp, _i],
        f"tsdx_hip_gesv_emu_batched_{_t}": [_vp, _i, _i64, _vp, _i, _vp, _vp, _i],
    })
SIGNATURES["tsdx_hip_malloc"] = [C.POINTER(_vp), C.c_size_t]
SIGNATURES["tsdx_hip_free"] = [_vp]
SIGNATURES["tsdx_hip_memcpy"] = [_vp, _vp, C.c_size_t]
SIGNATURES["tsdx_hip_device_synchronize"] = []
SIGNATURES["tsdx_hip_get_device"] = [C.POINTER(_i)]
SIGNATURES["tsdx_hip_pointer_device"] = [_vp, C.POINTER(_i)]
SIGNATURES["tsdx_hip_posv_ir_workspace_size"] = [_i64]
SIGNATURES["tsdx_hip_posv_ir_f64"] = [_vp, _i, _i64, _vp, _vp, _i, _i, _i, _d, _vp, _vp, _vp, C.c_size_t, _vp]
SIGNATURES["tsdx_hip_posv_ir_noalloc_f64"] = [_vp, _i, _i64, _vp, _vp, _i, _i, _i, _d, _vp, _vp, _vp]
SIGNATURES["tsdx_hip_gesv_ir_workspace_size"] = [_i64]
SIGNATURES["tsdx_hip_gesv_ir_f64"] = [_vp, _i, _i64, _vp, _vp, _i, _i, _i, _d, _vp, _vp, _vp, C.c_size_t, _vp]
for _t in ("f32", "f64"):
    SIGNATURES[f"tsdx_hip_ragged_plan_create_{_t}"] = [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i, C.POINTER(_vp), _vp]
SIGNATURES["tsdx_hip_ragged_run"] = [_vp, _vp]
SIGNATURES["tsdx_hip_ragged_plan_report"] = [_vp]
SIGNATURES["tsdx_hip_ragged_plan_destroy"] = [_vp]
SIGNATURES["tsdx_hip_router_features_f64"] = [_vp, _i, _i64, _vp, _vp]
SIGNATURES["tsdx_hip_router_forward"] = [_vp, _vp, _i64, _vp, _vp]
SIGNATURES["tsdx_hip_router_infer_workspace_size"] = []
SIGNATURES["tsdx_hip_router_infer"] = [_vp, _vp, _i64, _vp, _vp, _vp, C.c_size_t, _vp]
SIGNATURES["tsdx_hip_router_grad_workspace_size"] = [_i64]
SIGNATURES["tsdx_hip_router_grad"] = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, C.c_size_t, _vp]
SIGNATURES["tsdx_hip_router_adam"] = [_vp, _vp, _vp, _vp, _i, C.c_float, _vp]
SIGNATURES["tsdx_hip_posv_routed_workspace_size"] = [_i64]
SIGNATURES["tsdx_hip_posv_routed_f64"] = [_vp, _i, _i64, _vp, _vp, _i, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp,
                                          C.c_size_t, _vp]
SIGNATURES["tsdx_hip_measure_bandwidth_gbs"] = [_i, C.POINTER(_d)]
SIGNATURES["tsdx_hip_measure_peak_gflops"] = [_i, C.POINTER(_d)]
SIGNATURES["tsdx_hip_measure_mfma_tflops"] = [_i, C.POINTER(_d)]
_RES = {"tsdx_hip_last_error": C.c_char_p, "tsdx_hip_posv_ir_workspace_size": C.c_size_t,
        "tsdx_hip_gesv_ir_workspace_size": C.c_size_t,
        "tsdx_hip_ragged_plan_report": C.c_char_p, "tsdx_hip_router_grad_workspace_size": C.c_size_t,
        "tsdx_hip_posv_routed_workspace_size": C.c_size_t, "tsdx_hip_router_infer_workspace_size": C.c_size_t}

KIND = {"cholesky": 0, "lu": 1}

EMU = {"fp16": 1, "bf16": 2}  # TSDX_HIP_EMU_FP16 / _BF16

FACTOR = {"auto": 0, "fp32": 1, "bf16": 2, "fp16": 3, "fp64": 4, "fp64_block": 5}

_lib = None


def bind_runtime():
    """One HIP runtime per process.  PyTorch-ROCm wheels bundle their own
    libamdhip64.so.7 and load it under a different file name; if this library
    loaded /opt/rocm's copy first, torch would bring a second runtime into the
    process and fail to see the GPU.  Loading torch first (when it is
    installed) makes libtsdx_hip.so bind to torch's copy, which has the same
    SONAME.  TSDX_NO_TORCH=1 skips this."""
    if os.environ.get("TSDX_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib():
    """The loaded library (loaded once)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run build.py (tinysolverdx-hip has no CPU fallback)")
        bind_runtime()
        L = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RES.get(name, C.c_int)
        _lib = L
    return _lib


def last_error():
    return lib().tsdx_hip_last_error().decode()


def check(rc, what=""):
    """Raises the exception the reference's bindings raise for the status."""
    if rc == OK:
        return
    msg = last_error() or what
    if rc == EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args), name)
