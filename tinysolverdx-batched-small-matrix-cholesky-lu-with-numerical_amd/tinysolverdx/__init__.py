"""tinysolverdx (MI355X / HIP build): batched small-matrix Cholesky / LU on NumPy arrays.

The public functions keep the reference package's interface
(python/tinysolverdx/__init__.py:1-117): matrices come as a (batch, n, n)
array or one (n, n) matrix, n in 1..64, and the array's dtype picks the
precision unless `precision` is given (float32 -> "fp32", float64 -> "fp64",
float16 -> "fp16"; "bf16" on a float32 array gives float32 results holding
bfloat16 values).  Inputs are never modified.  The work runs on the GPU
(backend "auto" = "hip"); the reference's CPU backends ("naive", "scalar",
"simd") are not part of this build and raise RuntimeError.  round_to and the
long-double error metrics run on the host, as in the reference.

Device arrays: every solver also takes matrices that already live in GPU
memory, from any producer of __cuda_array_interface__ or DLPack (torch
tensors, CuPy / Numba arrays, DeviceArray), without a trip over PCIe.  Then
the results are DeviceArray objects (library-owned GPU memory that exports
both protocols: torch.from_dlpack(x) or cupy.asarray(x) take them without a
copy, x.numpy() copies to the host); every array argument of the call must
then be a device array, float32 or float64 (int32 pivots), and the precision
is the array's own or the emulated fp16 / bf16.

Beyond the reference: posv / gesv run factor and solve as one device launch,
the machine probes measure the GPU, and `tinysolverdx.device` works on torch
device tensors with stream-ordered, allocation-free calls.
"""
import numpy as np

from . import _lib

_lib.bind_runtime()  # before _core pulls in libtsdx_hip.so: one HIP runtime per process

from . import _core  # noqa: E402
from ._core import (  # noqa: E402,F401  (re-exported)
    has_hip, has_cuda, has_openmp, num_threads, set_num_threads, unit_roundoff,
    chol_backward_error, lu_backward_error, solve_backward_error,
    solve_forward_error, reference_solve, measure_peak_gflops, measure_bandwidth_gbs,
    measure_mfma_tflops, DeviceArray,
)

__all__ = ["cholesky", "lu", "cholesky_solve", "lu_solve", "posv", "gesv", "round_to",
           "has_hip", "has_cuda", "has_openmp", "num_threads", "set_num_threads", "unit_roundoff",
           "chol_backward_error", "lu_backward_error", "solve_backward_error",
           "solve_forward_error", "reference_solve", "measure_peak_gflops", "measure_bandwidth_gbs",
           "measure_mfma_tflops", "DeviceArray"]

# dtype -> precision used when none is given
_IMPLIED = {np.dtype(np.float16): "fp16", np.dtype(np.float32): "fp32", np.dtype(np.float64): "fp64"}
_KNOWN = frozenset(("fp16", "bf16", "fp32", "fp64"))


class _Stack:
    """One call's matrices as the C-contiguous (batch, n, n) array the core
    takes: float64 storage stays float64 (the core rounds into a narrower
    precision itself), everything else is stored as float32."""

    def __init__(self, A, precision):
        a = np.asarray(A)
        self.single = a.ndim == 2
        self.caller_dtype = a.dtype
        if precision is None:
            precision = _IMPLIED.get(a.dtype)
            if precision is None:
                raise TypeError(f"unsupported dtype {a.dtype}; use float16/32/64")
        elif precision not in _KNOWN:
            raise ValueError("precision must be one of fp16, bf16, fp32, fp64")
        self.precision = precision
        storage = np.float64 if a.dtype == np.float64 else np.float32
        self.data = np.ascontiguousarray(a[np.newaxis] if self.single else a, dtype=storage)
        self._suffix = "f64" if storage is np.float64 else "f32"

    def call(self, op, *args):
        """_core.<op>_<storage>(*args)."""
        return getattr(_core, f"{op}_{self._suffix}")(*args)

    def rhs(self, B):
        """(right-hand sides as (batch, n, k) in the matrices' storage, whether
        the caller passed vectors).  One matrix takes b (n,) or B (n, k); a
        stack takes (batch, n) or (batch, n, k)."""
        b = np.asarray(B)
        vectors = b.ndim == (1 if self.single else 2)
        if vectors:
            b = b[..., np.newaxis]
        if self.single:
            b = b[np.newaxis]
        return np.ascontiguousarray(b, dtype=self.data.dtype), vectors

    def result(self, x, vectors=False):
        """A solver output in the caller's terms: float16 again for fp16 work on
        float16 input, the vector axis and the batch axis dropped where the
        caller had none."""
        if self.precision == "fp16" and self.caller_dtype == np.float16:
            x = x.astype(np.float16)
        if vectors:
            x = x[..., 0]
        return x[0] if self.single else x

    def flag(self, x):
        """info / piv for the caller's shape."""
        return x[0] if self.single else x


def cholesky(A, precision=None, backend="auto"):
    """Lower Cholesky factors and LAPACK-style info (0 = ok, k>0: pivot k not positive)."""
    if _core.is_device_array(A):
        return _core.cholesky_device(A, precision or "", backend)
    s = _Stack(A, precision)
    L, info = s.call("cholesky", s.data, s.precision, backend)
    return s.result(L), s.flag(info)


def lu(A, precision=None, backend="auto"):
    """Packed LU with partial pivoting: (LU, piv, info). piv is scipy/LAPACK style (0-based)."""
    if _core.is_device_array(A):
        return _core.lu_device(A, precision or "", backend)
    s = _Stack(A, precision)
    LU, piv, info = s.call("lu", s.data, s.precision, backend)
    return s.result(LU), s.flag(piv), s.flag(info)


def cholesky_solve(L, B, precision=None):
    """Solve L L^T x = B given factors from :func:`cholesky`."""
    if _core.is_device_array(L):
        return _core.cholesky_solve_device(L, B, precision or "")
    s = _Stack(L, precision)
    b, vectors = s.rhs(B)
    return s.result(s.call("cholesky_solve", s.data, b, s.precision), vectors)


def lu_solve(LU, piv, B, precision=None):
    """Solve A x = B given (LU, piv) from :func:`lu`."""
    if _core.is_device_array(LU):
        return _core.lu_solve_device(LU, piv, B, precision or "")
    s = _Stack(LU, precision)
    p = np.asarray(piv)
    p = np.ascontiguousarray(p[np.newaxis] if s.single else p, dtype=np.int32)
    b, vectors = s.rhs(B)
    return s.result(s.call("lu_solve", s.data, p, b, s.precision), vectors)


def posv(A, B, precision=None, backend="auto"):
    """Solve SPD systems A x = B (factor and solve fused on the device). Returns (x, info)."""
    if _core.is_device_array(A):
        return _core.posv_device(A, B, precision or "", backend)
    s = _Stack(A, precision)
    b, vectors = s.rhs(B)
    _, X, info = s.call("posv", s.data, b, s.precision, backend)
    return s.result(X, vectors), s.flag(info)


def gesv(A, B, precision=None, backend="auto"):
    """Solve general systems A x = B with partial pivoting (fused on the device). Returns (x, info)."""
    if _core.is_device_array(A):
        return _core.gesv_device(A, B, precision or "", backend)
    s = _Stack(A, precision)
    b, vectors = s.rhs(B)
    _, _, X, info = s.call("gesv", s.data, b, s.precision, backend)
    return s.result(X, vectors), s.flag(info)


def round_to(x, precision):
    """Round values to the nearest number representable in ``precision`` (float64 in/out)."""
    return _core.round_to_f64(np.ascontiguousarray(x, dtype=np.float64), precision)
