"""tinysolverdx (MI355X / HIP build): batched small-matrix Cholesky / LU on NumPy arrays.

Same API as the reference package (python/tinysolverdx/__init__.py:1-117):
inputs are (batch, n, n) arrays (a single (n, n) matrix is also accepted) with
n in 1..64; dtype selects the precision (float32 -> fp32, float64 -> fp64).
The work runs on the GPU (backend "auto" = "hip"); the reference's CPU
backends ("naive", "scalar", "simd") and its emulated fp16 / bf16 precisions
are not part of this build and raise RuntimeError.  round_to and the
long-double error metrics run on the host, as in the reference.

Differences that add to the reference's API: posv / gesv solve in one fused
device launch; `tinysolverdx.device` takes torch device tensors without copies.
"""
import numpy as np

from . import _lib

_lib.bind_runtime()  # before _core pulls in libtsdx_hip.so: one HIP runtime per process

from . import _core  # noqa: E402
from ._core import (  # noqa: E402,F401  (re-exported)
    has_hip, has_cuda, has_openmp, num_threads, set_num_threads, unit_roundoff,
    chol_backward_error, lu_backward_error, solve_backward_error,
    solve_forward_error, reference_solve,
)

__all__ = ["cholesky", "lu", "cholesky_solve", "lu_solve", "posv", "gesv", "round_to",
           "has_hip", "has_cuda", "has_openmp", "num_threads", "set_num_threads", "unit_roundoff",
           "chol_backward_error", "lu_backward_error", "solve_backward_error",
           "solve_forward_error", "reference_solve"]

_DEFAULT_PREC = {np.dtype(np.float32): "fp32", np.dtype(np.float64): "fp64", np.dtype(np.float16): "fp16"}
_PRECISIONS = ("fp16", "bf16", "fp32", "fp64")


def _fn(name, arr):
    # one entry point per storage dtype, no implicit casting (reference __init__.py:25-27)
    return getattr(_core, f"{name}_{'f64' if arr.dtype == np.float64 else 'f32'}")


def _prep(a, precision, ndim_single):
    """(contiguous work array, precision, was-single, caller dtype); reference __init__.py:30-45."""
    a = np.asarray(a)
    single = a.ndim == ndim_single
    if single:
        a = a[None]
    out_dtype = a.dtype
    if precision is None:
        if a.dtype not in _DEFAULT_PREC:
            raise TypeError(f"unsupported dtype {a.dtype}; use float16/32/64")
        precision = _DEFAULT_PREC[a.dtype]
    if precision not in _PRECISIONS:
        raise ValueError("precision must be one of fp16, bf16, fp32, fp64")
    work = np.float64 if (a.dtype == np.float64) else np.float32
    return np.ascontiguousarray(a, dtype=work), precision, single, out_dtype


def _finish(x, out_dtype, precision):
    if precision == "fp16" and out_dtype == np.float16:
        return x.astype(np.float16)
    return x


def _rhs(B, ref, single):
    B = np.asarray(B)
    if single:
        B = B[:, None] if B.ndim == 1 else B
        B = B[None]
    elif B.ndim == 2:
        B = B[:, :, None]
    return np.ascontiguousarray(B, dtype=ref.dtype)


def cholesky(A, precision=None, backend="auto"):
    """Lower Cholesky factors and LAPACK-style info (0 = ok, k>0: pivot k not positive)."""
    a, p, single, dt = _prep(A, precision, 2)
    L, info = _fn("cholesky", a)(a, p, backend)
    L = _finish(L, dt, p)
    return (L[0], info[0]) if single else (L, info)


def lu(A, precision=None, backend="auto"):
    """Packed LU with partial pivoting: (LU, piv, info). piv is scipy/LAPACK style (0-based)."""
    a, p, single, dt = _prep(A, precision, 2)
    LU, piv, info = _fn("lu", a)(a, p, backend)
    LU = _finish(LU, dt, p)
    return (LU[0], piv[0], info[0]) if single else (LU, piv, info)


def cholesky_solve(L, B, precision=None):
    """Solve L L^T x = B given factors from :func:`cholesky`."""
    l, p, single, dt = _prep(L, precision, 2)
    vec = np.asarray(B).ndim == (1 if single else 2)
    X = _fn("cholesky_solve", l)(l, _rhs(B, l, single), p)
    X = _finish(X, dt, p)
    X = X[..., 0] if vec else X
    return X[0] if single else X


def lu_solve(LU, piv, B, precision=None):
    """Solve A x = B given (LU, piv) from :func:`lu`."""
    l, p, single, dt = _prep(LU, precision, 2)
    pv = np.ascontiguousarray(np.asarray(piv)[None] if single else piv, dtype=np.int32)
    vec = np.asarray(B).ndim == (1 if single else 2)
    X = _fn("lu_solve", l)(l, pv, _rhs(B, l, single), p)
    X = _finish(X, dt, p)
    X = X[..., 0] if vec else X
    return X[0] if single else X


def posv(A, B, precision=None, backend="auto"):
    """Solve SPD systems A x = B (factor and solve fused on the device). Returns (x, info)."""
    a, p, single, dt = _prep(A, precision, 2)
    vec = np.asarray(B).ndim == (1 if single else 2)
    _, X, info = _fn("posv", a)(a, _rhs(B, a, single), p, backend)
    X = _finish(X, dt, p)
    X = X[..., 0] if vec else X
    return (X[0], info[0]) if single else (X, info)


def gesv(A, B, precision=None, backend="auto"):
    """Solve general systems A x = B with partial pivoting (fused on the device). Returns (x, info)."""
    a, p, single, dt = _prep(A, precision, 2)
    vec = np.asarray(B).ndim == (1 if single else 2)
    _, _, X, info = _fn("gesv", a)(a, _rhs(B, a, single), p, backend)
    X = _finish(X, dt, p)
    X = X[..., 0] if vec else X
    return (X[0], info[0]) if single else (X, info)


def round_to(x, precision):
    """Round values to the nearest number representable in ``precision`` (float64 in/out)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _core.round_to_f64(x, precision)
