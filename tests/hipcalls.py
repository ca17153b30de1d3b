"""NumPy-friendly wrappers over the C ABI of libtsdx_hip.so for the tests.

Host-pointer entry points (`*_batched_*`) take NumPy arrays; device entry
points are exercised with torch device tensors.  Every wrapper copies its
inputs, so results can be compared with the oracle's call for call.
"""
import ctypes as C

import numpy as np

from tinysolverdx import _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _t(dtype):
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}[np.dtype(dtype)]


def cholesky(A):
    A = np.array(A, copy=True, order="C")
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_cholesky_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(info))
    return A, info


def lu(A):
    A = np.array(A, copy=True, order="C")
    piv = np.zeros(A.shape[:2], np.int32)
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_lu_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(piv), _p(info))
    return A, piv, info


def cholesky_solve(L, B):
    L = np.ascontiguousarray(L)
    B = np.array(B, copy=True, order="C", dtype=L.dtype)
    _lib.call(f"tsdx_hip_cholesky_solve_batched_{_t(L.dtype)}", _p(L), L.shape[-1], L.shape[0], _p(B), B.shape[-1])
    return B


def lu_solve(LU, piv, B):
    LU = np.ascontiguousarray(LU)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    B = np.array(B, copy=True, order="C", dtype=LU.dtype)
    _lib.call(f"tsdx_hip_lu_solve_batched_{_t(LU.dtype)}", _p(LU), _p(piv), LU.shape[-1], LU.shape[0], _p(B),
              B.shape[-1])
    return B


def posv(A, B):
    A = np.array(A, copy=True, order="C")
    B = np.array(B, copy=True, order="C", dtype=A.dtype)
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_posv_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(B), B.shape[-1], _p(info))
    return A, B, info


def gesv(A, B):
    A = np.array(A, copy=True, order="C")
    B = np.array(B, copy=True, order="C", dtype=A.dtype)
    piv = np.zeros(A.shape[:2], np.int32)
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_gesv_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(B), B.shape[-1], _p(piv),
              _p(info))
    return A, B, piv, info


# ------------------------------------------- emulated fp16 / bf16 (host calls)
def cholesky_emu(A, prec):
    A = np.array(A, copy=True, order="C")
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_cholesky_emu_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(info), _lib.EMU[prec])
    return A, info


def lu_emu(A, prec):
    A = np.array(A, copy=True, order="C")
    piv = np.zeros(A.shape[:2], np.int32)
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_lu_emu_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(piv), _p(info),
              _lib.EMU[prec])
    return A, piv, info


def cholesky_solve_emu(L, B, prec):
    L = np.ascontiguousarray(L)
    B = np.array(B, copy=True, order="C", dtype=L.dtype)
    _lib.call(f"tsdx_hip_cholesky_solve_emu_batched_{_t(L.dtype)}", _p(L), L.shape[-1], L.shape[0], _p(B),
              B.shape[-1], _lib.EMU[prec])
    return B


def lu_solve_emu(LU, piv, B, prec):
    LU = np.ascontiguousarray(LU)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    B = np.array(B, copy=True, order="C", dtype=LU.dtype)
    _lib.call(f"tsdx_hip_lu_solve_emu_batched_{_t(LU.dtype)}", _p(LU), _p(piv), LU.shape[-1], LU.shape[0], _p(B),
              B.shape[-1], _lib.EMU[prec])
    return B


def posv_emu(A, B, prec):
    A = np.array(A, copy=True, order="C")
    B = np.array(B, copy=True, order="C", dtype=A.dtype)
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_posv_emu_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(B), B.shape[-1],
              _p(info), _lib.EMU[prec])
    return A, B, info


def gesv_emu(A, B, prec):
    A = np.array(A, copy=True, order="C")
    B = np.array(B, copy=True, order="C", dtype=A.dtype)
    piv = np.zeros(A.shape[:2], np.int32)
    info = np.zeros(A.shape[0], np.int32)
    _lib.call(f"tsdx_hip_gesv_emu_batched_{_t(A.dtype)}", _p(A), A.shape[-1], A.shape[0], _p(B), B.shape[-1],
              _p(piv), _p(info), _lib.EMU[prec])
    return A, B, piv, info
