"""Zero-copy calls on device-resident torch tensors (the stream-ordered half of
the C ABI, include/tsdx_hip.h).

Tensors must live on a ROCm device, be C-contiguous and have the dtype the
entry point names (float32 / float64 matrices, int32 pivots and info).  Work
is queued on torch's current stream of the tensor's device; nothing is copied
and nothing synchronises.  The functions write their outputs in place, like
the C ABI, and return them for convenience.
"""
from . import _lib


def _dt(A):
    import torch
    if A.dtype == torch.float32:
        return "f32"
    if A.dtype == torch.float64:
        return "f64"
    raise TypeError(f"unsupported dtype {A.dtype}; use float32 or float64")


def _chk(t, name, dtype=None):
    import torch
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}")
    return t.data_ptr()


def _stream(t):
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


def _shape(A):
    if A.dim() != 3 or A.shape[1] != A.shape[2]:
        raise ValueError("expected tensor of shape (batch, n, n)")
    return A.shape[0], A.shape[1]


def potrf(A, info=None):
    """A := lower Cholesky factor (in place); info (batch,) int32 or None."""
    import torch
    batch, n = _shape(A)
    _lib.call(f"tsdx_hip_potrf_{_dt(A)}", _chk(A, "A"), n, batch, _chk(info, "info", torch.int32), _stream(A))
    return A, info


def getrf(A, piv, info=None):
    """A := packed L\\U (in place), piv (batch, n) int32, info (batch,) int32 or None."""
    import torch
    batch, n = _shape(A)
    _lib.call(f"tsdx_hip_getrf_{_dt(A)}", _chk(A, "A"), n, batch, _chk(piv, "piv", torch.int32),
              _chk(info, "info", torch.int32), _stream(A))
    return A, piv, info


def potrs(L, B):
    """B := (L L^T)^-1 B in place; B is (batch, n, nrhs)."""
    batch, n = _shape(L)
    _lib.call(f"tsdx_hip_potrs_{_dt(L)}", _chk(L, "L", L.dtype), n, batch, _chk(B, "B", L.dtype), B.shape[-1],
              _stream(L))
    return B


def getrs(LU, piv, B):
    """B := A^-1 B in place from (LU, piv)."""
    import torch
    batch, n = _shape(LU)
    _lib.call(f"tsdx_hip_getrs_{_dt(LU)}", _chk(LU, "LU", LU.dtype), _chk(piv, "piv", torch.int32), n, batch,
              _chk(B, "B", LU.dtype), B.shape[-1], _stream(LU))
    return B


def posv(A, B, info=None):
    """A := L, B := X in one launch."""
    import torch
    batch, n = _shape(A)
    _lib.call(f"tsdx_hip_posv_{_dt(A)}", _chk(A, "A"), n, batch, _chk(B, "B", A.dtype), B.shape[-1],
              _chk(info, "info", torch.int32), _stream(A))
    return A, B, info


def gesv(A, B, piv, info=None):
    """A := L\\U, piv, B := X in one launch."""
    import torch
    batch, n = _shape(A)
    _lib.call(f"tsdx_hip_gesv_{_dt(A)}", _chk(A, "A"), n, batch, _chk(B, "B", A.dtype), B.shape[-1],
              _chk(piv, "piv", torch.int32), _chk(info, "info", torch.int32), _stream(A))
    return A, B, piv, info


def posv_solve(A, B, X, info=None):
    """X := A^-1 B for SPD A; A and B are only read (no factor is stored)."""
    import torch
    batch, n = _shape(A)
    _lib.call(f"tsdx_hip_posv_solve_{_dt(A)}", _chk(A, "A"), n, batch, _chk(B, "B", A.dtype), _chk(X, "X", A.dtype),
              B.shape[-1], _chk(info, "info", torch.int32), _stream(A))
    return X, info


def gesv_solve(A, B, X, piv=None, info=None):
    """X := A^-1 B with partial pivoting; A and B are only read."""
    import torch
    batch, n = _shape(A)
    _lib.call(f"tsdx_hip_gesv_solve_{_dt(A)}", _chk(A, "A"), n, batch, _chk(B, "B", A.dtype), _chk(X, "X", A.dtype),
              B.shape[-1], _chk(piv, "piv", torch.int32), _chk(info, "info", torch.int32), _stream(A))
    return X, piv, info
