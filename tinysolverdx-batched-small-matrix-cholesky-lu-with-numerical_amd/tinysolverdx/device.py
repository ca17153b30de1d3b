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


def posv_ir_workspace(batch, device="cuda"):
    """A device workspace for posv_ir (uint8 tensor)."""
    import torch
    nbytes = _lib.lib().tsdx_hip_posv_ir_workspace_size(batch)
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def posv_ir(A, B, X, factor="fp32", max_iter=10, tol=1e-14, info=None, stat=None, work=None):
    """Mixed-precision posv with float64 iterative refinement (float64 A, B, X).

    factor: "fp32", "bf16", "fp16" (low-precision factor + refinement), "fp64"
    (plain float64 solve for every matrix), "fp64_block" (the float64
    fallback tier for every matrix: block Cholesky on the matrix cores with
    float64 refinement for n > 32) or "auto".  Matrices that do not reach tol
    by refinement are solved again in float64.  stat (int32) gets the
    corrections used per matrix, -1 for the float64 fallback."""
    import torch
    batch, n = _shape(A)
    if A.dtype != torch.float64:
        raise TypeError("posv_ir takes float64 A, B and X")
    if work is None:
        work = posv_ir_workspace(batch, A.device)
    _lib.call("tsdx_hip_posv_ir_f64", _chk(A, "A", torch.float64), n, batch, _chk(B, "B", torch.float64),
              _chk(X, "X", torch.float64), B.shape[-1], _lib.FACTOR[factor], int(max_iter), float(tol),
              _chk(info, "info", torch.int32), _chk(stat, "stat", torch.int32), _chk(work, "work"), work.numel(),
              _stream(A))
    return X, info, stat


class RaggedPlan:
    """Matrices of different sizes and kinds factored by one graph launch.

    A is a flat float32 / float64 device tensor; matrix m (row-major n x n) is
    A[offsets[m] : offsets[m] + n*n] with n = sizes[m]; kinds[m] is 0 (Cholesky)
    or 1 (LU).  offsets / piv_offsets are int64, sizes / kinds / piv / info
    int32 device tensors; piv and info may be None.  Each matrix gets the
    result of the per-size potrf / getrf of its kind, bit for bit.  The plan
    is bound to these tensors (keep them alive); run() factors A in place on
    torch's current stream.  tune=True lets the plan time its compiled
    variants on synthetic matrices of the same sizes and keep the fastest.
    """

    def __init__(self, A, offsets, sizes, kinds, piv=None, piv_offsets=None, info=None, tune=True):
        import ctypes as C
        import json
        import torch
        if A.dim() != 1:
            raise ValueError("A must be a flat tensor")
        t = _dt(A)
        self._keep = (A, offsets, sizes, kinds, piv, piv_offsets, info)
        self._plan = C.c_void_p()
        _lib.call(f"tsdx_hip_ragged_plan_create_{t}", _chk(A, "A"), _chk(offsets, "offsets", torch.int64),
                  _chk(sizes, "sizes", torch.int32), _chk(kinds, "kinds", torch.int32), sizes.numel(),
                  _chk(piv, "piv", torch.int32), _chk(piv_offsets, "piv_offsets", torch.int64),
                  _chk(info, "info", torch.int32), 1 if tune else 0, C.byref(self._plan), _stream(A))
        self.report = json.loads(_lib.lib().tsdx_hip_ragged_plan_report(self._plan).decode())

    def run(self):
        _lib.call("tsdx_hip_ragged_run", self._plan, _stream(self._keep[0]))

    def close(self):
        if self._plan:
            _lib.lib().tsdx_hip_ragged_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
