"""Zero-copy calls on device-resident torch tensors (the stream-ordered half of
the C ABI, include/tsdx_hip.h).

Tensors must live on the ROCm device of the matrices, be C-contiguous and have
the dtype and shape the entry point names (float32 / float64 matrices, int32
pivots and info).  Every shape is checked on the host before anything is
launched: the kernels trust the sizes they are given, so a wrong-length
buffer would otherwise be written out of bounds.  Work is queued on torch's
current stream of the tensor's device; nothing is copied and nothing
synchronises.  The functions write their outputs in place, like the C ABI, and
return them for convenience.
"""
from . import _lib


def _torch():
    import torch
    return torch


def _dt(A):
    torch = _torch()
    if A.dtype == torch.float32:
        return "f32"
    if A.dtype == torch.float64:
        return "f64"
    raise TypeError(f"unsupported dtype {A.dtype}; use float32 or float64")


def _stream(t):
    return _torch().cuda.current_stream(t.device).cuda_stream


class Call:
    """Host-side validation of one call's tensors against its matrices.

    `Call(A)` takes the (batch, n, n) matrices (or, with `square=False`, any
    device tensor whose first dimension is the batch); every other operand is
    passed through one of the methods below, which check device, layout,
    dtype and exact shape and return the device address (None for an
    optional operand given as None)."""

    def __init__(self, A, name="A", dtypes=None, square=True):
        torch = _torch()
        self._dev(A, name)
        if square:
            if A.dim() != 3 or A.shape[1] != A.shape[2]:
                raise ValueError(f"{name} must have shape (batch, n, n), got {tuple(A.shape)}")
            self.batch, self.n = int(A.shape[0]), int(A.shape[1])
        else:
            self.batch, self.n = int(A.shape[0]) if A.dim() else 0, 0
        dtypes = dtypes or (torch.float32, torch.float64)
        if A.dtype not in dtypes:
            raise TypeError(f"{name} must be one of {', '.join(map(str, dtypes))}, got {A.dtype}")
        self.A, self.device, self.dtype = A, A.device, A.dtype

    @staticmethod
    def _dev(t, name):
        if not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")

    def tensor(self, t, name, dtype, shape, optional=False):
        """Address of `t` after checking it is (shape) of dtype on A's device."""
        if t is None:
            if optional:
                return None
            raise ValueError(f"{name} must not be None")
        self._dev(t, name)
        if t.device != self.device:
            raise ValueError(f"{name} is on {t.device}, the matrices on {self.device}")
        if t.dtype != dtype:
            raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
        return t.data_ptr()

    def mat(self):
        return self.A.data_ptr()

    def ints(self, t, name, optional=True):
        """int32 (batch,)."""
        return self.tensor(t, name, _torch().int32, (self.batch,), optional)

    def piv(self, t, optional=False):
        """int32 (batch, n)."""
        return self.tensor(t, "piv", _torch().int32, (self.batch, self.n), optional)

    def rhs(self, B, name="B"):
        """(batch, n, nrhs) of the matrices' dtype; returns (address, nrhs)."""
        if B is None or B.dim() != 3:
            raise ValueError(f"{name} must have shape (batch, n, nrhs)")
        nrhs = int(B.shape[2])
        if nrhs < 1:
            raise ValueError(f"{name} needs at least one right-hand side")
        return self.tensor(B, name, self.dtype, (self.batch, self.n, nrhs)), nrhs

    def like(self, X, B, name="X"):
        """X of B's shape and the matrices' dtype."""
        return self.tensor(X, name, self.dtype, tuple(B.shape))

    def work(self, w, need, name="work"):
        """A uint8 workspace of at least `need` bytes."""
        if w is None:
            raise ValueError(f"{name} must not be None")
        self._dev(w, name)
        if w.dtype != _torch().uint8 or w.dim() != 1:
            raise TypeError(f"{name} must be a flat uint8 tensor")
        if w.numel() < need:
            raise ValueError(f"{name} holds {w.numel()} bytes, the call needs {need}")
        return w.data_ptr(), int(w.numel())

    def stream(self):
        return _stream(self.A)


def potrf(A, info=None):
    """A := lower Cholesky factor (in place); info (batch,) int32 or None."""
    c = Call(A)
    _lib.call(f"tsdx_hip_potrf_{_dt(A)}", c.mat(), c.n, c.batch, c.ints(info, "info"), c.stream())
    return A, info


def getrf(A, piv, info=None):
    """A := packed L\\U (in place), piv (batch, n) int32, info (batch,) int32 or None."""
    c = Call(A)
    _lib.call(f"tsdx_hip_getrf_{_dt(A)}", c.mat(), c.n, c.batch, c.piv(piv), c.ints(info, "info"), c.stream())
    return A, piv, info


def potrs(L, B):
    """B := (L L^T)^-1 B in place; B is (batch, n, nrhs)."""
    c = Call(L, "L")
    b, nrhs = c.rhs(B)
    _lib.call(f"tsdx_hip_potrs_{_dt(L)}", c.mat(), c.n, c.batch, b, nrhs, c.stream())
    return B


def getrs(LU, piv, B):
    """B := A^-1 B in place from (LU, piv)."""
    c = Call(LU, "LU")
    b, nrhs = c.rhs(B)
    _lib.call(f"tsdx_hip_getrs_{_dt(LU)}", c.mat(), c.piv(piv), c.n, c.batch, b, nrhs, c.stream())
    return B


def posv(A, B, info=None):
    """A := L, B := X in one launch."""
    c = Call(A)
    b, nrhs = c.rhs(B)
    _lib.call(f"tsdx_hip_posv_{_dt(A)}", c.mat(), c.n, c.batch, b, nrhs, c.ints(info, "info"), c.stream())
    return A, B, info


def gesv(A, B, piv, info=None):
    """A := L\\U, piv, B := X in one launch."""
    c = Call(A)
    b, nrhs = c.rhs(B)
    _lib.call(f"tsdx_hip_gesv_{_dt(A)}", c.mat(), c.n, c.batch, b, nrhs, c.piv(piv), c.ints(info, "info"),
              c.stream())
    return A, B, piv, info


def posv_solve(A, B, X, info=None):
    """X := A^-1 B for SPD A; A and B are only read (no factor is stored)."""
    c = Call(A)
    b, nrhs = c.rhs(B)
    _lib.call(f"tsdx_hip_posv_solve_{_dt(A)}", c.mat(), c.n, c.batch, b, c.like(X, B), nrhs, c.ints(info, "info"),
              c.stream())
    return X, info


def gesv_solve(A, B, X, piv=None, info=None):
    """X := A^-1 B with partial pivoting; A and B are only read."""
    c = Call(A)
    b, nrhs = c.rhs(B)
    _lib.call(f"tsdx_hip_gesv_solve_{_dt(A)}", c.mat(), c.n, c.batch, b, c.like(X, B), nrhs,
              c.piv(piv, optional=True), c.ints(info, "info"), c.stream())
    return X, piv, info


def posv_ir_workspace(batch, device="cuda"):
    """A device workspace for posv_ir (uint8 tensor)."""
    nbytes = _lib.lib().tsdx_hip_posv_ir_workspace_size(batch)
    return _torch().empty(nbytes, dtype=_torch().uint8, device=device)


def posv_ir(A, B, X, factor="fp32", max_iter=10, tol=1e-14, info=None, stat=None, work=None):
    """Mixed-precision posv with float64 iterative refinement (float64 A, B, X).

    factor: "fp32", "bf16", "fp16" (low-precision factor + refinement), "fp64"
    (plain float64 solve for every matrix), "fp64_block" (the float64
    fallback tier for every matrix: block Cholesky on the matrix cores with
    float64 refinement for n > 32) or "auto".  Matrices that do not reach tol
    by refinement are solved again in float64.  stat (int32) gets the
    corrections used per matrix, -1 for the float64 fallback."""
    torch = _torch()
    c = Call(A, dtypes=(torch.float64,))
    b, nrhs = c.rhs(B)
    if factor not in _lib.FACTOR:
        raise ValueError(f"factor must be one of {', '.join(_lib.FACTOR)}")
    if work is None:
        work = posv_ir_workspace(c.batch, A.device)
    w, wn = c.work(work, _lib.lib().tsdx_hip_posv_ir_workspace_size(c.batch))
    _lib.call("tsdx_hip_posv_ir_f64", c.mat(), c.n, c.batch, b, c.like(X, B), nrhs, _lib.FACTOR[factor],
              int(max_iter), float(tol), c.ints(info, "info"), c.ints(stat, "stat"), w, wn, c.stream())
    return X, info, stat


def gesv_ir_workspace(batch, device="cuda"):
    """A device workspace for gesv_ir (uint8 tensor)."""
    nbytes = _lib.lib().tsdx_hip_gesv_ir_workspace_size(batch)
    return _torch().empty(nbytes, dtype=_torch().uint8, device=device)


def gesv_ir(A, B, X, factor="fp32", max_iter=10, tol=1e-14, info=None, stat=None, work=None):
    """Mixed-precision gesv with float64 iterative refinement (float64 A, B, X).

    factor: "fp32", "bf16", "fp16" (partial-pivoting LU in that precision +
    refinement), "fp64" (plain float64 gesv for every matrix) or "auto".  A
    matrix whose low-precision LU meets a zero pivot, or whose refinement does
    not reach tol, is solved again by the float64 gesv, which also gives its
    info; converged matrices get info 0.  stat (int32) gets the corrections
    used per matrix, -1 for the float64 solve.  A and B are only read."""
    torch = _torch()
    c = Call(A, dtypes=(torch.float64,))
    b, nrhs = c.rhs(B)
    if factor not in _lib.FACTOR or factor == "fp64_block":
        raise ValueError("factor must be one of auto, fp32, bf16, fp16, fp64")
    if work is None:
        work = gesv_ir_workspace(c.batch, A.device)
    w, wn = c.work(work, _lib.lib().tsdx_hip_gesv_ir_workspace_size(c.batch))
    _lib.call("tsdx_hip_gesv_ir_f64", c.mat(), c.n, c.batch, b, c.like(X, B), nrhs, _lib.FACTOR[factor],
              int(max_iter), float(tol), c.ints(info, "info"), c.ints(stat, "stat"), w, wn, c.stream())
    return X, info, stat


class RaggedPlan:
    """Matrices of different sizes and kinds factored by one graph launch.

    A is a flat float32 / float64 device tensor; matrix m (row-major n x n) is
    A[offsets[m] : offsets[m] + n*n] with n = sizes[m]; kinds[m] is 0 (Cholesky)
    or 1 (LU).  offsets / piv_offsets are int64, sizes / kinds / piv / info
    int32 device tensors; piv and info may be None.  Each matrix gets the
    result of the per-size potrf / getrf of its kind, bit for bit.  The plan
    is bound to these tensors (keep them alive; the plan keeps its own copy of
    sizes and kinds, so later changes to them do not reach it, while offsets
    and piv_offsets must not change); run() factors A in place on torch's
    current stream.  tune=True lets the plan time its compiled variants on
    synthetic matrices of the same sizes and keep the fastest.  Two runs of
    one plan must not overlap.
    """

    def __init__(self, A, offsets, sizes, kinds, piv=None, piv_offsets=None, info=None, tune=True):
        import ctypes as C
        import json
        torch = _torch()
        if A.dim() != 1:
            raise ValueError("A must be a flat tensor")
        if sizes.dim() != 1:
            raise ValueError("sizes must be a flat tensor")
        count = int(sizes.numel())
        c = Call(sizes, "sizes", dtypes=(torch.int32,), square=False)
        c.device = A.device
        Call._dev(A, "A")
        if A.device != sizes.device:
            raise ValueError("A and sizes must be on one device")
        _dt(A)
        po = c.tensor(offsets, "offsets", torch.int64, (count,))
        pk = c.tensor(kinds, "kinds", torch.int32, (count,))
        pi = c.tensor(info, "info", torch.int32, (count,), optional=True)
        if (piv is None) != (piv_offsets is None) and piv is not None:
            raise ValueError("piv needs piv_offsets")
        pp = ppo = None
        if piv is not None:
            Call._dev(piv, "piv")
            if piv.dim() != 1 or piv.dtype != torch.int32 or piv.device != A.device:
                raise ValueError("piv must be a flat int32 tensor on A's device")
            pp = piv.data_ptr()
            ppo = c.tensor(piv_offsets, "piv_offsets", torch.int64, (count,))
        # the kernels trust offsets and sizes: check every matrix (and its
        # pivots) lies inside its buffer before the plan is built
        if count:
            n = sizes.to("cpu", torch.int64)
            if int(n.min()) < 1 or int(n.max()) > 64:
                raise ValueError("a matrix size is outside supported range [1, 64]")
            k = kinds.to("cpu")
            if bool(((k != 0) & (k != 1)).any()):
                raise ValueError("unknown kind: kinds must be 0 (Cholesky) or 1 (LU)")
            off = offsets.to("cpu")
            if int(off.min()) < 0 or int((off + n * n).max()) > A.numel():
                raise ValueError("a matrix given by offsets and sizes lies outside A")
            if piv is not None:
                lu = k == 1
                if bool(lu.any()):
                    poff = piv_offsets.to("cpu")[lu]
                    if int(poff.min()) < 0 or int((poff + n[lu]).max()) > piv.numel():
                        raise ValueError("the pivots given by piv_offsets and sizes lie outside piv")
        self._keep = (A, offsets, sizes, kinds, piv, piv_offsets, info)
        self._plan = C.c_void_p()
        _lib.call(f"tsdx_hip_ragged_plan_create_{_dt(A)}", A.data_ptr(), po, sizes.data_ptr(), pk, count, pp, ppo,
                  pi, 1 if tune else 0, C.byref(self._plan), _stream(A))
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
