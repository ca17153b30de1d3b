"""ctypes access to the parity oracle (liboracle.so) and, where it was built,
to the compiled reference (_ref/libtsdx_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.

Both libraries expose the same call shapes, so `Oracle("port")` and
`Oracle("reference")` are interchangeable in tests.  Backend codes follow the
reference enum (precision.hpp:33): auto 0, naive 1, scalar 2, simd 3, cuda 4.
"""
import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PORT = HERE / "liboracle.so"
LIB_REF = HERE / "_ref" / "libtsdx_ref.so"

BACKENDS = {"auto": 0, "naive": 1, "scalar": 2, "simd": 3, "cuda": 4}
PRECISIONS = {"fp32": 0, "fp64": 1, "fp16": 2, "bf16": 3}

EINVAL, EUNAVAIL = -1, -2


class OracleError(RuntimeError):
    pass


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _sfx(dtype):
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}[np.dtype(dtype)]


class Oracle:
    """kind = "port" (the C restatement) or "reference" (compiled reference)."""

    def __init__(self, kind="port"):
        self.kind = kind
        path = LIB_PORT if kind == "port" else LIB_REF
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle{' ref' if kind != 'port' else ''})")
        self.lib = C.CDLL(str(path))
        self.pre = "orc_" if kind == "port" else "ref_"
        i64 = C.c_longlong
        vp, ci = C.c_void_p, C.c_int
        for sfx in ("f32", "f64"):
            self._sig(f"cholesky_batched_{sfx}", [vp, ci, i64, vp, ci])
            self._sig(f"lu_batched_{sfx}", [vp, ci, i64, vp, vp, ci])
            self._sig(f"chol_solve_batched_{sfx}", [vp, ci, i64, vp, ci])
            self._sig(f"lu_solve_batched_{sfx}", [vp, vp, ci, i64, vp, ci])
            self._sig(f"posv_batched_{sfx}", [vp, ci, i64, vp, ci, vp, ci])
            self._sig(f"gesv_batched_{sfx}", [vp, ci, i64, vp, ci, vp, vp, ci])
        for m in ("chol_backward_error",):
            self._sig(m, [vp, vp, ci, i64, vp], None)
        self._sig("lu_backward_error", [vp, vp, vp, ci, i64, vp], None)
        self._sig("solve_backward_error", [vp, vp, vp, ci, i64, vp], None)
        self._sig("solve_forward_error", [vp, vp, vp, ci, i64, vp], None)
        self._sig("reference_solve", [vp, vp, ci, i64, vp], None)
        self._sig("round_fp16", [C.c_float], C.c_float)
        self._sig("round_bf16", [C.c_float], C.c_float)
        self._sig("unit_roundoff", [ci], C.c_double)
        self._sig("set_num_threads", [ci], None)
        self._sig("num_threads", [], ci)
        for sfx in ("f32", "f64"):
            if kind == "port":
                self._sig(f"emu_cholesky_{sfx}", [ci, vp, ci, i64, vp])
                self._sig(f"emu_lu_{sfx}", [ci, vp, ci, i64, vp, vp])
                self._sig(f"emu_chol_solve_{sfx}", [ci, vp, ci, i64, vp, ci])
                self._sig(f"emu_lu_solve_{sfx}", [ci, vp, vp, ci, i64, vp, ci])
            else:
                self._sig(f"cholesky_{sfx}", [ci, vp, ci, i64, vp, ci])
                self._sig(f"lu_{sfx}", [ci, vp, ci, i64, vp, vp, ci])
                self._sig(f"cholesky_solve_{sfx}", [ci, vp, ci, i64, vp, ci])
                self._sig(f"lu_solve_{sfx}", [ci, vp, vp, ci, i64, vp, ci])
        if kind == "port":
            self._sig("round_to_f64", [ci, vp, i64], None)
            self._sig("chol_rm_f64", [vp, ci])
            self._sig("lu_rm_f64", [vp, ci, vp])
        else:
            self._sig("round_to_f64", [ci, vp, i64])

    def _sig(self, name, args, res=C.c_int):
        f = getattr(self.lib, self.pre + name)
        f.argtypes = args
        f.restype = res

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    @staticmethod
    def _check(rc):
        if rc == EINVAL:
            raise ValueError("invalid argument")
        if rc == EUNAVAIL:
            raise OracleError("backend unavailable")
        return rc

    # ---- factorizations (copies; returns new arrays) ----
    def cholesky(self, A, backend="auto"):
        A = np.array(A, copy=True, order="C")
        batch, n = A.shape[0], A.shape[-1]
        info = np.zeros(batch, np.int32)
        self._check(self._f(f"cholesky_batched_{_sfx(A.dtype)}")(_p(A), n, batch, _p(info), BACKENDS[backend]))
        return A, info

    def lu(self, A, backend="auto"):
        A = np.array(A, copy=True, order="C")
        batch, n = A.shape[0], A.shape[-1]
        piv = np.zeros((batch, n), np.int32)
        info = np.zeros(batch, np.int32)
        self._check(self._f(f"lu_batched_{_sfx(A.dtype)}")(_p(A), n, batch, _p(piv), _p(info), BACKENDS[backend]))
        return A, piv, info

    def cholesky_solve(self, L, B):
        L = np.ascontiguousarray(L)
        B = np.array(B, copy=True, order="C", dtype=L.dtype)
        self._check(self._f(f"chol_solve_batched_{_sfx(L.dtype)}")(_p(L), L.shape[-1], L.shape[0], _p(B), B.shape[-1]))
        return B

    def lu_solve(self, LU, piv, B):
        LU = np.ascontiguousarray(LU)
        piv = np.ascontiguousarray(piv, dtype=np.int32)
        B = np.array(B, copy=True, order="C", dtype=LU.dtype)
        self._check(self._f(f"lu_solve_batched_{_sfx(LU.dtype)}")(_p(LU), _p(piv), LU.shape[-1], LU.shape[0], _p(B), B.shape[-1]))
        return B

    def posv(self, A, B, backend="auto"):
        A = np.array(A, copy=True, order="C")
        B = np.array(B, copy=True, order="C", dtype=A.dtype)
        info = np.zeros(A.shape[0], np.int32)
        self._check(self._f(f"posv_batched_{_sfx(A.dtype)}")(_p(A), A.shape[-1], A.shape[0], _p(B), B.shape[-1], _p(info), BACKENDS[backend]))
        return A, B, info

    def gesv(self, A, B, backend="auto"):
        A = np.array(A, copy=True, order="C")
        B = np.array(B, copy=True, order="C", dtype=A.dtype)
        piv = np.zeros(A.shape[:2], np.int32)
        info = np.zeros(A.shape[0], np.int32)
        self._check(self._f(f"gesv_batched_{_sfx(A.dtype)}")(_p(A), A.shape[-1], A.shape[0], _p(B), B.shape[-1], _p(piv), _p(info), BACKENDS[backend]))
        return A, B, piv, info

    # ---- runtime precision (runtime.cpp:92-140): emulated fp16 / bf16 ----
    # The port restates only the emulated formats; the reference takes all four.
    def _emu(self, port_name, ref_name, prec, *args):
        if self.kind == "port":
            if prec not in ("fp16", "bf16"):
                raise ValueError("the port restates the emulated precisions fp16 / bf16 only")
            return self._check(self._f(port_name)(PRECISIONS[prec], *args))
        return self._check(self._f(ref_name)(PRECISIONS[prec], *args))

    def cholesky_p(self, A, prec):
        """tsdx::cholesky(precision, ...) on a copy of A (float32 or float64 storage)."""
        A = np.array(A, copy=True, order="C")
        batch, n = A.shape[0], A.shape[-1]
        info = np.zeros(batch, np.int32)
        extra = () if self.kind == "port" else (BACKENDS["auto"],)
        sfx = _sfx(A.dtype)
        self._emu(f"emu_cholesky_{sfx}", f"cholesky_{sfx}", prec, _p(A), n, batch, _p(info), *extra)
        return A, info

    def lu_p(self, A, prec):
        A = np.array(A, copy=True, order="C")
        batch, n = A.shape[0], A.shape[-1]
        piv = np.zeros((batch, n), np.int32)
        info = np.zeros(batch, np.int32)
        extra = () if self.kind == "port" else (BACKENDS["auto"],)
        sfx = _sfx(A.dtype)
        self._emu(f"emu_lu_{sfx}", f"lu_{sfx}", prec, _p(A), n, batch, _p(piv), _p(info), *extra)
        return A, piv, info

    def cholesky_solve_p(self, L, B, prec):
        L = np.ascontiguousarray(L)
        B = np.array(B, copy=True, order="C", dtype=L.dtype)
        sfx = _sfx(L.dtype)
        self._emu(f"emu_chol_solve_{sfx}", f"cholesky_solve_{sfx}", prec, _p(L), L.shape[-1], L.shape[0], _p(B),
                  B.shape[-1])
        return B

    def lu_solve_p(self, LU, piv, B, prec):
        LU = np.ascontiguousarray(LU)
        piv = np.ascontiguousarray(piv, dtype=np.int32)
        B = np.array(B, copy=True, order="C", dtype=LU.dtype)
        sfx = _sfx(LU.dtype)
        self._emu(f"emu_lu_solve_{sfx}", f"lu_solve_{sfx}", prec, _p(LU), _p(piv), LU.shape[-1], LU.shape[0],
                  _p(B), B.shape[-1])
        return B

    # ---- precision ----
    def round_fp16(self, x):
        return np.float32(self._f("round_fp16")(float(x)))

    def round_bf16(self, x):
        return np.float32(self._f("round_bf16")(float(x)))

    def unit_roundoff(self, p):
        return self._f("unit_roundoff")(PRECISIONS[p])

    def round_to(self, x, p):
        x = np.array(x, copy=True, order="C", dtype=np.float64)
        self._f("round_to_f64")(PRECISIONS[p], _p(x), x.size)
        return x

    def set_num_threads(self, n):
        self._f("set_num_threads")(n)

    def num_threads(self):
        return self._f("num_threads")()

    # ---- long-double metrics (double in, one value per matrix) ----
    def _metric(self, name, A, *rest):
        A = np.ascontiguousarray(A, dtype=np.float64)
        out = np.zeros(A.shape[0])
        args = [_p(A)] + [_p(np.ascontiguousarray(r)) for r in rest]
        self._f(name)(*args, A.shape[-1], A.shape[0], _p(out))
        return out

    def chol_backward_error(self, A, L):
        return self._metric("chol_backward_error", A, np.asarray(L, np.float64))

    def lu_backward_error(self, A, LU, piv):
        return self._metric("lu_backward_error", A, np.asarray(LU, np.float64), np.asarray(piv, np.int32))

    def solve_backward_error(self, A, x, b):
        return self._metric("solve_backward_error", A, np.asarray(x, np.float64), np.asarray(b, np.float64))

    def solve_forward_error(self, A, x, b):
        return self._metric("solve_forward_error", A, np.asarray(x, np.float64), np.asarray(b, np.float64))

    def reference_solve(self, A, b):
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        self._f("reference_solve")(_p(A), _p(b), A.shape[-1], A.shape[0], _p(x))
        return x


def available(kind):
    return (LIB_PORT if kind == "port" else LIB_REF).exists()
