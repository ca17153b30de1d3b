"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that holds the reference sources (they are compiled, where
they lie, by `make -C oracle ref` into oracle/_ref/libtsdx_ref.so):

    make -C oracle ref && python tests/golden/make_golden.py [--only emulated]

Every fixture stores the raw input bytes together with the reference's outputs
(factors, pivots, info, solutions, metric values), so the tests never need to
regenerate inputs (numpy's QR/BLAS bits differ between builds).  The reference
runs with backend "auto" (its default: the SIMD tile kernels for Cholesky, and
for LU up to n = 24, the scalar kernels for LU from n = 25) and, where failure
semantics differ, with "scalar" as well.
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
from oracle import Oracle  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent
SIZES = [1, 2, 3, 4, 5, 7, 8, 9, 12, 15, 16, 17, 20, 24, 25, 31, 32, 33, 40, 48, 63, 64]


def batch_for(n):
    return int(max(2, min(11, 3000 // (n * n))))


def orthogonal(rng, batch, n):
    q, r = np.linalg.qr(rng.standard_normal((batch, n, n)))
    return q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]


def spd(rng, batch, n, kappa):
    lam = np.geomspace(1.0, 1.0 / kappa, n)
    q = orthogonal(rng, batch, n)
    a = np.einsum("bij,j,bkj->bik", q, lam, q)
    return (a + a.transpose(0, 2, 1)) / 2


def factor_cases(rng, dt):
    """(name, A) pairs for potrf and getrf."""
    chol, lu = [], []
    for n in SIZES:
        b = batch_for(n)
        S = spd(rng, b, n, 1e3).astype(dt)
        if n >= 3 and b >= 3:
            S[1] = S[1].copy()
            S[1, 2, 2] = -1.0                      # pivot 3 (or earlier) breaks
            S[2] = 0                               # zero matrix: info 1
        chol.append((f"n{n}", S))
        G = rng.standard_normal((b, n, n)).astype(dt)
        if b >= 3 and n >= 2:
            j = int(rng.integers(0, n))
            G[1, :, j] = 0                          # zero column j: info j+1
            G[2] = 0                                # zero matrix: info 1
        lu.append((f"n{n}", G))
    # closed forms and edge values
    chol.append(("kat2", np.array([[[4.0, 2.0], [2.0, 3.0]]], dt)))
    chol.append(("negI5", -np.eye(5, dtype=dt)[None]))
    nanm = spd(rng, 2, 6, 10.0).astype(dt)
    nanm[0, 3, 1] = np.nan
    nanm[1, 4, 4] = np.nan
    chol.append(("nan6", nanm))
    lu.append(("kat2", np.array([[[0.0, 1.0], [1.0, 0.0]]], dt)))
    dup = rng.standard_normal((3, 4, 4)).astype(dt)
    dup[0, 2] = dup[0, 1]
    lu.append(("duprow4", dup))
    ties = rng.integers(-3, 4, size=(8, 9, 9)).astype(dt)   # many equal magnitudes
    ties[3, :, 0] = 2.0
    ties[4, :, 0] = -2.0
    ties[5, 0, 0] = -0.0
    lu.append(("ties9", ties))
    tiesb = rng.integers(-2, 3, size=(4, 33, 33)).astype(dt)
    lu.append(("ties33", tiesb))
    nanl = rng.standard_normal((4, 7, 7)).astype(dt)
    nanl[0, 0, 0] = np.nan                          # NaN diagonal is sticky
    nanl[1, 3, 0] = np.nan                          # NaN never wins
    nanl[2, :, 2] = np.nan
    nanl[3, 1, 1] = np.inf
    lu.append(("nan7", nanl))
    nanl33 = rng.standard_normal((2, 33, 33)).astype(dt)
    nanl33[0, 5, 0] = np.nan
    nanl33[1, 0, 0] = np.nan
    lu.append(("nan33", nanl33))
    sub = (rng.standard_normal((3, 8, 8)) * (1e-310 if dt == np.float64 else 1e-40)).astype(dt)
    lu.append(("subnormal8", sub))
    return chol, lu


def emulated(R):
    """Runtime-precision calls in fp16 / bf16 (the reference's Emu types under
    its Auto backend, i.e. the scalar kernels) on float32 and float64 storage:
    factors with their failure content, pivots, info and both solves."""
    rng = np.random.default_rng(20261204)
    out = {}
    for prec in ("fp16", "bf16"):
        for dt, sfx, sizes in ((np.float32, "f32", (1, 2, 3, 5, 8, 13, 16, 24, 33, 64)),
                               (np.float64, "f64", (3, 16, 40))):
            for n in sizes:
                b = 5
                S = spd(rng, b, n, 1e2).astype(dt)
                if n >= 3:
                    S[1, 2, 2] = -1.0                  # fails at pivot 3 or earlier
                S[2] *= 3e4                            # fp16: entries near the top of the range
                G = rng.standard_normal((b, n, n)).astype(dt)
                if n >= 2:
                    G[3, :, n // 2] = 0                # zero pivot: info n//2 + 1, column skipped
                G[4, 0, 0] = np.nan
                B = rng.standard_normal((b, n, 3)).astype(dt)
                key = f"{prec}_{sfx}_n{n}"
                L, info = R.cholesky_p(S, prec)
                LU, piv, linfo = R.lu_p(G, prec)
                out.update({f"{key}_S": S, f"{key}_L": L, f"{key}_info": info,
                            f"{key}_G": G, f"{key}_LU": LU, f"{key}_piv": piv, f"{key}_linfo": linfo,
                            f"{key}_B": B,
                            f"{key}_X": R.cholesky_solve_p(L, B, prec),
                            f"{key}_Y": R.lu_solve_p(LU, piv, B, prec)})
    np.savez_compressed(OUT / "emulated.npz", **out)


def main():
    if sys.argv[1:] == ["--only", "emulated"]:
        emulated(Oracle("reference"))
        return
    R = Oracle("reference")
    rng = np.random.default_rng(20261203)
    for dt, sfx in ((np.float32, "f32"), (np.float64, "f64")):
        chol, lu = factor_cases(rng, dt)
        out = {}
        for name, A in chol:
            L, info = R.cholesky(A, "auto")
            Ls, infos = R.cholesky(A, "scalar")
            out[f"{name}_A"], out[f"{name}_L"], out[f"{name}_info"] = A, L, info
            out[f"{name}_Lscalar"], out[f"{name}_infoscalar"] = Ls, infos
            n = A.shape[-1]
            B = rng.standard_normal((A.shape[0], n, 3)).astype(dt)
            out[f"{name}_B"] = B
            out[f"{name}_X"] = R.cholesky_solve(L, B)
        np.savez_compressed(OUT / f"potrf_{sfx}.npz", **out)
        out = {}
        for name, A in lu:
            LU, piv, info = R.lu(A, "auto")
            out[f"{name}_A"], out[f"{name}_LU"], out[f"{name}_piv"], out[f"{name}_info"] = A, LU, piv, info
            if A.shape[-1] <= 24:
                LUs, pivs, infos = R.lu(A, "scalar")
                out[f"{name}_LUscalar"], out[f"{name}_pivscalar"], out[f"{name}_infoscalar"] = LUs, pivs, infos
            n = A.shape[-1]
            B = rng.standard_normal((A.shape[0], n, 4)).astype(dt)
            out[f"{name}_B"] = B
            out[f"{name}_X"] = R.lu_solve(LU, piv, B)
        np.savez_compressed(OUT / f"getrf_{sfx}.npz", **out)

    # long-double metrics on small fp64 systems
    out = {}
    for n in (2, 5, 10, 16, 33):
        A = rng.standard_normal((4, n, n))
        S = spd(rng, 4, n, 1e4)
        b = rng.standard_normal((4, n))
        LU, piv, _ = R.lu(A)
        L, _ = R.cholesky(S)
        x = R.reference_solve(A, b)
        xp = x * (1 + 1e-3 * rng.standard_normal(x.shape))
        out.update({f"n{n}_A": A, f"n{n}_S": S, f"n{n}_b": b, f"n{n}_LU": LU, f"n{n}_piv": piv, f"n{n}_L": L,
                    f"n{n}_x": x, f"n{n}_xp": xp,
                    f"n{n}_lu_bwd": R.lu_backward_error(A, LU, piv),
                    f"n{n}_chol_bwd": R.chol_backward_error(S, L),
                    f"n{n}_solve_bwd": R.solve_backward_error(A, xp, b),
                    f"n{n}_solve_fwd": R.solve_forward_error(A, xp, b)})
    np.savez_compressed(OUT / "metrics_f64.npz", **out)

    # fp16 / bf16 rounding and round_to
    vals = np.concatenate([
        np.array([1.0, 1 + 1 / 2048, 1 + 3 / 2048, 65504.0, 65519.0, 65520.0, 6.0e-8, 2.9e-8, 3.0e-8, 1e-9, -0.1,
                  3.14159265, np.finfo(np.float32).max, -np.finfo(np.float32).max, 0.0, -0.0, np.inf, -np.inf, np.nan,
                  6.103515625e-05, 6.1e-05, 1 + 1 / 256, 1 + 3 / 256], np.float32),
        (rng.standard_normal(2000) * np.exp(rng.uniform(-30, 30, 2000))).astype(np.float32)])
    out = {"x": vals,
           "fp16": np.array([R.round_fp16(v) for v in vals], np.float32),
           "bf16": np.array([R.round_bf16(v) for v in vals], np.float32),
           "u": np.array([R.unit_roundoff(p) for p in ("fp32", "fp64", "fp16", "bf16")])}
    x64 = rng.standard_normal(500) * np.exp(rng.uniform(-20, 20, 500))
    out["x64"] = x64
    for p in ("fp32", "fp16", "bf16", "fp64"):
        out[f"round_to_{p}"] = R.round_to(x64, p)
    np.savez_compressed(OUT / "precision.npz", **out)
    emulated(R)
    for f in sorted(OUT.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
