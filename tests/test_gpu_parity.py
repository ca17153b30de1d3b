"""Parity of the HIP path with the reference, through the C ABI.

Factors, pivots and info must be bit-identical to the golden fixtures written
by the compiled reference and to the oracle on fresh inputs (any two NaNs
count as equal).  Solutions agree to rounding: the backward substitution runs
as a column sweep, so the bar is the reference's own accuracy, measured as a
solve backward error below 8*n*u (fp32/fp64) plus closeness to the oracle's
solution within c*kappa*n*u.
"""
import numpy as np
import pytest

import hipcalls as H
from conftest import bits_equal, first_bit_mismatch, golden, golden_cases

pytestmark = pytest.mark.gpu

U = {np.dtype(np.float32): 2.0 ** -24, np.dtype(np.float64): 2.0 ** -53}


def _spd(rng, b, n, dt, kappa=1e3):
    from tinysolverdx import matrices
    return matrices.spd(rng, b, n, kappa).astype(dt)


def _solution_ok(A, X, B, Xref, dt):
    """Backward error of each column below 8 n u and agreement with Xref."""
    import tinysolverdx as tsdx
    n = A.shape[-1]
    u = U[np.dtype(dt)]
    A64 = A.astype(np.float64)
    for r in range(B.shape[-1]):
        bw = tsdx.solve_backward_error(A64, X[..., r].astype(np.float64), B[..., r].astype(np.float64))
        assert np.nanmax(bw) < 8 * n * u, (n, r, np.nanmax(bw))
    kappa = np.linalg.cond(A64)
    rel = np.linalg.norm(X - Xref, axis=1) / np.maximum(np.linalg.norm(Xref, axis=1), 1e-300)
    assert np.all(rel <= 64 * n * u * kappa[:, None] + 1e-300), (n, rel.max())


# -------------------------------------------------------------- golden fixtures
@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_potrf_golden_bit_exact(hip, sfx):
    g = golden(f"potrf_{sfx}")
    for c in golden_cases(g, "_A"):
        print(f"[potrf golden] {sfx} {c}", flush=True)
        L, info = H.cholesky(g[f"{c}_A"])
        assert np.array_equal(info, g[f"{c}_info"]), (c, info, g[f"{c}_info"])
        assert bits_equal(L, g[f"{c}_L"]), (c, first_bit_mismatch(L, g[f"{c}_L"]))


@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_getrf_golden_bit_exact(hip, sfx):
    g = golden(f"getrf_{sfx}")
    for c in golden_cases(g, "_A"):
        print(f"[getrf golden] {sfx} {c}", flush=True)
        LU, piv, info = H.lu(g[f"{c}_A"])
        assert np.array_equal(piv, g[f"{c}_piv"]), (c, first_bit_mismatch(piv, g[f"{c}_piv"]))
        assert np.array_equal(info, g[f"{c}_info"]), (c, info, g[f"{c}_info"])
        assert bits_equal(LU, g[f"{c}_LU"]), (c, first_bit_mismatch(LU, g[f"{c}_LU"]))


@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_solves_golden(hip, sfx):
    for kind in ("potrf", "getrf"):
        g = golden(f"{kind}_{sfx}")
        for c in golden_cases(g, "_A"):
            A = g[f"{c}_A"]
            if not np.isfinite(A).all() or A.shape[-1] < 1:
                continue
            if kind == "potrf":
                if (g[f"{c}_info"] != 0).any():
                    continue
                X = H.cholesky_solve(g[f"{c}_L"], g[f"{c}_B"])
            else:
                if (g[f"{c}_info"] != 0).any() or c.startswith("subnormal"):
                    continue
                X = H.lu_solve(g[f"{c}_LU"], g[f"{c}_piv"], g[f"{c}_B"])
            _solution_ok(A, X, g[f"{c}_B"], g[f"{c}_X"], A.dtype)


# ---------------------------------------------------------- oracle, all sizes
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_factor_all_sizes_vs_oracle(hip, oracle, dt):
    rng = np.random.default_rng(7)
    for n in range(1, 65):
        b = 67  # not a multiple of any group count: exercises the tail group
        print(f"[factor sweep] {np.dtype(dt).name} n={n}", flush=True)
        S = _spd(rng, b, n, dt)
        S[3] = -np.eye(n, dtype=dt)
        if n > 2:
            S[5, n // 2, n // 2] = -1.0
        L, info = H.cholesky(S)
        Lo, infoo = oracle.cholesky(S)
        assert np.array_equal(info, infoo), n
        ok = infoo == 0
        assert bits_equal(L[ok], Lo[ok]), (n, first_bit_mismatch(L[ok], Lo[ok]))
        G = rng.standard_normal((b, n, n)).astype(dt)
        G[2, :, n // 2] = 0
        G[4] = 0
        LU, piv, info = H.lu(G)
        LUo, pivo, infoo = oracle.lu(G)
        assert np.array_equal(piv, pivo), (n, first_bit_mismatch(piv, pivo))
        assert np.array_equal(info, infoo), n
        assert bits_equal(LU, LUo), (n, first_bit_mismatch(LU, LUo))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_failed_cholesky_matches_tile_semantics(hip, oracle, dt):
    """After a failure the reference's Auto (SIMD tile) kernel substitutes 1 and
    continues; the HIP kernel does the same, so even failed factors match."""
    rng = np.random.default_rng(8)
    for n in (5, 16, 33, 64):
        S = _spd(rng, 9, n, dt)
        S[:, n // 3, n // 3] = -2.0
        L, info = H.cholesky(S)
        Lo, infoo = oracle.cholesky(S)
        assert np.array_equal(info, infoo) and (info > 0).all()
        assert bits_equal(L, Lo), (n, first_bit_mismatch(L, Lo))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_fused_drivers_vs_oracle(hip, oracle, dt):
    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 7, 8, 12, 16, 17, 24, 25, 32, 33, 40, 48, 49, 63, 64):
        for nrhs in (1, 3, 4, 9):
            print(f"[fused] {np.dtype(dt).name} n={n} nrhs={nrhs}", flush=True)
            S = _spd(rng, 23, n, dt)
            G = rng.standard_normal((23, n, n)).astype(dt)
            B = rng.standard_normal((23, n, nrhs)).astype(dt)
            L, X, info = H.posv(S, B)
            Lo, Xo, infoo = oracle.posv(S, B)
            assert np.array_equal(info, infoo) and bits_equal(L, Lo), n
            _solution_ok(S, X, B, Xo, dt)
            LU, X, piv, info = H.gesv(G, B)
            LUo, Xo, pivo, infoo = oracle.gesv(G, B)
            assert np.array_equal(piv, pivo) and np.array_equal(info, infoo) and bits_equal(LU, LUo), n
            _solution_ok(G, X, B, Xo, dt)
            Xs = H.cholesky_solve(Lo, B)
            _solution_ok(S, Xs, B, oracle.cholesky_solve(Lo, B), dt)
            Xg = H.lu_solve(LUo, pivo, B)
            _solution_ok(G, Xg, B, oracle.lu_solve(LUo, pivo, B), dt)


def test_forward_pass_is_bit_exact(hip, oracle):
    """With an identity upper factor only the forward pass does arithmetic, and
    that pass keeps the reference's operation order: results are bit-equal."""
    rng = np.random.default_rng(10)
    for dt in (np.float32, np.float64):
        for n in (3, 16, 31, 64):
            LU = np.tril(rng.standard_normal((5, n, n)), -1).astype(dt) + np.eye(n, dtype=dt)
            piv = np.tile(np.arange(n, dtype=np.int32), (5, 1))
            B = rng.standard_normal((5, n, 4)).astype(dt)
            assert bits_equal(H.lu_solve(LU, piv, B), oracle.lu_solve(LU, piv, B)), (dt, n)


@pytest.mark.parametrize("n", [33, 40, 64])
def test_getrf_multirow_batches_bit_exact(hip, oracle, n):
    """fp32 getrf batches of >= 16384 matrices at 32 < n <= 64 run the
    two-rows-per-lane kernel (k_getrf_rows); factors, pivots and info must be
    the reference's, bit for bit, like the one-row kernel's: zero columns,
    a zero matrix, duplicate rows, NaN and inf entries and ties included."""
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(1000 + n)
    batch = 16384
    G = rng.standard_normal((batch, n, n)).astype(np.float32)
    zc = rng.choice(batch, 160, replace=False)
    G[zc, :, rng.integers(0, n, 160)] = 0.0
    G[7] = 0.0
    G[11, 5] = G[11, 2]
    G[13, 3, 4] = np.nan
    G[17, :, 0] = 1.0  # ties in the first column: the first row wins
    G[19, 9, 9] = np.inf
    G[23, n - 1] = -G[23, n - 1]
    print(f"[getrf multirow] n={n} batch={batch}", flush=True)
    dG = torch.from_numpy(G).cuda()
    piv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    device.getrf(dG, piv, info)
    LUo, pivo, io = oracle.lu(G)
    torch.cuda.synchronize()
    assert np.array_equal(info.cpu().numpy(), io)
    assert np.array_equal(piv.cpu().numpy(), pivo)
    assert bits_equal(dG.cpu().numpy(), LUo)


@pytest.mark.parametrize("dt,n,batch", [(np.float32, 8, 4099), (np.float32, 16, 16387), (np.float32, 32, 65537),
                                         (np.float32, 64, 1029), (np.float64, 8, 2051), (np.float64, 16, 4097),
                                         (np.float64, 32, 32771), (np.float64, 64, 8195)])
def test_getrf_full_bucket_batches_bit_exact(hip, oracle, dt, n, batch):
    """n equal to its bucket (8, 16, 32, 64): the one-row LU kernel writes the
    factor through LDS in line order; factors, pivots and info bit-exact,
    zero columns, a zero matrix, NaN, inf and ties included, and a batch that
    leaves the last wave partly empty between two guard regions that must
    come back untouched."""
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(3000 + n)
    G = rng.standard_normal((batch, n, n)).astype(dt)
    zc = rng.choice(batch, batch // 100, replace=False)
    G[zc, :, rng.integers(0, n, batch // 100)] = 0.0
    G[7] = 0.0
    G[11, 5] = G[11, 2]
    G[13, 3, 4] = np.nan
    G[17, :, 0] = 1.0
    G[19, n - 1, n - 1] = np.inf
    print(f"[getrf full bucket] {np.dtype(dt).name} n={n} batch={batch}", flush=True)
    tdt = torch.float32 if dt == np.float32 else torch.float64
    buf = torch.full((batch + 2, n, n), 7.0, dtype=tdt, device="cuda")
    dG = buf[1:batch + 1]
    dG.copy_(torch.from_numpy(G))
    piv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    device.getrf(dG, piv, info)
    LUo, pivo, io = oracle.lu(G)
    torch.cuda.synchronize()
    assert np.array_equal(info.cpu().numpy(), io)
    assert np.array_equal(piv.cpu().numpy(), pivo)
    assert bits_equal(dG.cpu().numpy(), LUo)
    assert bool((buf[0] == 7.0).all()) and bool((buf[batch + 1] == 7.0).all())


@pytest.mark.parametrize("dt,n", [(np.float32, 8), (np.float32, 16), (np.float32, 32), (np.float32, 64),
                                  (np.float64, 8), (np.float64, 16), (np.float64, 32), (np.float64, 64)])
def test_getrf_config1_full_scale_bit_exact(hip, oracle, dt, n):
    """The LU workload at its full size: ~1 GiB of N(0,1) matrices per (dtype,
    n) (fp32: 4194304 / 1048576 / 262144 / 65536 matrices for n = 8 / 16 / 32 /
    64, fp64 half as many), 1% of them with one zeroed column.  Pivots, info
    and factors bit-exact against the oracle for every matrix (pivot choices
    are sensitive to the fused multiply-add chain at about 1 in 4000 fp32
    n=64 matrices, so only the full batch shows them all); the fused gesv with
    4 right-hand sides reproduces the same pivots and info and solves to the
    reference's accuracy."""
    import torch
    from tinysolverdx import device
    es = np.dtype(dt).itemsize
    batch = (1 << 30) // (es * n * n)
    rng = np.random.default_rng(4000 + n + es)
    print(f"[getrf config-1 scale] {np.dtype(dt).name} n={n} batch={batch}", flush=True)
    G = rng.standard_normal((batch, n, n), dtype=np.float32 if dt == np.float32 else np.float64)
    zc = rng.choice(batch, batch // 100, replace=False)
    G[zc, :, rng.integers(0, n, zc.size)] = 0.0
    tdt = torch.float32 if dt == np.float32 else torch.float64
    dA0 = torch.from_numpy(G).cuda()
    dG = dA0.clone()
    piv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    device.getrf(dG, piv, info)
    torch.cuda.synchronize()
    LUo, pivo, io = oracle.lu(G)
    got_i, got_p = info.cpu().numpy(), piv.cpu().numpy()
    assert np.array_equal(got_i, io), int((got_i != io).sum())
    assert np.array_equal(got_p, pivo), int((got_p != pivo).any(axis=1).sum())
    assert bits_equal(dG.cpu().numpy(), LUo), first_bit_mismatch(dG.cpu().numpy(), LUo)
    assert (io > 0).sum() >= zc.size // 2
    del LUo, dG
    B = torch.randn((batch, n, 4), generator=torch.Generator(device="cuda").manual_seed(n), device="cuda", dtype=tdt)
    X = torch.empty_like(B)
    piv2 = torch.zeros_like(piv)
    info2 = torch.zeros_like(info)
    device.gesv_solve(dA0, B, X, piv2, info2)
    torch.cuda.synchronize()
    assert torch.equal(piv2, piv) and torch.equal(info2, info)
    ok = info == 0
    A64, X64, B64 = dA0[ok].double(), X[ok].double(), B[ok].double()
    r = torch.linalg.vector_norm(B64 - A64 @ X64, dim=1)
    den = torch.linalg.matrix_norm(A64)[:, None] * torch.linalg.vector_norm(X64, dim=1) + \
        torch.linalg.vector_norm(B64, dim=1)
    bw = (r / den).max().item()
    u = 2.0 ** -24 if dt == np.float32 else 2.0 ** -53
    assert bw < 8 * n * u, bw


@pytest.mark.parametrize("dt,n,batch", [(np.float32, 32, 65536), (np.float32, 13, 131072), (np.float32, 50, 16384),
                                         (np.float64, 29, 32768), (np.float32, 32, 65531), (np.float32, 16, 65539),
                                         (np.float64, 32, 32771), (np.float32, 64, 16387), (np.float32, 8, 8195),
                                         (np.float64, 8, 4097), (np.float64, 16, 4099)])
def test_potrf_persistent_batches_bit_exact(hip, oracle, dt, n, batch):
    """Batches of several rounds of waves of the multi-row Cholesky
    (k_potrf_rows; fp32 n=32 is config 0): factors and info bit-exact against
    the oracle, failing matrices (tile semantics) included.  The full-bucket
    sizes (8, 16, 32, 64) take the line-ordered stores through LDS or the
    lower-triangle loads; batches that are not a multiple of the matrices per
    wave leave the last wave partly empty, whose absent matrices must not be
    written."""
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(2000 + n)
    base = _spd(rng, 4096, n, dt)
    S = np.concatenate([base] * (batch // 4096 + 1))[:batch]
    S[5] = -np.eye(n, dtype=dt)
    S[batch - 3, n // 2, n // 2] = -1.0
    S[batch // 2, 1, 0] = np.nan
    print(f"[potrf persistent] {np.dtype(dt).name} n={n} batch={batch}", flush=True)
    # the batch sits between two guard regions that must come back untouched
    buf = torch.full((batch + 2, n, n), 7.0, dtype=torch.float32 if dt == np.float32 else torch.float64,
                     device="cuda")
    dS = buf[1:batch + 1]
    dS.copy_(torch.from_numpy(S))
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    device.potrf(dS, info)
    Lo, io = oracle.cholesky(S)
    torch.cuda.synchronize()
    assert np.array_equal(info.cpu().numpy(), io)
    assert bits_equal(dS.cpu().numpy(), Lo)
    assert bool((buf[0] == 7.0).all()) and bool((buf[batch + 1] == 7.0).all())


# ------------------------------------------------------------ device entry points
def test_device_entry_points_and_unaligned_rows(hip, oracle):
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(12)
    for dt, tdt in ((np.float32, torch.float32), (np.float64, torch.float64)):
        for n in (5, 8, 32, 33, 64):
            S = _spd(rng, 50, n, dt)
            G = rng.standard_normal((50, n, n)).astype(dt)
            dS = torch.from_numpy(S).cuda()
            info = torch.zeros(50, dtype=torch.int32, device="cuda")
            device.potrf(dS, info)
            Lo, io = oracle.cholesky(S)
            torch.cuda.synchronize()
            assert bits_equal(dS.cpu().numpy(), Lo) and np.array_equal(info.cpu().numpy(), io)
            # rows starting at an odd element offset force the element-wise load path
            buf = torch.zeros(G.size + 1, dtype=tdt, device="cuda")
            dG = buf[1:].view(50, n, n)
            dG.copy_(torch.from_numpy(G))
            piv = torch.zeros((50, n), dtype=torch.int32, device="cuda")
            device.getrf(dG, piv, info)
            LUo, pivo, io = oracle.lu(G)
            torch.cuda.synchronize()
            assert bits_equal(dG.cpu().numpy(), LUo) and np.array_equal(piv.cpu().numpy(), pivo)
            # solve-only drivers leave A and B untouched
            B = rng.standard_normal((50, n, 2)).astype(dt)
            dA, dB = torch.from_numpy(S).cuda(), torch.from_numpy(B).cuda()
            dX = torch.empty_like(dB)
            device.posv_solve(dA, dB, dX, info)
            torch.cuda.synchronize()
            assert bits_equal(dA.cpu().numpy(), S) and bits_equal(dB.cpu().numpy(), B)
            _solution_ok(S, dX.cpu().numpy(), B, oracle.posv(S, B)[1], dt)
            dA = torch.from_numpy(G).cuda()
            device.gesv_solve(dA, dB, dX, None, info)
            torch.cuda.synchronize()
            assert bits_equal(dA.cpu().numpy(), G)
            _solution_ok(G, dX.cpu().numpy(), B, oracle.gesv(G, B)[1], dt)


def test_empty_and_null_info(hip):
    import ctypes as C
    from tinysolverdx import _lib
    A = np.eye(4, dtype=np.float32)[None].copy()
    assert _lib.lib().tsdx_hip_cholesky_batched_f32(A.ctypes.data_as(C.c_void_p), 4, 1, None) == 0
    np.testing.assert_array_equal(A[0], np.eye(4))
    assert _lib.lib().tsdx_hip_cholesky_batched_f32(A.ctypes.data_as(C.c_void_p), 4, 0, None) == 0
