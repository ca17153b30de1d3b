"""Mixed-precision posv with float64 refinement and float64 fallback.

Every returned solution must meet the normwise backward error bar 1e-14 of the
reference's fp64 solve tests (tests/cpp/test_solvers.cpp:189), judged here by
the long-double metric of the host API (reference runtime.cpp:230-242).
Matrices that refinement cannot handle (kappa * u_low too large, a failing
low-precision factorization) must come back through the float64 fallback
(stat == -1) and still meet the bar; a matrix that is not SPD in float64
reports info > 0 exactly like the plain posv.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(rng, batch, n, lo, hi, nrhs):
    from tinysolverdx import matrices
    kappa = 10.0 ** rng.uniform(lo, hi, batch)
    A = matrices.spd_kappas(rng, kappa, n)
    B = rng.standard_normal((batch, n, nrhs))
    return A, B, kappa


def _run(A, B, factor, max_iter=10, tol=1e-14):
    import torch
    from tinysolverdx import device
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    info = torch.full((A.shape[0],), -7, dtype=torch.int32, device="cuda")
    stat = torch.full((A.shape[0],), -7, dtype=torch.int32, device="cuda")
    device.posv_ir(dA, dB, dX, factor=factor, max_iter=max_iter, tol=tol, info=info, stat=stat)
    torch.cuda.synchronize()
    return dX.cpu().numpy(), info.cpu().numpy(), stat.cpu().numpy()


def _bwd(tsdx, A, X, B):
    return np.stack([tsdx.solve_backward_error(A, X[..., r], B[..., r]) for r in range(B.shape[-1])]).max(axis=0)


@pytest.mark.parametrize("factor", ["fp32", "bf16", "fp16", "fp64", "fp64_block"])
@pytest.mark.parametrize("n,nrhs", [(64, 1), (64, 4), (64, 8), (32, 2), (16, 8), (5, 1)])
def test_every_answer_meets_tolerance(tsdx, factor, n, nrhs):
    rng = np.random.default_rng(n * 10 + nrhs)
    A, B, kappa = _system(rng, 96, n, 1, 6, nrhs)
    X, info, stat = _run(A, B, factor)
    assert (info == 0).all()
    bw = _bwd(tsdx, A, X, B)
    assert bw.max() <= 1e-14, (factor, n, bw.max(), stat[bw.argmax()], kappa[bw.argmax()])
    if factor in ("fp64", "fp64_block"):
        assert (stat == -1).all()
    else:
        assert ((stat >= 0) & (stat <= 10) | (stat == -1)).all()
    if factor == "fp32" and n >= 16:
        # kappa <= 1e6 < 1/u_fp32: refinement converges for all of them
        assert (stat >= 0).mean() > 0.95, np.bincount(stat + 1)


def test_refinement_fails_over_to_fp64(tsdx):
    """kappa far beyond 1/u_bf16 cannot be refined from a bf16 factor; those
    matrices must take the float64 path and still meet the bar."""
    rng = np.random.default_rng(3)
    A, B, kappa = _system(rng, 64, 64, 5, 9, 1)
    X, info, stat = _run(A, B, "bf16")
    assert (stat == -1).sum() > 32, np.bincount(stat + 1)
    bw = _bwd(tsdx, A, X, B)
    assert bw.max() <= 1e-14


def test_not_spd_reports_info_like_plain_posv(tsdx, oracle):
    rng = np.random.default_rng(4)
    A, B, _ = _system(rng, 12, 20, 1, 3, 1)
    A[3] = -np.eye(20)
    A[7, 5, 5] = -3.0
    X, info, stat = _run(A, B, "fp32")
    _, _, info_o = oracle.posv(A, B)
    assert np.array_equal(info, info_o) and info[3] == 1 and info[7] > 0
    assert stat[3] == -1 and stat[7] == -1
    ok = info == 0
    assert _bwd(tsdx, A[ok], X[ok], B[ok]).max() <= 1e-14


def test_max_iter_zero_sends_unconverged_to_fp64(tsdx):
    rng = np.random.default_rng(5)
    A, B, _ = _system(rng, 40, 32, 3, 6, 1)
    X, info, stat = _run(A, B, "fp32", max_iter=0)
    assert (stat == -1).sum() > 30  # a float solve alone is not 1e-14 accurate
    assert _bwd(tsdx, A, X, B).max() <= 1e-14


def test_ir_argument_errors(tsdx):
    import ctypes as C
    from tinysolverdx import _lib
    L = _lib.lib()
    assert L.tsdx_hip_posv_ir_f64(None, 65, 1, None, None, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_posv_ir_f64(None, 8, 1, None, None, 1, 9, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_posv_ir_f64(None, 8, 0, None, None, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.OK


@pytest.mark.parametrize("factor", ["fp32", "bf16"])
def test_reads_lower_triangle_only(tsdx, factor):
    """Like the plain posv (and the reference's Cholesky), only the lower
    triangle of A is an input: garbage above the diagonal changes nothing."""
    rng = np.random.default_rng(6)
    A, B, _ = _system(rng, 32, 64, 1, 4, 2)
    G = A.copy()
    iu = np.triu_indices(64, 1)
    G[:, iu[0], iu[1]] = 1e30
    X, info, stat = _run(G, B, factor)
    Xr, info_r, stat_r = _run(A, B, factor)
    assert np.array_equal(X, Xr) and np.array_equal(stat, stat_r) and (info == 0).all()
    assert _bwd(tsdx, A, X, B).max() <= 1e-14


@pytest.mark.parametrize("n", [17, 33, 48, 63])
def test_padded_sizes(tsdx, n):
    """Sizes that leave part of the last 16 x 16 block as identity padding."""
    rng = np.random.default_rng(n)
    A, B, _ = _system(rng, 48, n, 1, 5, 3)
    X, info, stat = _run(A, B, "fp32")
    assert (info == 0).all() and (stat >= 0).mean() > 0.9, np.bincount(stat + 1)
    assert _bwd(tsdx, A, X, B).max() <= 1e-14


@pytest.mark.parametrize("n", [64, 40])
def test_fp64_block_tier_reports_info_like_plain_posv(tsdx, oracle, n):
    """For n > 32 the float64 tier is a block Cholesky on the matrix cores; a
    matrix it cannot factor (not SPD) is finished by the plain float64 posv in
    the same wave, so info is the reference's."""
    rng = np.random.default_rng(40 + n)
    A, B, _ = _system(rng, 16, n, 1, 3, 1)
    A[3] = -np.eye(n)
    A[7, n - 2, n - 2] = -3.0
    A[11, 20, 20] = 0.0
    A[11, 20, :20] = 0.0
    X, info, stat = _run(A, B, "fp32")
    _, _, info_o = oracle.posv(A, B)
    assert np.array_equal(info, info_o) and info[3] == 1 and info[7] > 0 and info[11] > 0
    assert stat[3] == -1 and stat[7] == -1 and stat[11] == -1
    ok = info == 0
    assert _bwd(tsdx, A[ok], X[ok], B[ok]).max() <= 1e-14


@pytest.mark.parametrize("n,nrhs", [(64, 3), (50, 1), (37, 5)])
def test_fp64_block_tier_meets_tolerance(tsdx, n, nrhs):
    """bf16 factors cannot be refined at kappa 1e5..1e10: most matrices go to
    the float64 tier, whose answers must still meet the bar for every
    right-hand side (kappa up to 1e10 also exercises its own refinement)."""
    rng = np.random.default_rng(70 + n)
    A, B, kappa = _system(rng, 96, n, 5, 10, nrhs)
    X, info, stat = _run(A, B, "bf16")
    assert (info == 0).all() and (stat == -1).sum() > 48, np.bincount(stat + 1)
    bw = _bwd(tsdx, A, X, B)
    assert bw.max() <= 1e-14, (bw.max(), kappa[bw.argmax()])


def test_fp64_block_tier_reads_lower_triangle_only(tsdx):
    rng = np.random.default_rng(8)
    A, B, _ = _system(rng, 32, 64, 6, 10, 2)
    G = A.copy()
    iu = np.triu_indices(64, 1)
    G[:, iu[0], iu[1]] = 1e30
    X, info, stat = _run(G, B, "bf16")
    Xr, info_r, stat_r = _run(A, B, "bf16")
    assert (stat == -1).sum() > 16
    assert np.array_equal(X, Xr) and np.array_equal(stat, stat_r) and (info == 0).all()
    assert _bwd(tsdx, A, X, B).max() <= 1e-14
