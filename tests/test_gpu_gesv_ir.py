"""Mixed-precision gesv with float64 refinement and the float64 gesv behind it.

Every returned solution must meet the normwise backward error bar 1e-14 of the
reference's fp64 solve tests (tests/cpp/test_solvers.cpp:189), judged by the
long-double metric of the host API (reference runtime.cpp:230-242).  Matrices
that refinement cannot handle (kappa * u_low too large, a zero pivot in the
low-precision LU, max_iter too small) must come back through the float64 gesv
(stat == -1): their X is then the plain float64 gesv's, bit for bit, and info
is the reference's (checked against the oracle), singular matrices included.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(rng, batch, n, lo, hi, nrhs):
    from tinysolverdx import matrices
    kappa = 10.0 ** rng.uniform(lo, hi, batch)
    A = matrices.general_kappas(rng, kappa, n)
    B = rng.standard_normal((batch, n, nrhs))
    return A, B, kappa


def _run(A, B, factor, max_iter=10, tol=1e-14):
    import torch
    from tinysolverdx import device
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    info = torch.full((A.shape[0],), -7, dtype=torch.int32, device="cuda")
    stat = torch.full((A.shape[0],), -7, dtype=torch.int32, device="cuda")
    device.gesv_ir(dA, dB, dX, factor=factor, max_iter=max_iter, tol=tol, info=info, stat=stat)
    torch.cuda.synchronize()
    return dX.cpu().numpy(), info.cpu().numpy(), stat.cpu().numpy()


def _plain(A, B):
    """The plain float64 gesv on the device (same kernels' chain as the list)."""
    import torch
    from tinysolverdx import device
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    info = torch.zeros(A.shape[0], dtype=torch.int32, device="cuda")
    device.gesv_solve(dA, dB, dX, info=info)
    torch.cuda.synchronize()
    return dX.cpu().numpy(), info.cpu().numpy()


def _bwd(tsdx, A, X, B):
    return np.stack([tsdx.solve_backward_error(A, X[..., r], B[..., r]) for r in range(B.shape[-1])]).max(axis=0)


@pytest.mark.parametrize("factor", ["fp32", "bf16", "fp16", "fp64"])
@pytest.mark.parametrize("n,nrhs", [(64, 1), (64, 4), (48, 2), (32, 2), (24, 3), (16, 8), (5, 1), (1, 1)])
def test_every_answer_meets_tolerance(tsdx, factor, n, nrhs):
    rng = np.random.default_rng(n * 10 + nrhs)
    A, B, kappa = _system(rng, 96, n, 1, 6, nrhs)
    X, info, stat = _run(A, B, factor)
    assert (info == 0).all()
    bw = _bwd(tsdx, A, X, B)
    assert bw.max() <= 1e-14, (factor, n, bw.max(), stat[bw.argmax()], kappa[bw.argmax()])
    if factor == "fp64":
        assert (stat == -1).all()
    else:
        assert ((stat >= 0) & (stat <= 10) | (stat == -1)).all()
    if factor == "fp32" and n >= 16:
        # kappa <= 1e6 < 1/u_fp32: refinement converges for nearly all of them
        assert (stat >= 0).mean() > 0.9, np.bincount(stat + 1)
    # the float64 answers are the plain gesv's
    Xp, info_p = _plain(A, B)
    f = stat == -1
    assert np.array_equal(X[f], Xp[f]) and np.array_equal(info, info_p)


def test_refinement_fails_over_to_fp64(tsdx):
    """kappa far beyond 1/u_bf16 cannot be refined from a bf16 LU: those
    matrices must take the float64 gesv and still meet the bar."""
    rng = np.random.default_rng(3)
    A, B, kappa = _system(rng, 64, 64, 5, 9, 1)
    X, info, stat = _run(A, B, "bf16")
    assert (stat == -1).sum() > 32, np.bincount(stat + 1)
    assert _bwd(tsdx, A, X, B).max() <= 1e-14


@pytest.mark.parametrize("n", [8, 20, 30, 64])
def test_singular_matrices_report_info_like_plain_gesv(tsdx, oracle, n):
    """Zero columns make the low-precision LU meet a zero pivot: the float64
    gesv takes over and gives the reference's info (the zero-pivot rule of
    n <= 24 continues past the pivot, larger n skip the column)."""
    rng = np.random.default_rng(90 + n)
    A, B, _ = _system(rng, 12, n, 1, 3, 2)
    A[2][:, 0] = 0.0
    A[5][:, n // 2] = 0.0
    A[9] = 0.0
    X, info, stat = _run(A, B, "fp32")
    _, Xo, _, info_o = oracle.gesv(A, B)
    assert np.array_equal(info, info_o), (info, info_o)
    assert info[2] == 1 and info[5] == n // 2 + 1 and info[9] == 1
    assert stat[2] == stat[5] == stat[9] == -1
    Xp, _ = _plain(A, B)
    f = stat == -1
    assert np.array_equal(X[f], Xp[f], equal_nan=True)
    ok = info == 0
    assert _bwd(tsdx, A[ok], X[ok], B[ok]).max() <= 1e-14


def test_nan_and_overflow_go_to_fp64(tsdx):
    """A NaN entry, and entries beyond the half range (65504), cannot be
    refined from a low-precision LU: the float64 gesv answers them."""
    rng = np.random.default_rng(11)
    A, B, _ = _system(rng, 16, 12, 1, 3, 1)
    A[3, 4, 7] = np.nan
    A[6] *= 1e8
    X, info, stat = _run(A, B, "fp16")
    assert stat[3] == -1 and stat[6] == -1
    Xp, info_p = _plain(A, B)
    f = stat == -1
    assert np.array_equal(X[f], Xp[f], equal_nan=True) and np.array_equal(info, info_p)
    ok = np.ones(16, bool)
    ok[3] = False
    assert _bwd(tsdx, A[ok], X[ok], B[ok]).max() <= 1e-14


def test_max_iter_zero_sends_unconverged_to_fp64(tsdx):
    rng = np.random.default_rng(5)
    A, B, _ = _system(rng, 40, 32, 3, 6, 1)
    X, info, stat = _run(A, B, "fp32", max_iter=0)
    assert (stat == -1).sum() > 30  # a float solve alone is not 1e-14 accurate
    assert _bwd(tsdx, A, X, B).max() <= 1e-14


@pytest.mark.parametrize("n", [3, 8, 16, 64])
def test_large_batch(tsdx, n):
    """Many workgroups (and groups of several matrices per wave for n <= 16):
    a sample of the answers meets the bar, nearly all by refinement."""
    rng = np.random.default_rng(200 + n)
    batch = 40000 if n <= 16 else 8000
    A, B, _ = _system(rng, batch, n, 0, 4, 1)
    X, info, stat = _run(A, B, "fp32")
    assert (info == 0).all() and (stat >= 0).mean() > 0.97, np.bincount(stat + 1)
    pick = rng.choice(batch, 400, replace=False)
    assert _bwd(tsdx, A[pick], X[pick], B[pick]).max() <= 1e-14
    assert np.isfinite(X).all()


def test_inputs_are_only_read(tsdx):
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(12)
    A, B, _ = _system(rng, 20, 40, 1, 4, 2)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    device.gesv_ir(dA, dB, dX)
    torch.cuda.synchronize()
    assert np.array_equal(dA.cpu().numpy(), A) and np.array_equal(dB.cpu().numpy(), B)
    assert _bwd(tsdx, A, dX.cpu().numpy(), B).max() <= 1e-14
