"""The reference's Python test-suite (tests/python/test_solvers.py) run against
this package with the HIP backend ("auto" resolves to "hip")."""
import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu

SIZES = [2, 3, 5, 8, 16, 17, 32, 64]
BACKENDS = ["auto", "hip"]


def rng(seed=0):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-4)])
def test_cholesky_matches_numpy(tsdx, n, backend, dtype, tol):
    from tinysolverdx import matrices
    A = matrices.spd(rng(n), 21, n, 1e2).astype(dtype)
    L, info = tsdx.cholesky(A, backend=backend)
    assert L.dtype == dtype and (info == 0).all()
    np.testing.assert_allclose(L, np.linalg.cholesky(A.astype(np.float64)), atol=tol, rtol=tol)
    assert np.allclose(np.triu(L, 1), 0)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("backend", BACKENDS)
def test_lu_matches_scipy(tsdx, n, backend):
    from tinysolverdx import matrices
    A = matrices.general(rng(n), 13, n, 1e3)
    LU, piv, info = tsdx.lu(A, backend=backend)
    assert (info == 0).all()
    for b in range(A.shape[0]):
        ref_lu, ref_piv = sla.lu_factor(A[b])
        np.testing.assert_array_equal(piv[b], ref_piv)
        np.testing.assert_allclose(LU[b], ref_lu, atol=1e-10, rtol=1e-10)


@pytest.mark.parametrize("n", [2, 7, 32, 64])
def test_solvers_match_numpy_solve(tsdx, n):
    from tinysolverdx import matrices
    r = rng(1)
    A = matrices.general(r, 9, n, 1e3)
    S = matrices.spd(r, 9, n, 1e3)
    b = r.standard_normal((9, n))
    x, info = tsdx.gesv(A, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-8, atol=1e-8)
    x, info = tsdx.posv(S, b)
    assert (info == 0).all()
    np.testing.assert_allclose(x, np.linalg.solve(S, b[..., None])[..., 0], rtol=1e-8, atol=1e-8)
    B = r.standard_normal((9, n, 3))
    x, _ = tsdx.gesv(A, B)
    np.testing.assert_allclose(A @ x, B, atol=1e-9)
    LU, piv, _ = tsdx.lu(A)
    np.testing.assert_allclose(tsdx.lu_solve(LU, piv, B), x, rtol=1e-8, atol=1e-8)
    L, _ = tsdx.cholesky(S)
    np.testing.assert_allclose(tsdx.cholesky_solve(L, B), np.linalg.solve(S, B), rtol=1e-8, atol=1e-8)


def test_single_matrix_and_vector_rhs(tsdx):
    from tinysolverdx import matrices
    A = matrices.spd(rng(), 1, 6, 10.0)[0]
    b = np.arange(6.0)
    x, info = tsdx.posv(A, b)
    assert x.shape == (6,) and info == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-12)
    x, info = tsdx.gesv(A, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-12)


def test_input_not_modified_and_noncontiguous(tsdx):
    from tinysolverdx import matrices
    A = matrices.spd(rng(), 4, 8, 10.0)
    keep = A.copy()
    tsdx.cholesky(A)
    np.testing.assert_array_equal(A, keep)
    big = np.zeros((4, 8, 16))
    big[:, :, ::2] = A
    L, _ = tsdx.cholesky(big[:, :, ::2])
    np.testing.assert_allclose(L, np.linalg.cholesky(A), atol=1e-12)


def test_not_positive_definite_reports_info(tsdx):
    from tinysolverdx import matrices
    A = matrices.spd(rng(), 3, 5, 10.0)
    A[1] = -np.eye(5)
    _, info = tsdx.cholesky(A)
    assert list(info) == [0, 1, 0]


def test_singular_lu_reports_info(tsdx):
    from tinysolverdx import matrices
    A = matrices.general(rng(), 2, 4, 10.0)
    _, _, info = tsdx.lu(A)
    assert list(info) == [0, 0]
    A[1] = 0
    _, _, info = tsdx.lu(A)
    assert info[1] == 1


def test_fp32_precision_on_fp64_storage(tsdx):
    """precision='fp32' on float64 arrays factors a float32 copy (runtime.cpp:95-106)."""
    from tinysolverdx import matrices
    A = matrices.spd(rng(4), 10, 16, 10.0)
    L, info = tsdx.cholesky(A, precision="fp32")
    L32, _ = tsdx.cholesky(A.astype(np.float32))
    assert L.dtype == np.float64 and (info == 0).all()
    np.testing.assert_array_equal(L, L32.astype(np.float64))


def test_error_metrics_against_numpy(tsdx):
    from tinysolverdx import matrices
    r = rng(5)
    A = matrices.general(r, 6, 10, 1e4)
    b = r.standard_normal((6, 10))
    LU, piv, _ = tsdx.lu(A)
    assert tsdx.lu_backward_error(A, LU, piv).max() < 1e-14
    x = np.linalg.solve(A, b[..., None])[..., 0]
    assert tsdx.solve_backward_error(A, x, b).max() < 1e-14
    assert tsdx.solve_forward_error(A, x, b).max() < 1e-10
    np.testing.assert_allclose(tsdx.reference_solve(A, b), x, rtol=1e-9)
    assert tsdx.solve_forward_error(A, x * 1.01, b).min() > 5e-3


def test_fp64_solve_backward_error_matches_reference_bar(tsdx):
    """GesvAndPosvMatchReference (tests/cpp/test_solvers.cpp:173-196): fp64 solve
    backward error < 1e-14 and agreement with the long-double solve."""
    r = rng(21)
    for n in (2, 9, 32, 64):
        Ag = r.standard_normal((13, n, n))
        M = r.standard_normal((13, n, n))
        As = M @ M.transpose(0, 2, 1) + n * np.eye(n)
        b = r.standard_normal((13, n))
        x, info = tsdx.gesv(Ag, b)
        xref = tsdx.reference_solve(Ag, b)
        np.testing.assert_allclose(x, xref, rtol=0, atol=1e-8 * (1 + np.abs(xref).max()))
        assert tsdx.solve_backward_error(Ag, x, b).max() < 1e-14
        x, info = tsdx.posv(As, b)
        xref = tsdx.reference_solve(As, b)
        np.testing.assert_allclose(x, xref, rtol=0, atol=1e-9 * (1 + np.abs(xref).max()))
