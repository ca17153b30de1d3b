"""The NumPy API of the package (tinysolverdx.cholesky / lu / *_solve / posv /
gesv / round_to / metrics) on the GPU, checked against the interface the
reference documents (python/tinysolverdx/__init__.py:1-117,
src/python/bindings.cpp:15-177): shapes, dtypes, squeezing of a single matrix
and of vector right-hand sides, inputs left untouched, error types.  Numbers
are pinned to the oracle (bit for bit where the reference is bit-exact) and to
LAPACK (numpy / scipy)."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import bits_equal, first_bit_mismatch

pytestmark = pytest.mark.gpu

U = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}


def _mats(kind, seed, batch, n, kappa, dtype=np.float64):
    from tinysolverdx import matrices
    rng = np.random.default_rng(seed)
    make = matrices.spd if kind == "spd" else matrices.general
    return make(rng, batch, n, kappa).astype(dtype)


# ------------------------------------------------------------ factorizations
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 31, 32, 33, 63, 64])
def test_cholesky_api_equals_oracle_and_lapack(tsdx, oracle, dtype, n):
    A = _mats("spd", 100 + n, 19, n, 1e2, dtype)
    for backend in ("auto", "hip"):
        L, info = tsdx.cholesky(A, backend=backend)
        assert L.dtype == dtype and L.shape == A.shape and info.shape == (19,) and info.dtype == np.int32
        L0, info0 = oracle.cholesky(A)
        assert np.array_equal(info, info0) and not info.any()
        assert bits_equal(L, L0), first_bit_mismatch(L, L0)
    # the factor is lower, and close to LAPACK's within the forward bound
    assert not np.triu(L, 1).any()
    ref = np.linalg.cholesky(A.astype(np.float64))
    assert np.abs(L - ref).max() <= 50 * n * U[dtype] * 1e2 * np.abs(ref).max()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 4, 7, 16, 24, 25, 32, 48, 64])
def test_lu_api_pivots_are_lapack_pivots(tsdx, oracle, dtype, n):
    A = _mats("gen", 200 + n, 11, n, 1e3, dtype)
    LU, piv, info = tsdx.lu(A)
    assert piv.shape == (11, n) and piv.dtype == np.int32 and LU.dtype == dtype
    LU0, piv0, info0 = oracle.lu(A)
    assert np.array_equal(piv, piv0) and np.array_equal(info, info0)
    assert bits_equal(LU, LU0), first_bit_mismatch(LU, LU0)
    # sequential 0-based swaps: scipy's getrf convention
    for m in range(A.shape[0]):
        lu_ref, piv_ref = sla.lu_factor(A[m].astype(np.float64))
        if dtype == np.float64:
            assert np.array_equal(piv[m], piv_ref)
        P = np.eye(n)
        for k, p in enumerate(piv[m]):
            P[[k, p]] = P[[p, k]]
        Lf = np.tril(LU[m].astype(np.float64), -1) + np.eye(n)
        Uf = np.triu(LU[m].astype(np.float64))
        assert np.abs(P @ A[m] - Lf @ Uf).max() <= 16 * n * U[dtype] * np.abs(A[m]).max() * 10


def test_failure_codes(tsdx):
    # Cholesky: a matrix that is not positive definite stops at its first bad pivot
    S = _mats("spd", 1, 4, 6, 10.0)
    S[2] = -np.eye(6)
    S[3][4, 4] = -5.0
    _, info = tsdx.cholesky(S)
    assert info[0] == 0 and info[1] == 0 and info[2] == 1 and info[3] == 5
    # LU: zero column j -> info j+1 with piv[j] = j; a zero matrix -> 1
    G = _mats("gen", 2, 3, 5, 10.0)
    G[1][:, 3] = 0.0
    G[2][:] = 0.0
    _, piv, info = tsdx.lu(G)
    assert list(info) == [0, 4, 1] and piv[1][3] == 3


# ------------------------------------------------------------------ solves
@pytest.mark.parametrize("n", [2, 7, 32, 64])
def test_solvers_agree_with_numpy_solve(tsdx, n):
    A = _mats("gen", 300 + n, 9, n, 1e3)
    S = _mats("spd", 400 + n, 9, n, 1e3)
    rng = np.random.default_rng(n)
    b = rng.standard_normal((9, n))
    B = rng.standard_normal((9, n, 3))
    want = lambda M, R: np.linalg.solve(M, R)
    x, info = tsdx.gesv(A, b)
    assert x.shape == (9, n) and not info.any()
    assert np.allclose(x, want(A, b[..., None])[..., 0], rtol=1e-9, atol=1e-9)
    x, info = tsdx.posv(S, b)
    assert x.shape == (9, n) and not info.any()
    assert np.allclose(x, want(S, b[..., None])[..., 0], rtol=1e-9, atol=1e-9)
    X = tsdx.gesv(A, B)[0]
    assert X.shape == (9, n, 3) and np.allclose(X, want(A, B), rtol=1e-9, atol=1e-9)
    # the two-call forms give the same solutions as the fused drivers
    LU, piv, _ = tsdx.lu(A)
    assert np.allclose(tsdx.lu_solve(LU, piv, B), X, rtol=1e-10, atol=1e-12)
    L, _ = tsdx.cholesky(S)
    assert np.allclose(tsdx.cholesky_solve(L, B), want(S, B), rtol=1e-9, atol=1e-9)
    assert tsdx.solve_backward_error(S, tsdx.posv(S, b)[0], b).max() < 1e-14


def test_one_matrix_and_vector_shapes(tsdx):
    S = _mats("spd", 3, 1, 6, 10.0)[0]
    G = _mats("gen", 4, 1, 6, 10.0)[0]
    b = np.linspace(-1.0, 1.0, 6)
    B = np.stack([b, 2 * b], axis=1)
    x, info = tsdx.posv(S, b)
    assert x.shape == (6,) and np.ndim(info) == 0 and info == 0
    assert np.allclose(S @ x, b, atol=1e-12)
    X, _ = tsdx.gesv(G, B)
    assert X.shape == (6, 2) and np.allclose(G @ X, B, atol=1e-12)
    L, info = tsdx.cholesky(S)
    assert L.shape == (6, 6) and np.ndim(info) == 0
    LU, piv, info = tsdx.lu(G)
    assert LU.shape == (6, 6) and piv.shape == (6,)
    assert tsdx.lu_solve(LU, piv, b).shape == (6,)
    assert tsdx.cholesky_solve(L, B).shape == (6, 2)


def test_inputs_untouched_and_strided_views(tsdx):
    S = _mats("spd", 5, 4, 8, 10.0)
    b = np.ones((4, 8))
    S0, b0 = S.copy(), b.copy()
    tsdx.cholesky(S)
    tsdx.posv(S, b)
    tsdx.gesv(S, b)
    assert np.array_equal(S, S0) and np.array_equal(b, b0)
    wide = np.zeros((4, 8, 24))
    wide[:, :, 1::3] = S
    L, _ = tsdx.cholesky(wide[:, :, 1::3])
    assert np.array_equal(L, tsdx.cholesky(S)[0])


def test_precision_on_wider_storage(tsdx):
    """precision='fp32' on float64 arrays factors a float32 copy and returns
    float64 values that are float32 numbers (runtime.cpp:95-106)."""
    S = _mats("spd", 6, 10, 16, 10.0)
    L, info = tsdx.cholesky(S, precision="fp32")
    L32, _ = tsdx.cholesky(S.astype(np.float32))
    assert L.dtype == np.float64 and not info.any()
    assert np.array_equal(L, L32.astype(np.float64))


def test_argument_errors(tsdx):
    bad = [
        (ValueError, lambda: tsdx.cholesky(np.zeros((2, 65, 65)))),       # n outside [1, 64]
        (ValueError, lambda: tsdx.lu(np.zeros((2, 3, 4)))),               # not square
        (ValueError, lambda: tsdx.cholesky(np.eye(4)[None], precision="fp8")),
        (ValueError, lambda: tsdx.cholesky(np.eye(4)[None], backend="metal")),
        (TypeError, lambda: tsdx.lu(np.eye(4, dtype=np.int64))),          # no implied precision
        (ValueError, lambda: tsdx.posv(np.eye(4)[None], np.ones((1, 5)))),  # rhs rows != n
        (ValueError, lambda: tsdx.lu_solve(np.eye(4)[None], np.zeros((1, 3), np.int32), np.ones((1, 4)))),
        (RuntimeError, lambda: tsdx.cholesky(np.eye(4)[None], backend="cuda")),   # not in this build
        (RuntimeError, lambda: tsdx.cholesky(np.eye(4)[None], backend="simd")),   # CPU backends are not
    ]
    for exc, call in bad:
        with pytest.raises(exc):
            call()
    assert not tsdx.has_cuda() and tsdx.has_hip()


# ------------------------------------------------------------------ metrics
def test_error_metrics_on_known_solutions(tsdx):
    A = _mats("gen", 7, 6, 10, 1e4)
    rng = np.random.default_rng(7)
    x_true = rng.standard_normal((6, 10))
    b = np.einsum("bij,bj->bi", A, x_true)
    x = tsdx.reference_solve(A, b)
    assert np.abs(x - x_true).max() < 1e-9
    assert tsdx.solve_backward_error(A, x, b).max() < 1e-15
    assert tsdx.solve_forward_error(A, x, b).max() < 1e-15
    # a 1e-3 relative perturbation is a 1e-3 forward error
    assert np.allclose(tsdx.solve_forward_error(A, x * (1 + 1e-3), b), 1e-3, rtol=1e-6)
    LU, piv, _ = tsdx.lu(A)
    assert tsdx.lu_backward_error(A, LU, piv).max() < 1e-15 * 10
    S = _mats("spd", 8, 3, 9, 1e2)
    L, _ = tsdx.cholesky(S)
    assert tsdx.chol_backward_error(S, L).max() < 4 * 9 * U[np.float64]


def test_machine_probes(tsdx):
    """The GPU probes give plausible MI355X ceilings (published: 8 TB/s HBM,
    78.6 TFLOP/s fp64 vector and matrix, 2.5 PFLOP/s bf16 matrix)."""
    bw = tsdx.measure_bandwidth_gbs()
    assert 1500.0 < bw < 8000.0, bw
    f64 = tsdx.measure_peak_gflops("fp64")
    f32 = tsdx.measure_peak_gflops("fp32")
    assert 10_000 < f64 < 90_000 and 10_000 < f32 < 180_000, (f32, f64)
    bf16 = tsdx.measure_mfma_tflops("bf16")
    assert 200 < bf16 < 2600, bf16
