"""Pins the oracle (oracle/tsdx_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a separate multiply and add instead of one fused operation, a wrong tie
rule, a transposed operand) fails one of them:
  * golden fixtures written by the compiled reference (bit-exact);
  * the compiled reference itself on fresh random inputs, when built here;
  * closed forms (2x2 Cholesky, the row-swap example, zero columns, -I);
  * library routines (numpy Cholesky, scipy lu_factor pivots, numpy solve);
  * invariants (piv[k] in [k, n), backward error bounds 4nu / 8nu, upper
    triangle zero, thread-count independence).
"""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import bits_equal, first_bit_mismatch, golden, golden_cases


# --------------------------------------------------------------- golden fixtures
@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_potrf_matches_golden(oracle, sfx):
    g = golden(f"potrf_{sfx}")
    for c in golden_cases(g, "_A"):
        A = g[f"{c}_A"]
        L, info = oracle.cholesky(A, "auto")
        assert np.array_equal(info, g[f"{c}_info"]), c
        assert bits_equal(L, g[f"{c}_L"]), (c, first_bit_mismatch(L, g[f"{c}_L"]))
        Ls, infos = oracle.cholesky(A, "scalar")
        assert np.array_equal(infos, g[f"{c}_infoscalar"]), c
        assert bits_equal(Ls, g[f"{c}_Lscalar"]), c
        X = oracle.cholesky_solve(g[f"{c}_L"], g[f"{c}_B"])
        assert bits_equal(X, g[f"{c}_X"]), c


@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_getrf_matches_golden(oracle, sfx):
    g = golden(f"getrf_{sfx}")
    for c in golden_cases(g, "_A"):
        A = g[f"{c}_A"]
        LU, piv, info = oracle.lu(A, "auto")
        assert np.array_equal(piv, g[f"{c}_piv"]), c
        assert np.array_equal(info, g[f"{c}_info"]), c
        assert bits_equal(LU, g[f"{c}_LU"]), (c, first_bit_mismatch(LU, g[f"{c}_LU"]))
        if f"{c}_LUscalar" in g.files:
            LUs, pivs, infos = oracle.lu(A, "scalar")
            assert np.array_equal(pivs, g[f"{c}_pivscalar"]) and np.array_equal(infos, g[f"{c}_infoscalar"]), c
            assert bits_equal(LUs, g[f"{c}_LUscalar"]), c
        X = oracle.lu_solve(g[f"{c}_LU"], g[f"{c}_piv"], g[f"{c}_B"])
        assert bits_equal(X, g[f"{c}_X"]), c


def test_metrics_match_golden(oracle):
    g = golden("metrics_f64")
    for c in golden_cases(g, "_A"):
        A, S, b = g[f"{c}_A"], g[f"{c}_S"], g[f"{c}_b"]
        np.testing.assert_array_equal(oracle.lu_backward_error(A, g[f"{c}_LU"], g[f"{c}_piv"]), g[f"{c}_lu_bwd"])
        np.testing.assert_array_equal(oracle.chol_backward_error(S, g[f"{c}_L"]), g[f"{c}_chol_bwd"])
        np.testing.assert_array_equal(oracle.solve_backward_error(A, g[f"{c}_xp"], b), g[f"{c}_solve_bwd"])
        np.testing.assert_array_equal(oracle.solve_forward_error(A, g[f"{c}_xp"], b), g[f"{c}_solve_fwd"])
        np.testing.assert_array_equal(oracle.reference_solve(A, b), g[f"{c}_x"])


def test_precision_matches_golden_and_kats(oracle):
    g = golden("precision")
    got16 = np.array([oracle.round_fp16(v) for v in g["x"]], np.float32)
    got8 = np.array([oracle.round_bf16(v) for v in g["x"]], np.float32)
    assert bits_equal(got16, g["fp16"]) and bits_equal(got8, g["bf16"])
    for p in ("fp32", "fp16", "bf16", "fp64"):
        assert bits_equal(oracle.round_to(g["x64"], p), g[f"round_to_{p}"]), p
    # known answers of tests/cpp/test_precision.cpp:10-42
    r16, rb = oracle.round_fp16, oracle.round_bf16
    assert r16(1.0 + 1 / 2048) == 1.0 and r16(1.0 + 3 / 2048) == 1.0 + 4 / 2048
    assert r16(65504.0) == 65504.0 and np.isinf(r16(65520.0))
    assert r16(6.0e-8) == np.float32(5.9604644775390625e-08) and r16(1e-9) == 0.0
    assert r16(-0.1) == np.float32(-0.0999755859375) and np.isnan(r16(np.nan))
    assert rb(1.0 + 1 / 256) == 1.0 and rb(1.0 + 3 / 256) == 1.0 + 4 / 256
    assert rb(3.14159265) == np.float32(3.140625) and np.isinf(rb(np.finfo(np.float32).max))
    assert [oracle.unit_roundoff(p) for p in ("fp32", "fp64", "fp16", "bf16")] == [2.0 ** -24, 2.0 ** -53, 2.0 ** -11,
                                                                                    2.0 ** -8]


# ------------------------------------------------ compiled reference, fresh inputs
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_oracle_equals_compiled_reference(oracle, ref_lib, dt):
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 6, 8, 13, 16, 17, 24, 25, 31, 32, 40, 64):
        b = 19
        G = rng.standard_normal((b, n, n)).astype(dt)
        M = rng.standard_normal((b, n, n))
        S = (M @ M.transpose(0, 2, 1) + n * np.eye(n)).astype(dt)
        for be in ("naive", "scalar", "simd", "auto"):
            L1, i1 = oracle.cholesky(S, be)
            L2, i2 = ref_lib.cholesky(S, be)
            assert bits_equal(L1, L2) and np.array_equal(i1, i2), (n, be)
            F1, p1, j1 = oracle.lu(G, be)
            F2, p2, j2 = ref_lib.lu(G, be)
            assert bits_equal(F1, F2) and np.array_equal(p1, p2) and np.array_equal(j1, j2), (n, be)
        B = rng.standard_normal((b, n, 2)).astype(dt)
        assert bits_equal(oracle.posv(S, B)[1], ref_lib.posv(S, B)[1]), n
        assert bits_equal(oracle.gesv(G, B)[1], ref_lib.gesv(G, B)[1]), n


# -------------------------------------------------- emulated fp16 / bf16 (Emu<F>)
def _emu_keys():
    g = golden("emulated")
    return g, golden_cases(g, "_S")


def test_emulated_matches_golden(oracle):
    """Runtime-precision fp16 / bf16 calls of the port against the compiled
    reference's outputs: factors (failure content included), pivots, info and
    both solves, bit for bit."""
    g, keys = _emu_keys()
    assert len(keys) == 26
    for key in keys:
        prec = key.split("_")[0]
        L, info = oracle.cholesky_p(g[f"{key}_S"], prec)
        assert bits_equal(L, g[f"{key}_L"]) and np.array_equal(info, g[f"{key}_info"]), \
            (key, first_bit_mismatch(L, g[f"{key}_L"]))
        LU, piv, linfo = oracle.lu_p(g[f"{key}_G"], prec)
        assert bits_equal(LU, g[f"{key}_LU"]) and np.array_equal(piv, g[f"{key}_piv"]), key
        assert np.array_equal(linfo, g[f"{key}_linfo"]), key
        assert bits_equal(oracle.cholesky_solve_p(g[f"{key}_L"], g[f"{key}_B"], prec), g[f"{key}_X"]), key
        assert bits_equal(oracle.lu_solve_p(g[f"{key}_LU"], g[f"{key}_piv"], g[f"{key}_B"], prec), g[f"{key}_Y"]), key


def test_emulated_fixture_cases_cover_failures():
    g, keys = _emu_keys()
    assert any((g[f"{k}_info"] > 0).any() for k in keys)          # failed Cholesky
    assert any((g[f"{k}_linfo"] > 0).any() for k in keys)         # zero LU pivot
    assert any(np.isnan(g[f"{k}_LU"]).any() for k in keys)        # NaN propagation


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_emulated_equals_compiled_reference(oracle, ref_lib, dt):
    rng = np.random.default_rng(12)
    for prec in ("fp16", "bf16"):
        for n in (1, 2, 4, 7, 16, 21, 32, 50, 64):
            G = rng.standard_normal((7, n, n)).astype(dt)
            M = rng.standard_normal((7, n, n))
            S = (M @ M.transpose(0, 2, 1) / n + np.eye(n)).astype(dt)
            S[3] = -S[3]
            L1, i1 = oracle.cholesky_p(S, prec)
            L2, i2 = ref_lib.cholesky_p(S, prec)
            assert bits_equal(L1, L2) and np.array_equal(i1, i2), (prec, n)
            F1, p1, j1 = oracle.lu_p(G, prec)
            F2, p2, j2 = ref_lib.lu_p(G, prec)
            assert bits_equal(F1, F2) and np.array_equal(p1, p2) and np.array_equal(j1, j2), (prec, n)
            B = rng.standard_normal((7, n, 2)).astype(dt)
            assert bits_equal(oracle.cholesky_solve_p(L2, B, prec), ref_lib.cholesky_solve_p(L2, B, prec))
            assert bits_equal(oracle.lu_solve_p(F2, p2, B, prec), ref_lib.lu_solve_p(F2, p2, B, prec))


def test_emulated_closed_forms(oracle):
    """Values that every step keeps exact in the format give the exact factors;
    a failure leaves the rows after it as the rounded input; the upper
    triangle is zeroed only on success."""
    A = np.array([[[4.0, 2.0, -2.0], [2.0, 10.0, 5.0], [-2.0, 5.0, 11.0]]], np.float32)
    for prec in ("fp16", "bf16"):
        L, info = oracle.cholesky_p(A, prec)
        assert info[0] == 0
        assert L[0, 0, 0] == 2 and L[0, 1, 0] == 1 and L[0, 1, 1] == 3 and L[0, 2, 0] == -1 and L[0, 2, 1] == 2
        assert L[0, 2, 2] == oracle.round_to(np.sqrt(6.0), prec)[()] and (np.triu(L[0], 1) == 0).all()
        M = np.array([[[0.0, 2.0], [4.0, 8.0]]], np.float32)
        LU, piv, info = oracle.lu_p(M, prec)
        assert list(piv[0]) == [1, 1] and info[0] == 0
        np.testing.assert_array_equal(LU[0], [[4.0, 8.0], [0.0, 2.0]])
        F = np.array([[[1.0, 0.1], [0.1, -1.0]]], np.float32)   # second pivot fails
        L, info = oracle.cholesky_p(F, prec)
        r = oracle.round_to(F[0].astype(np.float64), prec)
        assert info[0] == 2 and L[0, 0, 0] == 1 and L[0, 1, 0] == r[1, 0]
        assert L[0, 0, 1] == r[0, 1] and L[0, 1, 1] == r[1, 1]    # untouched entries come back rounded


# ------------------------------------------------------------------ closed forms
def test_closed_forms(oracle):
    L, info = oracle.cholesky(np.array([[[4.0, 2.0], [2.0, 3.0]]]))
    assert info[0] == 0
    np.testing.assert_array_equal(L[0], [[2.0, 0.0], [1.0, np.sqrt(2.0)]])
    for be in ("naive", "scalar", "simd", "auto"):
        LU, piv, info = oracle.lu(np.array([[[0.0, 1.0], [1.0, 0.0]]]), be)
        assert list(piv[0]) == [1, 1] and info[0] == 0 and LU[0, 0, 0] == 1 and LU[0, 1, 1] == 1
        _, info = oracle.cholesky(-np.eye(5)[None], be)
        assert info[0] == 1
    rng = np.random.default_rng(3)
    for n in (4, 9, 30, 64):
        for j in (0, n // 2, n - 1):
            A = rng.standard_normal((1, n, n))
            A[0, :, j] = 0.0
            _, piv, info = oracle.lu(A)
            assert info[0] == j + 1 and piv[0, j] == j, (n, j)


# ------------------------------------------------------------- library routines
@pytest.mark.parametrize("n", [2, 5, 16, 33, 64])
def test_against_lapack(oracle, n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((7, n, n))
    S = M @ M.transpose(0, 2, 1) + n * np.eye(n)
    G = rng.standard_normal((7, n, n))
    L, info = oracle.cholesky(S)
    assert (info == 0).all()
    np.testing.assert_allclose(L, np.linalg.cholesky(S), rtol=1e-12, atol=1e-12)
    assert np.all(np.triu(L, 1) == 0)
    LU, piv, info = oracle.lu(G)
    for b in range(7):
        ref_lu, ref_piv = sla.lu_factor(G[b])
        np.testing.assert_array_equal(piv[b], ref_piv)
        np.testing.assert_allclose(LU[b], ref_lu, rtol=1e-10, atol=1e-10)
    B = rng.standard_normal((7, n, 3))
    np.testing.assert_allclose(oracle.lu_solve(LU, piv, B), np.linalg.solve(G, B), rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(oracle.cholesky_solve(L, B), np.linalg.solve(S, B), rtol=1e-8, atol=1e-8)


# -------------------------------------------------------------------- invariants
@pytest.mark.parametrize("dt,u", [(np.float32, 2.0 ** -24), (np.float64, 2.0 ** -53)])
def test_backward_error_bounds(oracle, dt, u):
    rng = np.random.default_rng(5)
    for n in (1, 3, 8, 17, 32, 64):
        M = rng.standard_normal((11, n, n))
        S = (M @ M.transpose(0, 2, 1) + n * np.eye(n)).astype(dt)
        G = rng.standard_normal((11, n, n)).astype(dt)
        L, _ = oracle.cholesky(S)
        assert oracle.chol_backward_error(S.astype(np.float64), L).max() < 4 * u * n
        LU, piv, info = oracle.lu(G)
        assert (info == 0).all()
        for k in range(n):
            assert ((piv[:, k] >= k) & (piv[:, k] < n)).all()
        assert oracle.lu_backward_error(G.astype(np.float64), LU, piv).max() < 8 * u * n


def test_result_independent_of_thread_count(oracle):
    rng = np.random.default_rng(4)
    M = rng.standard_normal((3000, 12, 12))
    S = M @ M.transpose(0, 2, 1) + 12 * np.eye(12)
    oracle.set_num_threads(1)
    a1 = oracle.cholesky(S)[0]
    oracle.set_num_threads(4)
    a4 = oracle.cholesky(S)[0]
    oracle.set_num_threads(0)
    assert bits_equal(a1, a4)


def test_error_codes(oracle):
    with pytest.raises(ValueError):
        oracle.cholesky(np.zeros((1, 65, 65)))
    with pytest.raises(RuntimeError):
        oracle.cholesky(np.eye(3)[None], "cuda")


def test_router_oracle_gradient_matches_finite_differences():
    """The CPU MLP that pins the router kernels: its analytic gradient against
    central differences of its loss, at entries of every layer."""
    import mlp_oracle as M
    rng = np.random.default_rng(0)
    p = rng.standard_normal(M.PARAMS) * 0.3
    p[12:24] = rng.uniform(0.5, 2.0, 12)
    feat = rng.standard_normal((40, 12))
    label = rng.integers(0, 3, 40)
    g = M.grad(p, feat, label)
    assert np.all(g[:24] == 0)
    idx = np.concatenate([rng.integers(lo, hi, 6) for lo, hi in
                          ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                           (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.PARAMS))])
    fd = M.finite_difference(p, feat, label, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8), np.abs(g[idx] - fd).max()
    # a wrong sign or a dropped ReLU mask would fail here: the loss decreases along -g
    assert M.loss(p - 1e-3 * g, feat, label) < M.loss(p, feat, label)


def test_router_oracle_features_closed_forms():
    """Features of diagonal and 2x2 matrices, worked by hand."""
    import mlp_oracle as M
    f = M.features(np.diag([4.0, 1.0, 0.25, 1.0])[None])[0]
    assert f[0] == 4 / 64 and np.isclose(f[1], np.log10(0.25)) and np.isclose(f[2], np.log10(4.0))
    assert np.isclose(f[3], np.log10(np.sqrt(16 + 1 + 1 / 16 + 1))) and np.isclose(f[7], np.log10(16.0))
    assert f[6] == -30.0 and f[10] == 0.0 and f[11] == 1.0
    # only the lower triangle counts: the upper 99 is ignored
    a = np.array([[2.0, 99.0], [-1.0, 3.0]])
    f = M.features(a[None])[0]
    assert np.isclose(f[4], np.log10(4.0)) and np.isclose(f[5], (2.0 - 1.0) / np.sqrt(4 + 9 + 2)) and f[11] == 1.0
    assert M.features(np.array([[[-1.0, 0.0], [0.0, 1.0]]]))[0][10] == 1.0


def test_router_oracle_bf16_inference_pinned():
    """infer_bf16 (the matrix-core inference path) agrees with the float64
    forward pass to the bfloat16 rounding of its hidden activations, with the
    identity normalisation and bfloat16 weights and features as well as with a
    folded normalisation (a wrong fold shows as errors of order one); with
    zero second and third layers it returns b3 exactly; bfloat16 rounding is
    pinned by known values (ties to even)."""
    import mlp_oracle as M
    rb = M.round_bf16(np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 1.0 + 2.0 ** -7 + 2.0 ** -8,
                                3.0e38], np.float32))
    assert rb.tolist() == [1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 1.0 + 2.0 ** -6, float(M.round_bf16(np.float32(3.0e38)))]
    assert np.isinf(M.round_bf16(np.array([3.4e38], np.float32))[0])  # past the largest bfloat16: infinity
    rng = np.random.default_rng(9)
    p = np.zeros(M.PARAMS, np.float32)
    p[12:24] = 1.0
    p[M.OFF_W1:] = M.round_bf16((rng.standard_normal(M.PARAMS - M.OFF_W1) * 0.3).astype(np.float32))
    feat = M.round_bf16(rng.standard_normal((50, 12)).astype(np.float32))
    # hidden activations rounded to bfloat16 make the two differ by at most
    # 2^-9 relative per activation; with exact bf16 activations they agree:
    z_ref = M.forward(p, feat)[3]
    z_bf = M.infer_bf16(p, feat)
    assert np.allclose(z_bf, z_ref, rtol=3e-2, atol=3e-2)
    # one hidden layer of zero width (W2 = W3 = 0): logits are exactly b3
    q = p.copy()
    q[M.OFF_W2:M.OFF_B2] = 0.0
    q[M.OFF_W3:M.OFF_B3] = 0.0
    assert np.array_equal(M.infer_bf16(q, feat), np.tile(q[M.OFF_B3:].astype(np.float64), (50, 1)))
    assert M.tiers(np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 0.0], [1.0, np.nan, 0.0]])).tolist() == [1, 0, 2]
    r = p.copy()
    r[0:12] = rng.standard_normal(12)
    r[12:24] = rng.uniform(0.5, 2.0, 12)
    raw = (feat * r[12:24] + r[0:12]).astype(np.float32)  # features whose normalised values are `feat`
    assert np.allclose(M.infer_bf16(r, raw), M.forward(r, raw)[3], rtol=3e-2, atol=3e-2)
