"""Emulated fp16 / bf16 on the GPU (the reference's Emu<F> arithmetic under its
Auto backend, precision.hpp:71-97 and kernels.hpp:18-227) against the
compiled reference's fixtures and the oracle: factors (the content a failed
Cholesky leaves included), pivots, info and both solves are bit-identical,
through the host calls, the stream-ordered device calls and the Python API.
The reference's own tests of the emulated path
(tests/python/test_solvers.py:91-105) are repeated at the end."""
import numpy as np
import pytest

import hipcalls as H
from conftest import bits_equal, first_bit_mismatch, golden, golden_cases

pytestmark = pytest.mark.gpu


def _keys():
    g = golden("emulated")
    return g, golden_cases(g, "_S")


def test_golden_host_calls(hip):
    g, keys = _keys()
    for key in keys:
        prec = key.split("_")[0]
        L, info = H.cholesky_emu(g[f"{key}_S"], prec)
        assert bits_equal(L, g[f"{key}_L"]) and np.array_equal(info, g[f"{key}_info"]), \
            (key, first_bit_mismatch(L, g[f"{key}_L"]), info, g[f"{key}_info"])
        LU, piv, linfo = H.lu_emu(g[f"{key}_G"], prec)
        assert bits_equal(LU, g[f"{key}_LU"]), (key, first_bit_mismatch(LU, g[f"{key}_LU"]))
        assert np.array_equal(piv, g[f"{key}_piv"]) and np.array_equal(linfo, g[f"{key}_linfo"]), key
        X = H.cholesky_solve_emu(g[f"{key}_L"], g[f"{key}_B"], prec)
        assert bits_equal(X, g[f"{key}_X"]), (key, first_bit_mismatch(X, g[f"{key}_X"]))
        Y = H.lu_solve_emu(g[f"{key}_LU"], g[f"{key}_piv"], g[f"{key}_B"], prec)
        assert bits_equal(Y, g[f"{key}_Y"]), (key, first_bit_mismatch(Y, g[f"{key}_Y"]))


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_every_size_vs_oracle(hip, oracle, prec, dt):
    """n = 1..64 (every bucket and padding), several matrices per wave for
    n <= 16, with failures, NaN and zero pivots mixed in."""
    rng = np.random.default_rng(7)
    for n in range(1, 65):
        b = 9
        M = rng.standard_normal((b, n, n))
        S = (M @ M.transpose(0, 2, 1) / n + np.eye(n)).astype(dt)
        S[4] = -S[4]
        if n > 2:
            S[6, n // 2, n // 2] = -5.0
        G = rng.standard_normal((b, n, n)).astype(dt)
        G[2, :, n // 3] = 0.0
        G[5, n - 1, 0] = np.nan
        L, info = H.cholesky_emu(S, prec)
        Lo, info_o = oracle.cholesky_p(S, prec)
        assert bits_equal(L, Lo) and np.array_equal(info, info_o), (n, first_bit_mismatch(L, Lo))
        LU, piv, linfo = H.lu_emu(G, prec)
        LUo, pivo, linfo_o = oracle.lu_p(G, prec)
        assert bits_equal(LU, LUo) and np.array_equal(piv, pivo) and np.array_equal(linfo, linfo_o), n
        B = rng.standard_normal((b, n, 3)).astype(dt)
        assert bits_equal(H.cholesky_solve_emu(Lo, B, prec), oracle.cholesky_solve_p(Lo, B, prec)), n
        assert bits_equal(H.lu_solve_emu(LUo, pivo, B, prec), oracle.lu_solve_p(LUo, pivo, B, prec)), n


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_drivers_are_factor_then_solve(hip, oracle, prec):
    """posv / gesv (the reference's posv_batched / gesv_batched: factor, then
    solve with that factor) on many matrices at once."""
    rng = np.random.default_rng(8)
    for n, b in ((3, 5000), (16, 3000), (40, 700)):
        M = rng.standard_normal((b, n, n))
        S = (M @ M.transpose(0, 2, 1) / n + np.eye(n)).astype(np.float32)
        G = rng.standard_normal((b, n, n)).astype(np.float32)
        B = rng.standard_normal((b, n, 2)).astype(np.float32)
        L, X, info = H.posv_emu(S, B, prec)
        Lo, info_o = oracle.cholesky_p(S, prec)
        assert bits_equal(L, Lo) and np.array_equal(info, info_o)
        assert bits_equal(X, oracle.cholesky_solve_p(Lo, B, prec)), n
        LU, Y, piv, linfo = H.gesv_emu(G, B, prec)
        LUo, pivo, linfo_o = oracle.lu_p(G, prec)
        assert bits_equal(LU, LUo) and np.array_equal(piv, pivo) and np.array_equal(linfo, linfo_o)
        assert bits_equal(Y, oracle.lu_solve_p(LUo, pivo, B, prec)), n


def test_device_calls_match_host_calls(hip):
    import torch
    from tinysolverdx import _lib
    rng = np.random.default_rng(9)
    n, b = 24, 300
    M = rng.standard_normal((b, n, n))
    S = (M @ M.transpose(0, 2, 1) / n + np.eye(n)).astype(np.float64)
    G = rng.standard_normal((b, n, n))
    B = rng.standard_normal((b, n, 4))
    for prec in ("fp16", "bf16"):
        f = _lib.EMU[prec]
        dS, dG, dB1, dB2 = (torch.from_numpy(x.copy()).cuda() for x in (S, G, B, B))
        info = torch.zeros(b, dtype=torch.int32, device="cuda")
        piv = torch.zeros((b, n), dtype=torch.int32, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        _lib.call("tsdx_hip_potrf_emu_f64", dS.data_ptr(), n, b, info.data_ptr(), f, s)
        _lib.call("tsdx_hip_potrs_emu_f64", dS.data_ptr(), n, b, dB1.data_ptr(), 4, f, s)
        _lib.call("tsdx_hip_getrf_emu_f64", dG.data_ptr(), n, b, piv.data_ptr(), info.data_ptr(), f, s)
        _lib.call("tsdx_hip_getrs_emu_f64", dG.data_ptr(), piv.data_ptr(), n, b, dB2.data_ptr(), 4, f, s)
        torch.cuda.synchronize()
        L, X, _ = H.posv_emu(S, B, prec)
        LU, Y, p, _ = H.gesv_emu(G, B, prec)
        assert bits_equal(dS.cpu().numpy(), L) and bits_equal(dB1.cpu().numpy(), X)
        assert bits_equal(dG.cpu().numpy(), LU) and bits_equal(dB2.cpu().numpy(), Y)
        assert np.array_equal(piv.cpu().numpy(), p)


def test_format_errors(hip):
    from tinysolverdx import _lib
    L = _lib.lib()
    A = np.eye(4, dtype=np.float32)[None].copy()
    info = np.zeros(1, np.int32)
    assert L.tsdx_hip_cholesky_emu_batched_f32(H._p(A), 4, 1, H._p(info), 0) == _lib.EINVAL
    assert "format" in _lib.last_error()
    assert L.tsdx_hip_cholesky_emu_batched_f32(H._p(A), 65, 1, H._p(info), 1) == _lib.EINVAL
    assert L.tsdx_hip_lu_emu_batched_f32(H._p(A), 4, 1, None, None, 1) == _lib.EINVAL


# ---------------------------------------------- the Python API (runtime precision)
@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_emulated_precisions_have_expected_accuracy(tsdx, prec):
    """The reference's test (tests/python/test_solvers.py:91-100)."""
    from tinysolverdx import matrices
    u = tsdx.unit_roundoff(prec)
    A = tsdx.round_to(matrices.spd(np.random.default_rng(3), 20, 16, 10.0), prec)
    L, info = tsdx.cholesky(A.astype(np.float32), precision=prec)
    assert (info == 0).all()
    assert (tsdx.round_to(L, prec) == L).all()  # outputs representable in the format
    err = tsdx.chol_backward_error(A, L.astype(np.float64))
    assert err.max() < 20 * u * 16 and err.max() > tsdx.unit_roundoff("fp32")


def test_float16_dtype_roundtrip(tsdx):
    """The reference's test (tests/python/test_solvers.py:103-106)."""
    from tinysolverdx import matrices
    A = matrices.spd(np.random.default_rng(0), 5, 8, 10.0).astype(np.float16)
    L, info = tsdx.cholesky(A)
    assert L.dtype == np.float16 and (info == 0).all()


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_python_api_matches_oracle(tsdx, oracle, prec):
    rng = np.random.default_rng(10)
    from tinysolverdx import matrices
    S = matrices.spd(rng, 30, 12, 1e2)
    G = matrices.general(rng, 30, 12, 1e2)
    b = rng.standard_normal((30, 12))
    for dt in (np.float32, np.float64):
        L, info = tsdx.cholesky(S.astype(dt), precision=prec)
        Lo, info_o = oracle.cholesky_p(S.astype(dt), prec)
        assert bits_equal(L, Lo) and np.array_equal(info, info_o)
        LU, piv, linfo = tsdx.lu(G.astype(dt), precision=prec)
        LUo, pivo, _ = oracle.lu_p(G.astype(dt), prec)
        assert bits_equal(LU, LUo) and np.array_equal(piv, pivo)
        x = tsdx.cholesky_solve(L, b.astype(dt), precision=prec)
        assert bits_equal(x, oracle.cholesky_solve_p(Lo, b.astype(dt)[..., None], prec)[..., 0])
        y = tsdx.lu_solve(LU, piv, b.astype(dt), precision=prec)
        assert bits_equal(y, oracle.lu_solve_p(LUo, pivo, b.astype(dt)[..., None], prec)[..., 0])
        xs, info = tsdx.posv(S.astype(dt), b.astype(dt), precision=prec)
        assert bits_equal(xs, x)
        ys, info = tsdx.gesv(G.astype(dt), b.astype(dt), precision=prec)
        assert bits_equal(ys, y)
    # float16 arrays: fp16 implied, float16 results
    L16, _ = tsdx.cholesky(S.astype(np.float16))
    assert L16.dtype == np.float16
    assert np.array_equal(L16.astype(np.float32), oracle.cholesky_p(S.astype(np.float16).astype(np.float32), "fp16")[0])
