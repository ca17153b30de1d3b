"""CPU-side checks of the product: the C ABI library loads and exports every
symbol include/tsdx_hip.h declares, argument errors come back with the
reference's error types before any device work, and the host-side parts of
the API (rounding, unit roundoff, long-double metrics) match the fixtures
written by the compiled reference.  No compute call needs a GPU here.
"""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

from conftest import ROOT, bits_equal, golden, golden_cases

HEADER = ROOT / "include" / "tsdx_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(tsdx_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from tinysolverdx import _lib
    L = C.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding table covers the same set
    assert set(_lib.SIGNATURES) == set(names)


def test_core_module_surface(tsdx):
    core = tsdx._core
    for name in ["cholesky_f32", "cholesky_f64", "lu_f32", "lu_f64", "cholesky_solve_f32", "cholesky_solve_f64",
                 "lu_solve_f32", "lu_solve_f64", "round_to_f32", "round_to_f64", "posv_f32", "posv_f64", "gesv_f32",
                 "gesv_f64", "has_cuda", "has_hip", "has_openmp", "num_threads", "set_num_threads",
                 "unit_roundoff", "chol_backward_error", "lu_backward_error", "solve_backward_error",
                 "solve_forward_error", "reference_solve", "measure_peak_gflops", "measure_bandwidth_gbs",
                 "measure_mfma_tflops"]:
        assert hasattr(core, name), name
    assert tsdx.has_cuda() is False


def test_abi_argument_errors_without_device():
    from tinysolverdx import _lib
    L = _lib.lib()
    a = np.zeros(64 * 64, np.float32)
    p = a.ctypes.data_as(C.c_void_p)
    assert L.tsdx_hip_potrf_f32(p, 65, 1, None, None) == _lib.EINVAL
    assert "outside supported range" in _lib.last_error()
    assert L.tsdx_hip_potrf_f32(p, 0, 1, None, None) == _lib.EINVAL
    assert L.tsdx_hip_potrf_f32(p, 4, -1, None, None) == _lib.EINVAL
    assert L.tsdx_hip_getrf_f32(p, 4, 1, None, None, None) == _lib.EINVAL
    assert "piv" in _lib.last_error()
    assert L.tsdx_hip_potrs_f32(p, 4, 1, p, -1, None) == _lib.EINVAL
    # batch 0 is a no-op that needs no device
    assert L.tsdx_hip_potrf_f64(None, 8, 0, None, None) == _lib.OK
    assert L.tsdx_hip_cholesky_batched_f32(p, 65, 1, None) == _lib.EINVAL
    # machine probes: unknown probe kinds and null outputs are argument errors
    v = C.c_double(0.0)
    assert L.tsdx_hip_measure_bandwidth_gbs(7, C.byref(v)) == _lib.EINVAL
    assert L.tsdx_hip_measure_peak_gflops(0, None) == _lib.EINVAL
    assert L.tsdx_hip_measure_mfma_tflops(9, C.byref(v)) == _lib.EINVAL
    assert L.tsdx_hip_lu_batched_f64(p, 3, 1, None, None) == _lib.EINVAL
    assert L.tsdx_hip_posv_ir_workspace_size(1000) >= 4 * 1000
    assert L.tsdx_hip_posv_ir_f64(p, 65, 1, p, p, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_posv_ir_f64(p, 8, 1, p, p, 1, 7, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert "factor" in _lib.last_error()
    assert L.tsdx_hip_posv_ir_f64(p, 8, 1, p, p, 1, 1, 10, -1.0, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_posv_ir_f64(p, 8, 1, p, p, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert "workspace" in _lib.last_error()
    assert L.tsdx_hip_gesv_ir_workspace_size(1000) >= 4 * 1000
    assert L.tsdx_hip_gesv_ir_f64(p, 65, 1, p, p, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_gesv_ir_f64(p, 8, 1, p, p, 1, 5, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert "factor" in _lib.last_error()
    assert L.tsdx_hip_gesv_ir_f64(p, 8, 1, p, p, 1, 1, 1001, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_gesv_ir_f64(p, 8, 1, p, p, 1, 1, 10, 0.0, None, None, None, 0, None) == _lib.EINVAL
    assert L.tsdx_hip_gesv_ir_f64(p, 8, 1, p, p, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.EINVAL
    assert "workspace" in _lib.last_error()
    assert L.tsdx_hip_gesv_ir_f64(None, 8, 0, None, None, 1, 1, 10, 1e-14, None, None, None, 0, None) == _lib.OK


def test_python_error_types(tsdx):
    # tests/python/test_solvers.py:108-119 of the reference
    with pytest.raises(ValueError):
        tsdx.cholesky(np.zeros((1, 65, 65), np.float64))
    with pytest.raises(ValueError):
        tsdx.cholesky(np.zeros((1, 3, 4), np.float64))
    with pytest.raises(ValueError):
        tsdx.cholesky(np.eye(3)[None], precision="fp8")
    with pytest.raises(TypeError):
        tsdx.cholesky(np.eye(3, dtype=np.int32))
    with pytest.raises(RuntimeError):
        tsdx.cholesky(np.eye(3)[None], backend="cuda")
    for be in ("naive", "scalar", "simd"):
        with pytest.raises(RuntimeError, match="CPU backend"):
            tsdx.lu(np.eye(3)[None], backend=be)
    with pytest.raises(ValueError):
        tsdx.lu(np.eye(3)[None], backend="tpu")
    for be in ("naive", "scalar", "simd"):  # emulated formats take the same backend rules
        with pytest.raises(RuntimeError, match="CPU backend"):
            tsdx.cholesky(np.eye(3, dtype=np.float32)[None], precision="bf16", backend=be)


def test_no_device_fails_loudly(tsdx):
    if tsdx.has_hip():
        pytest.skip("a HIP device is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        tsdx.cholesky(np.eye(3)[None])
    with pytest.raises(RuntimeError, match="no HIP device"):
        tsdx.gesv(np.eye(3)[None], np.ones((1, 3)))
    with pytest.raises(RuntimeError, match="no HIP device"):
        tsdx.cholesky(np.eye(3, dtype=np.float16)[None])           # emulated fp16: no CPU path either

    class DeviceLike:  # an array-interface producer (the pointer is never dereferenced here)
        __cuda_array_interface__ = dict(shape=(2, 3, 3), typestr="<f4", data=(4096, False), strides=None, version=2)
    with pytest.raises(RuntimeError, match="no HIP device"):
        tsdx.cholesky(DeviceLike())


def test_device_array_detection(tsdx):
    import torch
    assert not tsdx._core.is_device_array(np.eye(3))
    assert not tsdx._core.is_device_array(torch.eye(3))            # CPU tensors stay on the host path

    class DeviceLike:
        __cuda_array_interface__ = dict(shape=(3, 3), typestr="<f8", data=(4096, False), version=3)
    assert tsdx._core.is_device_array(DeviceLike())
    assert tsdx.DeviceArray.__name__ == "DeviceArray"


def test_round_to_and_unit_roundoff_match_reference(tsdx):
    g = golden("precision")
    for p in ("fp32", "fp16", "bf16", "fp64"):
        assert bits_equal(tsdx.round_to(g["x64"], p), g[f"round_to_{p}"]), p
    x = g["x"].astype(np.float32)
    assert bits_equal(tsdx._core.round_to_f32(x, "fp16"), g["fp16"])
    assert bits_equal(tsdx._core.round_to_f32(x, "bf16"), g["bf16"])
    assert [tsdx.unit_roundoff(p) for p in ("fp32", "fp64", "fp16", "bf16")] == list(g["u"])


def test_host_metrics_match_reference(tsdx):
    g = golden("metrics_f64")
    for c in golden_cases(g, "_A"):
        A, S, b = g[f"{c}_A"], g[f"{c}_S"], g[f"{c}_b"]
        np.testing.assert_array_equal(tsdx.lu_backward_error(A, g[f"{c}_LU"], g[f"{c}_piv"]), g[f"{c}_lu_bwd"])
        np.testing.assert_array_equal(tsdx.chol_backward_error(S, g[f"{c}_L"]), g[f"{c}_chol_bwd"])
        np.testing.assert_array_equal(tsdx.solve_backward_error(A, g[f"{c}_xp"], b), g[f"{c}_solve_bwd"])
        np.testing.assert_array_equal(tsdx.solve_forward_error(A, g[f"{c}_xp"], b), g[f"{c}_solve_fwd"])
        np.testing.assert_array_equal(tsdx.reference_solve(A, b), g[f"{c}_x"])


def test_metric_sanity_closed_forms(tsdx):
    # exact solutions of diagonal systems have zero residual and zero error
    d = 2.0 ** np.arange(5)  # powers of two: x = b / d is exact
    A = np.stack([np.diag(d)] * 2)
    b = np.ones((2, 5))
    x = b / d
    assert tsdx.solve_backward_error(A, x, b).max() == 0.0
    assert tsdx.solve_forward_error(A, x, b).max() == 0.0
    assert tsdx.solve_forward_error(A, 1.01 * x, b).min() > 5e-3


def test_matrix_generators_have_prescribed_condition():
    from tinysolverdx import matrices
    rng = np.random.default_rng(0)
    S = matrices.spd(rng, 5, 12, 1e4)
    assert np.allclose(S, S.transpose(0, 2, 1))
    np.testing.assert_allclose(np.linalg.cond(S), 1e4, rtol=1e-6)
    G = matrices.general_kappas(rng, np.array([10.0, 1e3]), 9)
    np.testing.assert_allclose(np.linalg.cond(G), [10.0, 1e3], rtol=1e-6)


def test_ragged_plan_argument_errors(tsdx):
    """Checks that need no device: null buffers, negative count, null plan."""
    import ctypes as C
    from tinysolverdx import _lib
    L = _lib.lib()
    plan = C.c_void_p()
    assert L.tsdx_hip_ragged_plan_create_f32(None, None, None, None, 3, None, None, None, 1, C.byref(plan), None) == _lib.EINVAL
    assert "null" in _lib.last_error()
    assert L.tsdx_hip_ragged_plan_create_f64(None, None, None, None, -1, None, None, None, 1, C.byref(plan), None) == _lib.EINVAL
    assert L.tsdx_hip_ragged_plan_create_f32(None, None, None, None, 0, None, None, None, 1, None, None) == _lib.EINVAL
    assert L.tsdx_hip_ragged_run(None, None) == _lib.EINVAL
    assert L.tsdx_hip_ragged_plan_report(None) == b""
    assert L.tsdx_hip_ragged_plan_destroy(None) == _lib.OK
