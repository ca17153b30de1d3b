"""The allocation-free refinement posv (tsdx_hip_posv_ir_noalloc_f64) against
the workspace form (tsdx_hip_posv_ir_f64) through the C ABI: identical
solutions, info and stat bit for bit, on two calls in a row (the second with a
smaller batch, so the shared scratch array holds the first call's list), with
matrices that need the float64 fallback in both."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits_equal, first_bit_mismatch

pytestmark = pytest.mark.gpu


def _case(n, batch, seed):
    import torch
    from tinysolverdx import matrices
    rng = np.random.default_rng(seed)
    kappa = 10.0 ** rng.uniform(1.0, 12.0, batch)  # the top of the range defeats fp32 refinement
    A = torch.from_numpy(matrices.spd_kappas(rng, kappa, n)).cuda()
    B = torch.from_numpy(rng.standard_normal((batch, n, 2))).cuda()
    return A, B


def _run(lib, A, B, factor, noalloc):
    import torch
    from tinysolverdx import _lib
    batch, n, nrhs = A.shape[0], A.shape[1], B.shape[2]
    X = torch.full_like(B, -3.0)
    info = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    stat = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = lib.tsdx_hip_posv_ir_noalloc_f64(A.data_ptr(), n, batch, B.data_ptr(), X.data_ptr(), nrhs,
                                              _lib.FACTOR[factor], 10, 1e-14, info.data_ptr(), stat.data_ptr(), s)
    else:
        work = torch.empty(lib.tsdx_hip_posv_ir_workspace_size(batch), dtype=torch.uint8, device="cuda")
        rc = lib.tsdx_hip_posv_ir_f64(A.data_ptr(), n, batch, B.data_ptr(), X.data_ptr(), nrhs, _lib.FACTOR[factor],
                                      10, 1e-14, info.data_ptr(), stat.data_ptr(), work.data_ptr(), work.numel(), s)
    _lib.check(rc)
    torch.cuda.synchronize()
    return X.cpu().numpy(), info.cpu().numpy(), stat.cpu().numpy()


@pytest.mark.parametrize("n,factor", [(64, "fp32"), (24, "fp32"), (64, "bf16"), (40, "fp64")])
def test_noalloc_equals_workspace_form(hip, n, factor):
    lib = hip.lib()
    for batch, seed in ((3000, 1), (700, 2)):  # second call smaller than the first
        print(f"[noalloc] n={n} factor={factor} batch={batch}", flush=True)
        A, B = _case(n, batch, seed + n)
        want = _run(lib, A, B, factor, noalloc=False)
        got = _run(lib, A, B, factor, noalloc=True)
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
        assert bits_equal(got[0], want[0]), first_bit_mismatch(got[0], want[0])
        if factor != "fp64":
            assert (want[2] == -1).any() and (want[2] >= 0).any()  # both the refinement and the fallback ran


def test_noalloc_rejects_batches_beyond_its_array(hip):
    from tinysolverdx import _lib
    lib = hip.lib()
    big = (1 << 22) + 1
    rc = lib.tsdx_hip_posv_ir_noalloc_f64(None, 64, big, None, None, 1, 1, 10, 1e-14, None, None, None)
    assert rc == _lib.EINVAL and "allocation-free limit" in _lib.last_error()
