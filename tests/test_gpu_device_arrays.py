"""Device arrays in the public API: matrices already in GPU memory go through
tinysolverdx.cholesky / lu / *_solve / posv / gesv without a host copy, from
any producer of __cuda_array_interface__ or DLPack, and come back as
DeviceArray (library-owned GPU memory exporting both protocols).  Results are
bit-identical to the host-array calls, inputs are never written, and the
exports are zero-copy (same device pointer)."""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _spd(rng, b, n, dt):
    from tinysolverdx import matrices
    return matrices.spd(rng, b, n, 1e2).astype(dt)


class DLPackOnly:
    """A producer that speaks DLPack only (no array interface)."""

    def __init__(self, t):
        self.t = t

    def __dlpack__(self, *args, **kwargs):
        return self.t.__dlpack__(*args, **kwargs)

    def __dlpack_device__(self):
        return self.t.__dlpack_device__()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 5, 16, 33, 64])
def test_factor_and_solve_match_host_calls(tsdx, dt, n):
    import torch
    rng = np.random.default_rng(n)
    S = _spd(rng, 40, n, dt)
    G = rng.standard_normal((40, n, n)).astype(dt)
    B = rng.standard_normal((40, n, 3)).astype(dt)
    dS, dG, dB = (torch.from_numpy(x).cuda() for x in (S, G, B))
    L, info = tsdx.cholesky(dS)
    Lh, infoh = tsdx.cholesky(S)
    assert isinstance(L, tsdx.DeviceArray) and L.shape == S.shape and L.dtype == dt
    assert bits_equal(L.numpy(), Lh) and np.array_equal(info.numpy(), infoh)
    LU, piv, linfo = tsdx.lu(dG)
    LUh, pivh, linfoh = tsdx.lu(G)
    assert bits_equal(LU.numpy(), LUh) and np.array_equal(piv.numpy(), pivh) and np.array_equal(linfo.numpy(), linfoh)
    X = tsdx.cholesky_solve(L, dB)
    assert bits_equal(X.numpy(), tsdx.cholesky_solve(Lh, B))
    Y = tsdx.lu_solve(LU, piv, dB)
    assert bits_equal(Y.numpy(), tsdx.lu_solve(LUh, pivh, B))
    xs, sinfo = tsdx.posv(dS, dB)
    assert bits_equal(xs.numpy(), tsdx.posv(S, B)[0]) and (sinfo.numpy() == 0).all()
    xg, ginfo = tsdx.gesv(dG, dB)
    assert bits_equal(xg.numpy(), tsdx.gesv(G, B)[0]) and np.array_equal(ginfo.numpy(), linfoh)
    # inputs untouched
    assert np.array_equal(dS.cpu().numpy(), S) and np.array_equal(dB.cpu().numpy(), B)


def test_zero_copy_exports(tsdx):
    import torch
    rng = np.random.default_rng(1)
    S = _spd(rng, 8, 12, np.float64)
    L, info = tsdx.cholesky(torch.from_numpy(S).cuda())
    ptr = L.__cuda_array_interface__["data"][0]
    t = torch.from_dlpack(L)
    assert t.is_cuda and t.data_ptr() == ptr and t.dtype == torch.float64 and tuple(t.shape) == S.shape
    u = torch.as_tensor(L, device="cuda")           # through the array interface
    assert u.data_ptr() == ptr
    del L                                           # the torch views keep the memory alive
    assert bits_equal(t.cpu().numpy(), tsdx.cholesky(S)[0]) and bits_equal(u.cpu().numpy(), t.cpu().numpy())
    assert info.__dlpack_device__() == (10, torch.cuda.current_device())
    assert repr(info).startswith("DeviceArray(shape=(8,), dtype=int32")


def test_dlpack_only_producer_and_chaining(tsdx):
    import torch
    rng = np.random.default_rng(2)
    G = rng.standard_normal((20, 9, 9)).astype(np.float32)
    b = rng.standard_normal((20, 9)).astype(np.float32)
    LU, piv, info = tsdx.lu(DLPackOnly(torch.from_numpy(G).cuda()))
    y = tsdx.lu_solve(LU, piv, DLPackOnly(torch.from_numpy(b).cuda()))   # DeviceArray inputs chain directly
    assert y.shape == (20, 9)
    assert bits_equal(y.numpy(), tsdx.lu_solve(*tsdx.lu(G)[:2], b))


def test_single_matrix_and_vector_shapes(tsdx):
    import torch
    rng = np.random.default_rng(3)
    S = _spd(rng, 1, 7, np.float64)[0]
    b = rng.standard_normal(7)
    L, info = tsdx.cholesky(torch.from_numpy(S).cuda())
    assert L.shape == (7, 7) and info.shape == () and int(info.numpy()) == 0
    x, info = tsdx.posv(torch.from_numpy(S).cuda(), torch.from_numpy(b).cuda())
    assert x.shape == (7,) and bits_equal(x.numpy(), tsdx.posv(S, b)[0])
    LU, piv, _ = tsdx.lu(torch.from_numpy(S).cuda())
    assert piv.shape == (7,)
    y = tsdx.lu_solve(LU, piv, torch.from_numpy(np.stack([b, 2 * b], 1)).cuda())
    assert y.shape == (7, 2)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_emulated_precisions_on_device_arrays(tsdx, oracle, prec):
    import torch
    rng = np.random.default_rng(4)
    S = _spd(rng, 16, 10, np.float32)
    G = rng.standard_normal((16, 10, 10)).astype(np.float32)
    B = rng.standard_normal((16, 10, 2)).astype(np.float32)
    L, info = tsdx.cholesky(torch.from_numpy(S).cuda(), precision=prec)
    assert bits_equal(L.numpy(), oracle.cholesky_p(S, prec)[0])
    x, _ = tsdx.posv(torch.from_numpy(S).cuda(), torch.from_numpy(B).cuda(), precision=prec)
    assert bits_equal(x.numpy(), tsdx.posv(S, B, precision=prec)[0])
    y, _ = tsdx.gesv(torch.from_numpy(G).cuda(), torch.from_numpy(B).cuda(), precision=prec)
    assert bits_equal(y.numpy(), tsdx.gesv(G, B, precision=prec)[0])


def test_device_array_errors(tsdx):
    import torch
    A = torch.eye(4, device="cuda", dtype=torch.float32)[None].repeat(3, 1, 1)
    with pytest.raises(ValueError, match="C-contiguous"):
        tsdx.cholesky(torch.eye(8, device="cuda")[None].repeat(3, 1, 1)[:, ::2, ::2])
    with pytest.raises(TypeError):
        tsdx.cholesky(A.half())
    with pytest.raises(ValueError, match="precision"):
        tsdx.cholesky(A, precision="fp64")                    # float32 storage: fp32 or emulated only
    with pytest.raises(ValueError):
        tsdx.cholesky(torch.zeros((2, 3, 4), device="cuda"))
    with pytest.raises(ValueError):
        tsdx.posv(A, torch.zeros((3, 5), device="cuda"))       # rhs rows
    with pytest.raises(TypeError):
        tsdx.posv(A, torch.zeros((3, 4), device="cuda", dtype=torch.float64))
    with pytest.raises(RuntimeError, match="CPU backend"):
        tsdx.cholesky(A, backend="scalar")
    with pytest.raises(TypeError):
        tsdx.cholesky_solve(A, np.zeros((3, 4), np.float32))   # host rhs with device factors
    with pytest.raises(ValueError, match="piv"):
        LU, piv, _ = tsdx.lu(A)
        tsdx.lu_solve(LU, torch.zeros((3, 5), device="cuda", dtype=torch.int32), torch.zeros((3, 4), device="cuda"))
