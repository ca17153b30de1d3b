"""Host-side argument checks of tinysolverdx.device and the router: every
wrong shape, dtype, device or workspace size is rejected with ValueError /
TypeError before anything is launched (the kernels trust the sizes they are
given, so these are the guard against out-of-bounds device writes)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(shape, dtype="float64"):
    import torch
    return torch.zeros(shape, dtype=getattr(torch, dtype), device="cuda")


def test_factor_and_solve_shapes():
    from tinysolverdx import device
    A, B = _t((4, 8, 8)), _t((4, 8, 2))
    cases = [
        (ValueError, lambda: device.potrf(_t((4, 8, 7)))),                       # not square
        (ValueError, lambda: device.potrf(A, _t((5,), "int32"))),                # info length
        (TypeError, lambda: device.potrf(A, _t((4,), "int64"))),                 # info dtype
        (ValueError, lambda: device.getrf(A, _t((4, 7), "int32"))),              # piv shape
        (ValueError, lambda: device.getrf(A, None)),                             # piv required
        (ValueError, lambda: device.potrs(A, _t((4, 8)))),                       # rhs must be 3-D
        (ValueError, lambda: device.potrs(A, _t((4, 7, 2)))),                    # rhs rows
        (ValueError, lambda: device.posv(A, _t((3, 8, 2)))),                     # rhs batch
        (TypeError, lambda: device.posv(A, _t((4, 8, 2), "float32"))),           # rhs dtype
        (ValueError, lambda: device.posv(A, _t((4, 8, 0)))),                     # no rhs
        (ValueError, lambda: device.gesv(A, B, _t((4, 9), "int32"))),            # piv shape
        (ValueError, lambda: device.posv_solve(A, B, _t((4, 8, 3)))),            # X shape
        (ValueError, lambda: device.gesv_solve(A, B, _t((4, 8, 2)), _t((4, 8, 1), "int32"))),
        (ValueError, lambda: device.posv(A.cpu(), B)),                           # host tensor
        (ValueError, lambda: device.posv(_t((4, 8, 16))[:, :, ::2], B)),         # not contiguous
        (TypeError, lambda: device.potrf(_t((4, 8, 8), "float16"))),
    ]
    for exc, call in cases:
        with pytest.raises(exc):
            call()


def test_posv_ir_and_router_shapes():
    from tinysolverdx import device, router
    A, B, X = _t((4, 64, 64)), _t((4, 64, 1)), _t((4, 64, 1))
    with pytest.raises(TypeError):
        device.posv_ir(_t((4, 64, 64), "float32"), B, X)
    with pytest.raises(ValueError):
        device.posv_ir(A, B, _t((4, 64, 2)))
    with pytest.raises(ValueError):
        device.posv_ir(A, B, X, stat=_t((3,), "int32"))
    with pytest.raises(ValueError):
        device.posv_ir(A, B, X, work=_t((8,), "uint8"))                          # workspace too small
    with pytest.raises(ValueError):
        device.posv_ir(A, B, X, factor="fp8")
    with pytest.raises(TypeError):
        device.gesv_ir(_t((4, 64, 64), "float32"), B, X)
    with pytest.raises(ValueError):
        device.gesv_ir(A, B, _t((4, 64, 2)))
    with pytest.raises(ValueError):
        device.gesv_ir(A, B, X, info=_t((5,), "int32"))
    with pytest.raises(ValueError):
        device.gesv_ir(A, B, X, work=_t((8,), "uint8"))                          # workspace too small
    with pytest.raises(ValueError):
        device.gesv_ir(A, B, X, factor="fp64_block")                              # no block LU tier
    r = router.Router(device="cuda")
    with pytest.raises(ValueError):
        r.posv(A, B, X, tier=_t((5,), "int32"))
    with pytest.raises(ValueError):
        r.posv(A, B, X, feat=_t((4, 11), "float32"))
    with pytest.raises(ValueError):
        r.forward(_t((4, 11), "float32"))
    with pytest.raises(ValueError):
        r.grad(_t((4, 12), "float32"), _t((3,), "int32"))
