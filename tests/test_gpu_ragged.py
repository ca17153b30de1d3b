"""Ragged batches: sizes 1..64 and both kinds in one graph launch.

Every factor, pivot and info must be bit-identical to the oracle (the
reference's Auto backend restated: kernels.hpp:18-180, batched.hpp:121-244)
run on each (kind, n) group as a dense batch, and to the per-size launches
(tsdx_hip_potrf / tsdx_hip_getrf) of the same matrices.  The inputs include
matrices that are not SPD, zero columns and NaNs.  The graph is replayed
several times with eager work and synchronisations in between, each replay
checked.
"""
import numpy as np
import pytest

from conftest import bits_equal, first_bit_mismatch

pytestmark = pytest.mark.gpu


def _zipf_sizes(rng, count, lo=1, hi=64):
    n = np.arange(lo, hi + 1)
    p = 1.0 / n
    return rng.choice(n, size=count, p=p / p.sum()).astype(np.int32)


def _batch(rng, count, dtype):
    """Flat packed storage (offsets not aligned), half Cholesky half LU, with
    failing matrices of both kinds."""
    from tinysolverdx import matrices
    sizes = _zipf_sizes(rng, count)
    kinds = (rng.random(count) < 0.5).astype(np.int32)  # 1 = LU
    offs = np.zeros(count, np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)[:-1]
    A = np.zeros(int((sizes.astype(np.int64) ** 2).sum()), dtype)
    for m in range(count):
        n = int(sizes[m])
        if kinds[m] == 0:
            M = matrices.spd_kappas(rng, 10.0 ** rng.uniform(0, 4, 1), n)[0]
            if m % 37 == 3:
                M[n // 2, n // 2] = -1.0  # not SPD
        else:
            M = rng.standard_normal((n, n))
            if m % 41 == 5:
                M[:, rng.integers(n)] = 0.0  # zero column: info > 0
        if m % 97 == 11:
            M[0, n - 1] = np.nan
        A[offs[m]:offs[m] + n * n] = M.astype(dtype).ravel()
    poffs = np.zeros(count, np.int64)
    poffs[1:] = np.cumsum(sizes.astype(np.int64))[:-1]
    return A, offs, sizes, kinds, poffs


def _per_size_reference(A, offs, sizes, kinds, poffs):
    import torch
    from tinysolverdx import device
    out = A.copy()
    piv = np.zeros(int(sizes.sum()), np.int32)
    info = np.zeros(len(sizes), np.int32)
    for kind in (0, 1):
        for n in np.unique(sizes[kinds == kind]):
            ids = np.nonzero((sizes == n) & (kinds == kind))[0]
            dense = np.stack([A[offs[m]:offs[m] + n * n].reshape(n, n) for m in ids])
            dA = torch.from_numpy(dense).cuda()
            dI = torch.zeros(len(ids), dtype=torch.int32, device="cuda")
            if kind == 0:
                device.potrf(dA, dI)
            else:
                dP = torch.zeros((len(ids), int(n)), dtype=torch.int32, device="cuda")
                device.getrf(dA, dP, dI)
                P = dP.cpu().numpy()
                for r, m in enumerate(ids):
                    piv[poffs[m]:poffs[m] + n] = P[r]
            torch.cuda.synchronize()
            F, I = dA.cpu().numpy(), dI.cpu().numpy()
            for r, m in enumerate(ids):
                out[offs[m]:offs[m] + n * n] = F[r].ravel()
                info[m] = I[r]
    return out, piv, info


def _oracle_reference(oracle, A, offs, sizes, kinds, poffs):
    """The oracle on every (kind, n) group of the ragged batch, scattered back
    to the ragged layout."""
    out = A.copy()
    piv = np.zeros(int(sizes.sum()), np.int32)
    info = np.zeros(len(sizes), np.int32)
    for kind in (0, 1):
        for n in np.unique(sizes[kinds == kind]):
            n = int(n)
            ids = np.nonzero((sizes == n) & (kinds == kind))[0]
            dense = np.stack([A[offs[m]:offs[m] + n * n].reshape(n, n) for m in ids])
            if kind == 0:
                F, I = oracle.cholesky(dense)
            else:
                F, P, I = oracle.lu(dense)
                for r, m in enumerate(ids):
                    piv[poffs[m]:poffs[m] + n] = P[r]
            for r, m in enumerate(ids):
                out[offs[m]:offs[m] + n * n] = F[r].ravel()
                info[m] = I[r]
    return out, piv, info


def _check(got, want, sizes, kinds):
    gA, gP, gI = got
    wA, wP, wI = want
    assert np.array_equal(gI, wI), np.nonzero(gI != wI)[0][:10]
    mask = np.concatenate([np.full(int(n), bool(k)) for n, k in zip(sizes, kinds)])
    assert np.array_equal(gP[mask], wP[mask])
    assert bits_equal(gA, wA), first_bit_mismatch(gA, wA)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("tune", [False, True])
def test_ragged_matches_oracle_and_per_size_launches(oracle, dtype, tune):
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(31 + (dtype == np.float64) + 2 * tune)
    A, offs, sizes, kinds, poffs = _batch(rng, 3000, dtype)
    print(f"[ragged] {np.dtype(dtype).name} tune={tune} count={len(sizes)}", flush=True)
    want = _oracle_reference(oracle, A, offs, sizes, kinds, poffs)
    assert (want[2] > 0).any() and ((want[2] > 0) & (kinds == 1)).any()  # failures of both kinds are in the batch
    _check(_per_size_reference(A, offs, sizes, kinds, poffs), want, sizes, kinds)
    dA0 = torch.from_numpy(A).cuda()
    dA = dA0.clone()
    d_off, d_poff = torch.from_numpy(offs).cuda(), torch.from_numpy(poffs).cuda()
    d_size, d_kind = torch.from_numpy(sizes).cuda(), torch.from_numpy(kinds).cuda()
    d_piv = torch.full((int(sizes.sum()),), -5, dtype=torch.int32, device="cuda")
    d_info = torch.full((len(sizes),), -5, dtype=torch.int32, device="cuda")
    plan = device.RaggedPlan(dA, d_off, d_size, d_kind, d_piv, d_poff, d_info, tune=tune)
    rep = plan.report
    assert rep["count"] == len(sizes)
    assert sum(rep[f"class{c}"]["matrices"] for c in range(3)) == len(sizes)
    for replay in range(3):
        plan.run()
        torch.cuda.synchronize()
        got = (dA.cpu().numpy(), d_piv.cpu().numpy(), d_info.cpu().numpy())
        _check(got, want, sizes, kinds)
        # eager work between replays: restore the input, scribble on outputs
        dA.copy_(dA0)
        d_piv.fill_(-7)
        d_info.fill_(-7)
        torch.cuda.synchronize()
    plan.close()


def test_ragged_plan_keeps_its_sizes(oracle):
    """The plan reads sizes and kinds once: changing the caller's arrays after
    planning changes nothing in later runs."""
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(77)
    A, offs, sizes, kinds, poffs = _batch(rng, 400, np.float64)
    want = _oracle_reference(oracle, A, offs, sizes, kinds, poffs)
    dA0 = torch.from_numpy(A).cuda()
    dA = dA0.clone()
    d_size, d_kind = torch.from_numpy(sizes).cuda(), torch.from_numpy(kinds).cuda()
    d_piv = torch.zeros(int(sizes.sum()), dtype=torch.int32, device="cuda")
    d_info = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
    plan = device.RaggedPlan(dA, torch.from_numpy(offs).cuda(), d_size, d_kind, d_piv,
                             torch.from_numpy(poffs).cuda(), d_info, tune=False)
    d_size.fill_(1)
    d_kind.fill_(0)
    plan.run()
    torch.cuda.synchronize()
    _check((dA.cpu().numpy(), d_piv.cpu().numpy(), d_info.cpu().numpy()), want, sizes, kinds)
    plan.close()


def test_ragged_rejects_buffers_too_small():
    import torch
    from tinysolverdx import device
    A = torch.zeros(20, device="cuda")
    sizes = torch.tensor([3, 4], dtype=torch.int32, device="cuda")
    kinds = torch.tensor([0, 1], dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="outside A"):
        device.RaggedPlan(A, torch.tensor([0, 9], dtype=torch.int64, device="cuda"), sizes, kinds)
    with pytest.raises(ValueError, match="outside piv"):
        device.RaggedPlan(torch.zeros(25, device="cuda"), torch.tensor([0, 9], dtype=torch.int64, device="cuda"),
                          sizes, kinds, torch.zeros(4, dtype=torch.int32, device="cuda"),
                          torch.tensor([0, 3], dtype=torch.int64, device="cuda"))
    with pytest.raises(ValueError, match="shape"):
        device.RaggedPlan(A, torch.zeros(3, dtype=torch.int64, device="cuda"), sizes, kinds)


def test_ragged_single_size_and_empty():
    import torch
    from tinysolverdx import device
    rng = np.random.default_rng(5)
    # one size, one kind: the same as one per-size launch
    n, count = 32, 500
    sizes = np.full(count, n, np.int32)
    kinds = np.ones(count, np.int32)
    A = rng.standard_normal(count * n * n).astype(np.float32)
    offs = np.arange(count, dtype=np.int64) * n * n
    poffs = np.arange(count, dtype=np.int64) * n
    want = _per_size_reference(A, offs, sizes, kinds, poffs)
    dA = torch.from_numpy(A).cuda()
    d_piv = torch.zeros(count * n, dtype=torch.int32, device="cuda")
    d_info = torch.zeros(count, dtype=torch.int32, device="cuda")
    plan = device.RaggedPlan(dA, torch.from_numpy(offs).cuda(), torch.from_numpy(sizes).cuda(),
                             torch.from_numpy(kinds).cuda(), d_piv, torch.from_numpy(poffs).cuda(), d_info)
    plan.run()
    torch.cuda.synchronize()
    _check((dA.cpu().numpy(), d_piv.cpu().numpy(), d_info.cpu().numpy()), want, sizes, kinds)
    plan.close()
    e64, e32 = torch.zeros(0, dtype=torch.int64, device="cuda"), torch.zeros(0, dtype=torch.int32, device="cuda")
    empty = device.RaggedPlan(torch.zeros(1, device="cuda"), e64, e32, e32)
    empty.run()
    assert empty.report["count"] == 0
    empty.close()


def test_ragged_rejects_bad_sizes_and_kinds():
    import torch
    from tinysolverdx import device
    A = torch.zeros(100, device="cuda")
    off = torch.zeros(2, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="outside supported range"):
        device.RaggedPlan(A, off, torch.tensor([3, 65], dtype=torch.int32, device="cuda"),
                          torch.zeros(2, dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError, match="unknown kind"):
        device.RaggedPlan(A, off, torch.tensor([3, 4], dtype=torch.int32, device="cuda"),
                          torch.tensor([0, 2], dtype=torch.int32, device="cuda"))
