"""Learned precision router: HIP features / forward / gradient / Adam against
the float64 CPU oracle (oracle/mlp_oracle.py, itself pinned by finite
differences), training, and the routed posv.

The gain test times the shipped router's real calls (features and network
included) against the fixed tiers and the network alone on fresh matrices.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(rng, batch, n, lo, hi):
    from tinysolverdx import matrices
    kappa = 10.0 ** rng.uniform(lo, hi, batch)
    return matrices.spd_kappas(rng, kappa, n), kappa


def test_features_match_oracle():
    import torch
    import mlp_oracle as M
    from tinysolverdx import router
    rng = np.random.default_rng(1)
    for n in (3, 16, 33, 64):
        A, _ = _spd(rng, 20, n, 0, 10)
        A[3] = -A[3]                      # negative diagonal
        A[5] = np.diag(rng.uniform(1, 2, n))
        iu = np.triu_indices(n, 1)
        A[7][iu] = 1e20                   # upper triangle must not matter
        got = router.features(torch.from_numpy(A).cuda()).cpu().numpy()
        want = M.features(A)
        assert np.allclose(got, want, rtol=1e-4, atol=2e-4), (n, np.abs(got - want).max(axis=0))


def test_forward_and_gradient_match_oracle():
    import torch
    import mlp_oracle as M
    from tinysolverdx import router
    rng = np.random.default_rng(2)
    p = router.init_params(3)
    p[0:12] = rng.standard_normal(12)
    p[12:24] = rng.uniform(0.5, 2, 12)
    p[router.OFF_B1:router.OFF_W2] = rng.standard_normal(64) * 0.1
    r = router.Router(p)
    for count in (1, 63, 64, 1000):  # partial and several chunks
        feat = rng.standard_normal((count, 12)).astype(np.float32)
        label = rng.integers(0, 3, count).astype(np.int32)
        f, l = torch.from_numpy(feat).cuda(), torch.from_numpy(label).cuda()
        z = r.forward(f).cpu().numpy()
        zo = M.forward(p, feat)[3]
        assert np.allclose(z, zo, rtol=1e-4, atol=1e-4), np.abs(z - zo).max()
        g, loss = r.grad(f, l)
        go = M.grad(p, feat, label)
        g = g.cpu().numpy()
        assert np.all(g[:24] == 0)
        scale = np.abs(go).max()
        assert np.abs(g - go).max() <= 1e-4 * scale + 1e-6, (count, np.abs(g - go).max(), scale)
        assert abs(loss.item() - M.loss(p, feat, label)) < 1e-4


def test_matrix_core_inference_matches_oracle():
    """tsdx_hip_router_infer (bfloat16 network on the matrix cores, the
    routed solve's decision path) against the oracle's restatement, and its
    logits within bfloat16 rounding of the float32 forward pass."""
    import torch
    import mlp_oracle as M
    from tinysolverdx import router
    rng = np.random.default_rng(12)
    p = router.init_params(5)
    p[0:12] = rng.standard_normal(12)
    p[12:24] = rng.uniform(0.5, 2, 12)
    p[router.OFF_B1:router.OFF_W2] = rng.standard_normal(64) * 0.1
    r = router.Router(p)
    for count in (1, 15, 16, 17, 1000):  # partial and several 16-sample columns
        feat = (rng.standard_normal((count, 12)) * 2).astype(np.float32)
        f = torch.from_numpy(feat).cuda()
        tier, z = r.infer(f)
        z, tier = z.cpu().numpy(), tier.cpu().numpy()
        zo = M.infer_bf16(p, feat)
        assert np.allclose(z, zo, rtol=1e-5, atol=1e-5 * np.abs(zo).max()), np.abs(z - zo).max()
        top2 = np.sort(zo, axis=1)[:, -2:]
        clear = (top2[:, 1] - top2[:, 0]) > 1e-4 * np.abs(zo).max()
        assert np.array_equal(tier[clear], M.tiers(zo)[clear])
        assert np.allclose(z, r.forward(f).cpu().numpy(), rtol=5e-2, atol=5e-2 * np.abs(zo).max())


def test_adam_step_matches_oracle():
    import torch
    import mlp_oracle as M
    from tinysolverdx import _lib, router
    rng = np.random.default_rng(4)
    p = router.init_params(5)
    g = rng.standard_normal(router.PARAMS).astype(np.float32)
    g[:24] = 0
    m = rng.standard_normal(router.PARAMS).astype(np.float32) * 0.1
    v = rng.uniform(0, 0.1, router.PARAMS).astype(np.float32)
    dp, dg, dm, dv = (torch.from_numpy(x.copy()).cuda() for x in (p, g, m, v))
    _lib.call("tsdx_hip_router_adam", dp.data_ptr(), dg.data_ptr(), dm.data_ptr(), dv.data_ptr(), 7, 1e-3,
              torch.cuda.current_stream().cuda_stream)
    # the kernel's constants are float32 (1 - beta2 differs from 1e-3 by 1.3e-5 relative)
    b1, b2 = float(np.float32(0.9)), float(np.float32(0.999))
    po, mo, vo = M.adam(p, g.astype(np.float64), m, v, 7, float(np.float32(1e-3)), b1=b1, b2=b2)
    assert np.allclose(dp.cpu().numpy(), po, rtol=1e-5, atol=1e-6)
    assert np.allclose(dm.cpu().numpy(), mo, rtol=1e-5, atol=1e-7)
    assert np.allclose(dv.cpu().numpy(), vo, rtol=1e-5, atol=1e-9)
    assert np.array_equal(dp.cpu().numpy()[:24], p[:24])


def test_training_learns_a_separable_rule():
    """Label = which of three feature bands feature 1 falls in; the loss must
    fall and the accuracy be high after a short fit."""
    import torch
    from tinysolverdx import router
    rng = np.random.default_rng(6)
    feat = rng.standard_normal((8192, 12)).astype(np.float32)
    label = np.digitize(feat[:, 1], [-0.5, 0.5]).astype(np.int32)
    f, l = torch.from_numpy(feat).cuda(), torch.from_numpy(label).cuda()
    r = router.Router(seed=1)
    hist = r.fit(f, l, epochs=30, batch=1024, lr=3e-3)
    assert hist[-1] < 0.5 * hist[0], hist
    acc = (r.predict(f).cpu().numpy() == label).mean()
    assert acc > 0.95, acc


def _bwd(tsdx, A, X, B):
    return np.stack([tsdx.solve_backward_error(A, X[..., r], B[..., r]) for r in range(B.shape[-1])]).max(axis=0)


def test_routed_posv_every_answer_meets_tolerance(tsdx):
    import torch
    from tinysolverdx import router
    rng = np.random.default_rng(7)
    A, _ = _spd(rng, 96, 32, 0, 12)
    A[10] = -np.eye(32)  # not SPD: info like the plain posv
    B = rng.standard_normal((96, 32, 2))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    for tiers in (0, 1, 2, None):
        p = router.init_params(0)
        if tiers is not None:  # a network that always answers `tiers`
            p[router.OFF_W3:router.OFF_B3] = 0
            p[router.OFF_B3:router.OFF_B3 + 3] = 0
            p[router.OFF_B3 + tiers] = 1.0
        r = router.Router(p)
        X = torch.zeros_like(dB)
        info = torch.full((96,), -9, dtype=torch.int32, device="cuda")
        stat = torch.full((96,), -9, dtype=torch.int32, device="cuda")
        tier = torch.full((96,), -9, dtype=torch.int32, device="cuda")
        r.posv(dA, dB, X, info=info, stat=stat, tier=tier)
        torch.cuda.synchronize()
        info, stat, tier = info.cpu().numpy(), stat.cpu().numpy(), tier.cpu().numpy()
        assert set(np.unique(tier)) <= {0, 1, 2}
        if tiers is not None:
            assert (tier == tiers).all()
        if tiers == 2:
            assert (stat == -1).all()
        assert info[10] == 1 and (np.delete(info, 10) == 0).all()
        ok = info == 0
        Xh = X.cpu().numpy()
        assert _bwd(tsdx, A[ok], Xh[ok], B[ok]).max() <= 1e-14, tiers


def _best_ms(fn, reps=3):
    """Best of `reps` HIP-event timings of fn() (after one untimed call)."""
    import torch
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def test_router_gain_measured(tsdx):
    """What the shipped router gains, timed on the real calls (features and
    network included), on fresh matrices built like the routed-posv workload
    (kappa log-uniform [1, 1e12]; 16384 per size).  Policies per size:
      routed      Router.posv as shipped (its per-size policy: the network,
                  or one tier for every matrix where timing found that
                  faster; shipped: float64 at n <= 32, fp32 + refinement at 64),
      network     Router.posv with the network at every size,
      fp32        fp32 factor + refinement for every matrix, float64 fallback,
      fp64        the float64 tier for every matrix,
      rule        float64 for the float64-direct sizes, fp32 + refinement with
                  fallback otherwise, without the network.
    Asserted: every routed answer meets 1e-14, and over the mix of sizes the
    routed policy beats both fixed tiers.  Measured, not asserted as a gain:
    at n = 64 the network against fp32 for all (it does not win there: the
    cheap features cannot see the smallest eigenvalue of these matrices, so
    it both misses refinement failures and sends convergent matrices to
    float64); measured 5-15 % slower at this batch, it must stay within 30 %
    (a guard against a regression of the network's own cost)."""
    import torch
    from tinysolverdx import _lib, device, matrices, router
    r = router.default()
    shipped = (r.direct, r.fp32_direct)
    cnt = 16384
    times = {k: {} for k in ("routed", "network", "fp32", "fp64", "rule")}
    for n in (16, 32, 64):
        A, _ = matrices.spd_device(n, cnt, 1.0, 1e12, 9100 + n)
        B = torch.randn((cnt, n, 1), generator=torch.Generator(device="cuda").manual_seed(n), device="cuda",
                        dtype=torch.float64)
        X = torch.empty_like(B)
        wi = device.posv_ir_workspace(cnt)
        wr = torch.empty(_lib.lib().tsdx_hip_posv_routed_workspace_size(cnt), dtype=torch.uint8, device="cuda")
        fp64 = "fp64_block" if n > 32 else "fp64"
        r.direct, r.fp32_direct = shipped
        times["routed"][n] = _best_ms(lambda: r.posv(A, B, X, work=wr))
        Xr = X.clone()
        r.direct, r.fp32_direct = frozenset(), frozenset()
        times["network"][n] = _best_ms(lambda: r.posv(A, B, X, work=wr))
        r.direct, r.fp32_direct = shipped
        times["fp32"][n] = _best_ms(lambda: device.posv_ir(A, B, X, factor="fp32", work=wi))
        times["fp64"][n] = _best_ms(lambda: device.posv_ir(A, B, X, factor=fp64, work=wi))
        times["rule"][n] = times["fp64"][n] if router.bucket(n) in shipped[0] else times["fp32"][n]
        sel = torch.arange(0, cnt, 64, device="cuda")  # the bar, on a sample (the CPU metric is float64 numpy)
        assert _bwd(tsdx, A[sel].cpu().numpy(), Xr[sel].cpu().numpy(), B[sel].cpu().numpy()).max() <= 1e-14, n
    total = {k: sum(v.values()) for k, v in times.items()}
    print("router gain (ms per 16384 matrices):", {k: {n: round(t, 4) for n, t in v.items()} for k, v in times.items()})
    assert total["routed"] < 0.95 * min(total["fp32"], total["fp64"]), total
    assert times["network"][64] <= 1.3 * times["fp32"][64], times


def test_fused_features_match_standalone_kernel():
    """The routed kernel computes the features from the float64 blocks it
    holds; they agree with the standalone features kernel and the oracle."""
    import torch
    import mlp_oracle as M
    from tinysolverdx import router
    rng = np.random.default_rng(11)
    for n in (5, 16, 40, 64):
        A, _ = _spd(rng, 24, n, 0, 10)
        A[2] = -A[2]
        A[4][np.triu_indices(n, 1)] = 1e20
        dA = torch.from_numpy(A).cuda()
        B = torch.from_numpy(rng.standard_normal((24, n, 1))).cuda()
        feat = torch.zeros((24, 12), dtype=torch.float32, device="cuda")
        tier = torch.full((24,), -1, dtype=torch.int32, device="cuda")
        r = router.default()
        r.direct, r.fp32_direct = frozenset(), frozenset()
        r.posv(dA, B, torch.zeros_like(B), feat=feat, tier=tier)
        got = feat.cpu().numpy()
        # the routed kernel's choice is the inference path's on its features
        assert np.array_equal(tier.cpu().numpy(), r.predict(feat).cpu().numpy()), n
        assert np.allclose(got, router.features(dA).cpu().numpy(), rtol=1e-4, atol=2e-4), n
        assert np.allclose(got, M.features(A), rtol=1e-4, atol=2e-4), n


@pytest.mark.parametrize("n", [64, 50])
def test_direct_bucket_takes_the_float64_tier(tsdx, n):
    """A size bucket marked direct skips the network: every matrix goes to
    the float64 tier (tier 2, stat -1) and still meets the bar; info as the
    plain posv for a matrix that is not SPD."""
    import torch
    from tinysolverdx import router
    rng = np.random.default_rng(12 + n)
    A, _ = _spd(rng, 40, n, 0, 12)
    A[5] = -np.eye(n)
    B = rng.standard_normal((40, n, 1))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    r = router.Router(router.init_params(0), direct=[64])
    X = torch.zeros_like(dB)
    info = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    stat = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    tier = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    r.posv(dA, dB, X, info=info, stat=stat, tier=tier)
    torch.cuda.synchronize()
    info, stat, tier = info.cpu().numpy(), stat.cpu().numpy(), tier.cpu().numpy()
    assert (tier == 2).all() and (stat == -1).all()
    assert info[5] == 1 and (np.delete(info, 5) == 0).all()
    ok = info == 0
    assert _bwd(tsdx, A[ok], X.cpu().numpy()[ok], B[ok]).max() <= 1e-14


@pytest.mark.parametrize("n", [64, 40])
def test_fp32_direct_bucket_takes_the_refinement_tier(tsdx, n):
    """A size bucket marked fp32 direct skips the network: every matrix takes
    the fp32 refinement tier (tier 1), whose float64 fallback still solves
    what refinement cannot (stat -1, among them a matrix that is not SPD,
    with the plain posv's info); every other answer meets the bar and equals
    tsdx_hip_posv_ir_f64's with the fp32 factor bit for bit."""
    import torch
    from tinysolverdx import device, router
    rng = np.random.default_rng(31 + n)
    A, _ = _spd(rng, 40, n, 0, 12)
    A[5] = -np.eye(n)
    B = rng.standard_normal((40, n, 1))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    r = router.Router(router.init_params(0), fp32_direct=[64])
    X = torch.zeros_like(dB)
    info = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    stat = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    tier = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    r.posv(dA, dB, X, info=info, stat=stat, tier=tier)
    Xw = torch.zeros_like(dB)
    info_w = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    stat_w = torch.full((40,), -9, dtype=torch.int32, device="cuda")
    device.posv_ir(dA, dB, Xw, factor="fp32", info=info_w, stat=stat_w)
    torch.cuda.synchronize()
    assert (tier.cpu().numpy() == 1).all()
    assert np.array_equal(info.cpu().numpy(), info_w.cpu().numpy())
    assert np.array_equal(stat.cpu().numpy(), stat_w.cpu().numpy())
    assert np.array_equal(X.cpu().numpy().view(np.uint64), Xw.cpu().numpy().view(np.uint64))
    info = info.cpu().numpy()
    assert info[5] == 1 and stat.cpu().numpy()[5] == -1
    ok = info == 0
    assert _bwd(tsdx, A[ok], X.cpu().numpy()[ok], B[ok]).max() <= 1e-14
