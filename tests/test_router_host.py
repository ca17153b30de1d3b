"""Host logic of the learned router (no GPU): size buckets, the direct
float64 buckets and their save/load beside the parameters."""
import numpy as np
import pytest


def test_size_buckets():
    from tinysolverdx import router
    assert [router.bucket(n) for n in (1, 4, 5, 8, 9, 16, 17, 32, 33, 48, 64)] == [4, 4, 8, 8, 16, 16, 32, 32, 64, 64, 64]


def test_direct_buckets_round_trip(tmp_path):
    from tinysolverdx import router
    r = router.Router(router.init_params(3), device="cpu", direct=[32, 16])
    assert r.direct == {16, 32}
    path = tmp_path / "r.npy"
    r.save(path, {"note": "test"})
    q = router.Router.load(path, device="cpu")
    assert q.direct == {16, 32}
    assert np.array_equal(q.params.numpy(), r.params.numpy())
    (tmp_path / "r.json").unlink()  # parameters alone: the network for every size
    assert router.Router.load(path, device="cpu").direct == frozenset()


def test_direct_buckets_validated():
    from tinysolverdx import router
    with pytest.raises(ValueError):
        router.Router(router.init_params(0), device="cpu", direct=[24])
    with pytest.raises(ValueError):
        router.Router(router.init_params(0), device="cpu", fp32_direct=[48])
    with pytest.raises(ValueError):  # one policy per bucket
        router.Router(router.init_params(0), device="cpu", direct=[64], fp32_direct=[64])


def test_fp32_direct_buckets_round_trip(tmp_path):
    from tinysolverdx import router
    r = router.Router(router.init_params(4), device="cpu", direct=[16], fp32_direct=[64, 32])
    path = tmp_path / "r.npy"
    r.save(path, {"policy_ns_per_matrix": {"64": {"routed": 2.0, "fp64": 3.0, "fp32": 1.0}}})
    q = router.Router.load(path, device="cpu")
    assert q.direct == {16} and q.fp32_direct == {32, 64}
    assert q.policy_times["64"]["fp32"] == 1.0


def test_shipped_router_metadata():
    """The shipped router's direct buckets came from timing on the GPU; they
    are a subset of the buckets and stored with the training report."""
    import json
    from tinysolverdx import router
    meta = json.loads(router.DEFAULT_PARAMS.with_suffix(".json").read_text())
    assert set(meta["direct_buckets"]) <= set(router.BUCKETS)
    assert set(meta.get("fp32_direct_buckets", [])) <= set(router.BUCKETS)
    assert not set(meta["direct_buckets"]) & set(meta.get("fp32_direct_buckets", []))
    assert "cost_ns_per_matrix" in meta["training"]
    # every direct bucket is the fastest of the three timed policies for its size
    for n, t in meta["training"].get("policy_ns_per_matrix", {}).items():
        best = min(t, key=t.get)
        b = router.bucket(int(n))
        assert (b in meta["direct_buckets"]) == (best == "fp64"), (n, t)
        assert (b in meta.get("fp32_direct_buckets", [])) == (best == "fp32"), (n, t)


def test_calibrate_shifts_the_float64_logit():
    """theta = (c2 - c1) / c2: with float64 three times the fp32 cost the
    float64 tier must win at p > 2/3, i.e. its logit moves by -log 2."""
    from tinysolverdx import router
    r = router.Router(router.init_params(1), device="cpu")
    before = r.params.numpy().copy()
    shift = r.calibrate({0: 9.0, 1: 1.0, 2: 3.0})
    after = r.params.numpy()
    assert abs(shift + np.log(2.0)) < 1e-6
    assert abs(after[router.OFF_B3 + 2] - before[router.OFF_B3 + 2] - shift) < 1e-6
    diff = np.flatnonzero(after != before)
    assert list(diff) == [router.OFF_B3 + 2]
