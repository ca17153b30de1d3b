"""Learned precision router for posv, on device-resident torch tensors.

A network of 12 -> 64 -> 64 -> 3 (ReLU, float32 weights) predicts, for each
SPD matrix, the cheapest posv tier expected to reach the tolerance:
    0  bf16 factor + float64 refinement
    1  fp32 factor + float64 refinement
    2  float64 direct solve
from 12 features taken in one pass over the lower triangle of A.  Features,
forward pass, gradient and Adam step are the library's HIP kernels
(include/tsdx_hip.h, tsdx_hip_router_*); torch only holds the buffers and
shuffles minibatches.  Labels come from the pipeline itself (make_labels):
every tier runs on the training matrices, and a matrix's label is the tier
with the smallest cost for it, costs measured per size at throughput
(measure_costs; a refinement that does not converge pays for the float64
solve as well).

The routed solve (Router.posv) never trusts the network for accuracy: every
answer is judged by the float64 residual, refinement that stalls falls back
to the float64 solve, exactly as tsdx_hip_posv_ir_f64.

A router also carries the size buckets (4/8/16/32/64) it solves with one
tier for every matrix, without consulting the network: Router.direct (the
float64 tier) and Router.fp32_direct (fp32 factor + refinement, float64
fallback).  After training, choose_direct times, per size, the routed solve
against both fixed tiers on separately seeded matrices of the training
distribution and keeps the fastest policy: a bucket where a fixed tier beats
the network (its features and its wrong guesses cost more than its right
ones save) is marked direct for that tier.  The choice is stored next to the
parameters (router_default.json).
"""
import numpy as np

from . import _lib
from .device import Call, _stream

FEATURES, HIDDEN, TIERS = 12, 64, 3
PARAMS = 5211
OFF_W1 = 24
OFF_B1 = OFF_W1 + HIDDEN * FEATURES
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + TIERS * HIDDEN
TIER_NAMES = ("bf16+IR", "fp32+IR", "fp64")


def features(A):
    """(batch, 12) float32 features of float64 A (batch, n, n), on A's device."""
    import torch
    c = Call(A, dtypes=(torch.float64,))
    feat = torch.empty((c.batch, FEATURES), dtype=torch.float32, device=A.device)
    _lib.call("tsdx_hip_router_features_f64", c.mat(), c.n, c.batch, feat.data_ptr(), c.stream())
    return feat


def _timed(fn, reps=3):
    """Smallest HIP-event time (seconds) of fn over reps runs on torch's stream."""
    import torch
    best = float("inf")
    s = torch.cuda.current_stream()
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    return best


def measure_costs(n, count=32768, max_iter=10, tol=1e-14, seed=1):
    """Time per matrix (seconds) of each tier at size n, at throughput: the
    refinement tiers on well-conditioned matrices (kappa up to 1e3, so they
    converge and the float64 fallback stays empty), the float64 tier with the
    solve that tier-2 matrices and refinement failures actually take
    (posv_ir factor "fp64_block").  Returns {0: bf16, 1: fp32, 2: fp64}."""
    import torch
    from . import device, matrices
    A, _ = matrices.spd_device(n, count, 1.0, 1e3, seed)
    B = torch.randn((count, n, 1), generator=torch.Generator(device="cuda").manual_seed(seed), device="cuda",
                    dtype=torch.float64)
    X = torch.empty_like(B)
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    work = device.posv_ir_workspace(count)
    cost = {t: _timed(lambda: device.posv_ir(A, B, X, factor=f, max_iter=max_iter, tol=tol, work=work)) / count
            for t, f in ((0, "bf16"), (1, "fp32"))}
    cost[2] = _timed(lambda: device.posv_ir(A, B, X, factor="fp64_block", max_iter=max_iter, tol=tol,
                                            info=info, work=work)) / count
    return cost


def make_labels(A, B, cost, max_iter=10, tol=1e-14):
    """The cheapest tier for each matrix of one size under the per-matrix tier
    costs `cost` ({tier: seconds}, from measure_costs at that size): a
    refinement tier costs cost[t] when it converges on the matrix (its stat)
    and cost[t] + cost[2] when the float64 fallback has to follow; ties go to
    the lower precision."""
    import torch
    from . import device
    batch = A.shape[0]
    X = torch.empty_like(B)
    work = device.posv_ir_workspace(batch, A.device)
    exp = []
    for t, f in ((0, "bf16"), (1, "fp32")):
        s = torch.empty(batch, dtype=torch.int32, device=A.device)
        device.posv_ir(A, B, X, factor=f, max_iter=max_iter, tol=tol, stat=s, work=work)
        exp.append(torch.where(s >= 0, torch.tensor(cost[t], dtype=torch.float64, device=A.device),
                               torch.tensor(cost[t] + cost[2], dtype=torch.float64, device=A.device)))
    exp.append(torch.full((batch,), cost[2], device=A.device, dtype=torch.float64))
    return torch.stack(exp, dim=1).argmin(dim=1).to(torch.int32)  # first minimum: lower precision on ties


DEFAULT_PARAMS = __import__("pathlib").Path(__file__).resolve().parent / "router_default.npy"
BUCKETS = (4, 8, 16, 32, 64)


def bucket(n):
    """Size bucket of n (the smallest of 4, 8, 16, 32, 64 that holds it; 64
    for larger n, which the library then rejects)."""
    return next((b for b in BUCKETS if n <= b), BUCKETS[-1])


def choose_direct(r, sizes=(16, 32, 64), count=65536, kmax=1e12, seed=20261210, reps=3):
    """Per size, the fastest of three policies for Router.posv, timed on
    `count` SPD matrices with kappa log-uniform on [1, kmax] (the training
    distribution, fresh seed): the network, the float64 tier for every matrix,
    the fp32 refinement tier for every matrix (float64 fallback included).
    Returns (float64 buckets, fp32 buckets, {n: {"routed": s, "fp64": s,
    "fp32": s}} per matrix)."""
    import torch
    from . import matrices
    fp64, fp32, times = [], [], {}
    saved = (r.direct, r.fp32_direct)
    try:
        for n in sizes:
            A, _ = matrices.spd_device(n, count, 1.0, kmax, seed + n)
            B = torch.randn((count, n, 1), generator=torch.Generator(device="cuda").manual_seed(n), device="cuda",
                            dtype=torch.float64)
            X = torch.empty_like(B)
            wr = torch.empty(_lib.lib().tsdx_hip_posv_routed_workspace_size(count), dtype=torch.uint8, device="cuda")
            t = {}
            for name, d64, d32 in (("routed", (), ()), ("fp64", (bucket(n),), ()), ("fp32", (), (bucket(n),))):
                r.direct, r.fp32_direct = frozenset(d64), frozenset(d32)
                t[name] = _timed(lambda: r.posv(A, B, X, work=wr), reps) / count
            times[n] = t
            best = min(("routed", "fp64", "fp32"), key=lambda k: t[k])
            if best == "fp64":
                fp64.append(bucket(n))
            elif best == "fp32":
                fp32.append(bucket(n))
            del A, B, X
    finally:
        r.direct, r.fp32_direct = saved
    return sorted(fp64), sorted(fp32), times


# Training mix of the shipped router: per size, (count, kappa_min, kappa_max)
# blocks of SPD matrices built like the posv workloads.  Every size covers
# kappa up to 1e12 (the routed-posv workload); n = 64 adds the kappa range of
# the n = 64 mixed-precision posv workload (1e1..1e6), so the prior the network
# learns matches the mix it is used on there.
TRAINING_MIX = {16: [(20000, 1.0, 1e12)], 32: [(20000, 1.0, 1e12)], 64: [(20000, 1.0, 1e12), (20000, 1e1, 1e6)]}


def training_set(mix=None, seed=20261207, device="cuda"):
    """Features and labels of SPD matrices built like the posv workloads
    (matrices.spd_device: Q diag(lambda) Q^T with Q Haar, lambda_0 = 1,
    lambda_{n-1} = 1/kappa and the other eigenvalues log-uniform in between,
    kappa log-uniform per block), for the blocks of `mix` ({n: [(count,
    kappa_min, kappa_max), ...]}, default TRAINING_MIX), seeded from `seed`."""
    import torch
    from . import matrices
    mix = TRAINING_MIX if mix is None else mix
    feats, labels = [], []
    costs = {}
    for n, blocks in mix.items():
        costs[n] = measure_costs(n)
        for b, (count, kmin, kmax) in enumerate(blocks):
            s = seed + 101 * n + b
            A, _ = matrices.spd_device(n, count, kmin, kmax, s, device=device)
            B = torch.randn((count, n, 1), generator=torch.Generator(device=device).manual_seed(s + 1),
                            device=device, dtype=torch.float64)
            labels.append(make_labels(A, B, costs[n]))
            feats.append(features(A))
            del A, B
    training_set.costs = costs
    return torch.cat(feats), torch.cat(labels)


def train_default(path=DEFAULT_PARAMS, mix=None, epochs=60, seed=20261207):
    """Trains the router shipped with the package (TRAINING_MIX: kappa up to
    1e12 for n in {16, 32, 64}, plus the 1e1..1e6 range at n = 64) with the
    library's kernels and writes its parameters."""
    feat, label = training_set(mix, seed=seed)
    r = Router(seed=seed)
    hist = r.fit(feat, label, epochs=epochs, batch=4096, lr=3e-3, seed=seed)
    acc = float((r.predict(feat) == label).float().mean().item())
    shift = r.calibrate(training_set.costs[max(training_set.costs)])  # the largest size: where the network runs
    counts = np.bincount(label.cpu().numpy(), minlength=3).tolist()
    costs = {n: {t: round(c * 1e9, 3) for t, c in v.items()} for n, v in training_set.costs.items()}
    report = {"loss_first": hist[0], "loss_last": hist[-1], "label_counts": counts, "train_accuracy": acc,
              "fp64_logit_shift": round(shift, 4), "cost_ns_per_matrix": costs}
    report.update(direct_policy(r))
    r.save(path, report)
    return report


def direct_policy(r):
    """Runs choose_direct, sets r.direct / r.fp32_direct and returns the
    report entries of the choice."""
    d64, d32, times = choose_direct(r)
    r.direct, r.fp32_direct = frozenset(d64), frozenset(d32)
    return {"direct_buckets": d64, "fp32_direct_buckets": d32,
            "policy_ns_per_matrix": {n: {k: round(v * 1e9, 3) for k, v in t.items()} for n, t in times.items()}}


def default():
    """The router trained by train_default() and shipped with the package."""
    if not DEFAULT_PARAMS.exists():
        raise RuntimeError(f"{DEFAULT_PARAMS} is missing: run python -m tinysolverdx.router --train")
    return Router.load(DEFAULT_PARAMS)


def init_params(seed=0):
    """He-initialised weights, zero biases, identity feature scaling (numpy)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[12:24] = 1.0
    p[OFF_W1:OFF_B1] = rng.standard_normal(HIDDEN * FEATURES) * np.sqrt(2.0 / FEATURES)
    p[OFF_W2:OFF_B2] = rng.standard_normal(HIDDEN * HIDDEN) * np.sqrt(2.0 / HIDDEN)
    p[OFF_W3:OFF_B3] = rng.standard_normal(TIERS * HIDDEN) * np.sqrt(2.0 / HIDDEN)
    return p


class Router:
    """Parameters live in a float32 device tensor of PARAMS entries."""

    def __init__(self, params=None, device="cuda", seed=0, direct=(), fp32_direct=()):
        import torch
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (PARAMS,):
            raise ValueError(f"router parameters must have {PARAMS} entries")
        self.params = torch.from_numpy(p.copy()).to(device)
        bad = (set(direct) | set(fp32_direct)) - set(BUCKETS)
        if bad:
            raise ValueError(f"direct buckets must be among {BUCKETS}, got {sorted(bad)}")
        if set(direct) & set(fp32_direct):
            raise ValueError("a bucket is either float64 direct or fp32 direct")
        self.direct = frozenset(direct)  # size buckets solved in float64 without the network
        self.fp32_direct = frozenset(fp32_direct)  # size buckets solved by fp32 + refinement without it
        self.policy_times = {}  # {size: {"routed", "fp64", "fp32": ns per matrix}} behind the choice, as loaded

    # ---------------------------------------------------------------- inference
    def forward(self, feat):
        import torch
        c = self._features(feat)
        logits = torch.empty((c.batch, TIERS), dtype=torch.float32, device=feat.device)
        if c.batch:
            _lib.call("tsdx_hip_router_forward", self._params(c), c.mat(), c.batch, logits.data_ptr(), c.stream())
        return logits

    def _features(self, feat):
        """Checks a (count, 12) float32 feature tensor on the parameters' device."""
        import torch
        c = Call(feat, "feat", dtypes=(torch.float32,), square=False)
        if feat.dim() != 2 or feat.shape[1] != FEATURES:
            raise ValueError(f"feat must have shape (count, {FEATURES}), got {tuple(feat.shape)}")
        return c

    def _params(self, c):
        import torch
        return c.tensor(self.params, "params", torch.float32, (PARAMS,))

    def infer(self, feat, work=None):
        """(tiers, logits) from the network as the routed posv runs it: on the
        matrix cores with bfloat16 weights and activations
        (tsdx_hip_router_infer), so the tiers are the routed solve's choices
        for these features; logits within bfloat16 rounding of forward()."""
        import torch
        c = self._features(feat)
        tier = torch.empty(c.batch, dtype=torch.int32, device=feat.device)
        logits = torch.empty((c.batch, TIERS), dtype=torch.float32, device=feat.device)
        if c.batch:
            need = _lib.lib().tsdx_hip_router_infer_workspace_size()
            if work is None:
                work = torch.empty(need, dtype=torch.uint8, device=feat.device)
            w, wn = c.work(work, need)
            _lib.call("tsdx_hip_router_infer", self._params(c), c.mat(), c.batch, tier.data_ptr(), logits.data_ptr(),
                      w, wn, c.stream())
        return tier, logits

    def predict(self, feat):
        """Tiers the routed posv picks for these features."""
        return self.infer(feat)[0]

    # ----------------------------------------------------------------- training
    def grad(self, feat, label, work=None):
        """(gradient, mean cross-entropy) over the batch."""
        import torch
        c = self._features(feat)
        need = _lib.lib().tsdx_hip_router_grad_workspace_size(c.batch)
        if work is None:
            work = torch.empty(need, dtype=torch.uint8, device=feat.device)
        w, wn = c.work(work, need)
        g = torch.empty(PARAMS, dtype=torch.float32, device=feat.device)
        loss = torch.empty(1, dtype=torch.float32, device=feat.device)
        _lib.call("tsdx_hip_router_grad", self._params(c), c.mat(), c.ints(label, "label", optional=False), c.batch,
                  g.data_ptr(), loss.data_ptr(), w, wn, c.stream())
        return g, loss

    def calibrate(self, cost):
        """Expected-cost decision between the fp32 and float64 tiers.  The
        network estimates p = P(fp32 refinement fails | features); under the
        label cost model (a failed refinement pays c1 and then c2) choosing
        float64 is cheaper iff p > theta = (c2 - c1) / c2, while the argmax of
        the logits decides at p > 1/2.  Shifting the float64 logit by
        -log(theta / (1 - theta)) makes the argmax take the expected-cost
        decision.  cost: {tier: cost} at the size the network is used for.
        Returns the shift applied to the float64 logit's bias."""
        c1, c2 = float(cost[1]), float(cost[2])
        theta = min(max((c2 - c1) / c2, 0.05), 0.95) if c2 > 0 else 0.5
        shift = -float(np.log(theta / (1.0 - theta)))
        self.params[OFF_B3 + 2] += shift
        return shift

    def set_scaling(self, feat):
        """Feature means and scales from a training set."""
        mu = feat.mean(dim=0)
        sd = feat.std(dim=0).clamp_min(1e-3)
        self.params[0:12] = mu
        self.params[12:24] = sd

    def fit(self, feat, label, epochs=40, batch=4096, lr=3e-3, seed=0):
        """Adam on minibatches; returns the mean loss of every epoch."""
        import torch
        dev = feat.device
        self.set_scaling(feat)
        m = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        v = torch.zeros_like(m)
        count = feat.shape[0]
        bsz = min(batch, count)
        work = torch.empty(_lib.lib().tsdx_hip_router_grad_workspace_size(bsz), dtype=torch.uint8, device=dev)
        gen = torch.Generator(device=dev).manual_seed(seed)
        step, history = 0, []
        for _ in range(epochs):
            perm = torch.randperm(count, generator=gen, device=dev)
            losses = []
            for s in range(0, count - bsz + 1, bsz):
                idx = perm[s:s + bsz]
                f, l = feat[idx].contiguous(), label[idx].contiguous()
                g, loss = self.grad(f, l, work)
                step += 1
                _lib.call("tsdx_hip_router_adam", self.params.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(),
                          step, float(lr), _stream(f))
                losses.append(loss)
            history.append(float(torch.cat(losses).mean().item()))
        return history

    def save(self, path, report=None):
        """Parameters to `path` (.npy); the direct buckets (and the training
        report, if given) to the .json beside it."""
        import json
        import pathlib
        np.save(path, self.params.cpu().numpy())
        meta = {"direct_buckets": sorted(self.direct), "fp32_direct_buckets": sorted(self.fp32_direct)}
        if report is not None:
            meta["training"] = report
        pathlib.Path(path).with_suffix(".json").write_text(json.dumps(meta, indent=1, default=str) + "\n")

    @classmethod
    def load(cls, path, device="cuda"):
        import json
        import pathlib
        meta = pathlib.Path(path).with_suffix(".json")
        m = json.loads(meta.read_text()) if meta.exists() else {}
        r = cls(np.load(path, allow_pickle=False), device=device, direct=m.get("direct_buckets", []),
                fp32_direct=m.get("fp32_direct_buckets", []))
        r.policy_times = m.get("training", {}).get("policy_ns_per_matrix", {})
        return r

    # -------------------------------------------------------------- routed posv
    def posv(self, A, B, X, max_iter=10, tol=1e-14, info=None, stat=None, tier=None, feat=None, work=None):
        """Routed posv (float64): X := A^-1 B; stat / tier / feat as documented
        in include/tsdx_hip.h (tsdx_hip_posv_routed_f64).  For a size in
        self.direct every matrix takes the float64 tier without the network
        (tier 2, stat -1, feat left untouched); for a size in self.fp32_direct
        every matrix takes the fp32 refinement tier without it (tier 1, stat
        as tsdx_hip_posv_ir_f64: corrections used, -1 where the float64
        fallback solved the matrix; feat left untouched)."""
        import torch
        c = Call(A, dtypes=(torch.float64,))
        batch, n = c.batch, c.n
        b, nrhs = c.rhs(B)
        x = c.like(X, B)
        pinfo, pstat, ptier = c.ints(info, "info"), c.ints(stat, "stat"), c.ints(tier, "tier")
        pfeat = c.tensor(feat, "feat", torch.float32, (batch, FEATURES), optional=True)
        if bucket(n) in self.direct:
            from . import device
            if n <= 32:  # the float64 tier there is the plain float64 posv: one launch, no list
                device.posv_solve(A, B, X, info)
                if stat is not None:
                    stat.fill_(-1)
            else:
                wi = work if work is not None and work.numel() >= _lib.lib().tsdx_hip_posv_ir_workspace_size(batch) \
                    else None
                device.posv_ir(A, B, X, factor="fp64_block", max_iter=max_iter, tol=tol, info=info, stat=stat,
                               work=wi)
            if tier is not None:
                tier.fill_(2)
            return X, info, stat, tier
        if bucket(n) in self.fp32_direct:
            from . import device
            wi = work if work is not None and work.numel() >= _lib.lib().tsdx_hip_posv_ir_workspace_size(batch) \
                else None
            device.posv_ir(A, B, X, factor="fp32", max_iter=max_iter, tol=tol, info=info, stat=stat, work=wi)
            if tier is not None:
                tier.fill_(1)
            return X, info, stat, tier
        need = _lib.lib().tsdx_hip_posv_routed_workspace_size(batch)
        if work is None:
            work = torch.empty(need, dtype=torch.uint8, device=A.device)
        w, wn = c.work(work, need)
        _lib.call("tsdx_hip_posv_routed_f64", c.mat(), n, batch, b, x, nrhs, self._params(c), int(max_iter),
                  float(tol), pinfo, pstat, ptier, pfeat, w, wn, c.stream())
        return X, info, stat, tier


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the default precision router on the GPU")
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--choose-direct", action="store_true",
                    help="re-time the per-size policies of the shipped router (parameters unchanged)")
    ap.add_argument("--out", default=str(DEFAULT_PARAMS))
    ap.add_argument("--epochs", type=int, default=60)
    args = ap.parse_args()
    if args.train:
        print(json.dumps(train_default(args.out, epochs=args.epochs)), flush=True)
    elif args.choose_direct:
        import pathlib
        r = default()
        report = json.loads(DEFAULT_PARAMS.with_suffix(".json").read_text()).get("training", {})
        for k in ("direct_buckets", "fp32_direct_buckets", "routed_vs_fp64_ns_per_matrix", "policy_ns_per_matrix"):
            report.pop(k, None)
        report.update(direct_policy(r))
        r.save(pathlib.Path(args.out), report)
        print(json.dumps(report), flush=True)
