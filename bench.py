#!/usr/bin/env python3
"""Throughput benchmark of the batched posv path on one MI355X.

Workload (the headline metric): posv at n = 64 on 131072 SPD float64 matrices,
A = Q diag(lambda) Q^T with Q Haar-orthogonal, lambda_0 = 1, lambda_63 = 1/kappa
and the other eigenvalues log-uniform in between, kappa log-uniform on
[1e1, 1e6] per matrix; B ~ N(0, 1), nrhs = 1 (--nrhs changes it).  Inputs are
generated on the device and stay resident in HBM; one step = one call over
the whole batch.  Every solution is judged by its float64 normwise backward
error ||b - Ax|| / (||A||_F ||x|| + ||b||), which must be <= 1e-14.

Printed on stdout: one JSON line with the metric, the roofline of the dominant
kernel (HIP events on the launch stream) and the reference's own CPU posv
(compiled from its sources, oracle/_ref) timed on this host's cores.  Progress
goes to stderr.

    python bench.py [--steps K] [--warmup W] [--nrhs R] [--batch N] [--factor F]

The other workloads of the path (one JSON line each, same fields):
    python bench.py --workload potrf                       fp32 n=32 Cholesky, 65536 matrices
    python bench.py --workload getrf --n 64 --dtype f64    LU + gesv(4 RHS), ~1 GiB of A
    python bench.py --workload ragged --dtype f32          1,000,000 mixed sizes and kinds
    python bench.py --workload router                      2^20 matrices, kappa up to 1e12, routed vs not
"""
import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "tinysolverdx-batched-small-matrix-cholesky-lu-with-numerical_amd"
sys.path.insert(0, str(PKG_DIR))
sys.path.insert(0, str(ROOT / "oracle"))

HBM_PEAK_GBS = 8000.0
METRIC = "posv solves/s n=64: routed mixed+IR vs plain FP64 GPU at bwd err≤1e-14; HBM GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_spd(n, batch, kmin, kmax, seed, chunk=16384):
    """SPD float64 matrices on the device (tinysolverdx.matrices.spd_device):
    A = Q diag(lambda) Q^T symmetrised, Q Haar, lambda_0 = 1, lambda_{n-1} =
    1/kappa, the rest log-uniform in between, kappa log-uniform on [kmin, kmax]."""
    from tinysolverdx import matrices
    return matrices.spd_device(n, batch, kmin, kmax, seed, chunk=chunk)


def backward_error(A, X, B):
    """float64 normwise backward error per matrix (max over right-hand sides)."""
    import torch
    R = B - torch.matmul(A, X)
    num = torch.linalg.vector_norm(R, dim=1)
    den = torch.linalg.matrix_norm(A, "fro")[:, None] * torch.linalg.vector_norm(X, dim=1) + \
        torch.linalg.vector_norm(B, dim=1)
    return (num / den).amax(dim=1)


def kappa_decades(kappa, X64, tiers):
    """Per decade of kappa: the matrix count and, for every solve in tiers
    (name -> (X, A, B, stat or None, tier or None)), the largest forward error
    against the plain float64 solve X64 (max over right-hand sides of the
    relative 2-norm difference), the largest backward error, the share of
    refinement fallbacks (stat -1 where the routed choice was not float64)
    and, given tier, the share the network sent to float64."""
    import torch
    dec = torch.floor(torch.log10(kappa.to(torch.float64))).to(torch.int64)
    ref = torch.linalg.vector_norm(X64, dim=1)
    out = {}
    stats = {}
    for name, (X, A, B, st, tr) in tiers.items():
        fw = (torch.linalg.vector_norm(X - X64, dim=1) / ref).amax(dim=1)
        bw = backward_error(A, X, B)
        fb = None
        if st is not None:
            fb = st < 0
            if tr is not None:
                fb = fb & (tr != 2)
        stats[name] = (fw, bw, fb, tr)
    for d in range(int(dec.min().item()), int(dec.max().item()) + 1):
        sel = dec == d
        cnt = int(sel.sum().item())
        if cnt == 0:
            continue
        row = {"matrices": cnt}
        for name, (fw, bw, fb, tr) in stats.items():
            r = {"max_fwd_err_vs_fp64": float(f"{fw[sel].max().item():.3e}"),
                 "max_bwd_err": float(f"{bw[sel].max().item():.3e}")}
            if fb is not None:
                r["fallback_rate"] = round(int((fb & sel).sum().item()) / cnt, 5)
            if tr is not None:
                r["fp64_tier_share"] = round(int(((tr == 2) & sel).sum().item()) / cnt, 5)
            row[name] = r
        out[f"1e{d}-1e{d + 1}"] = row
    return out


def time_steps(fn, steps, warmup):
    """Warmup, then exactly `steps` calls bracketed by device synchronisation;
    returns (wall seconds, HIP-event seconds) for the K calls."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) * 1e-3


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bit_mismatches(got, want):
    """Number of matrices whose factor differs from `want` in any bit (every
    NaN counts as equal to every NaN: NaN payloads differ between hosts)."""
    import numpy as np
    g = np.ascontiguousarray(got).reshape(got.shape[0], -1)
    w = np.ascontiguousarray(want).reshape(want.shape[0], -1)
    iv = {4: np.uint32, 8: np.uint64}[g.dtype.itemsize]
    gb, wb = g.view(iv).copy(), w.view(iv).copy()
    gb[np.isnan(g)] = 0
    wb[np.isnan(w)] = 0
    return int(((gb != wb) | (np.isnan(g) != np.isnan(w))).any(axis=1).sum())


def cpu_baseline(A, B, nrhs):
    """The reference's posv (its Auto backend: SIMD Cholesky + scalar solve),
    compiled from its own sources, on the host cores; best of reps like the
    reference's bench.py:24-32.  Falls back to the oracle restatement ("port")
    when the compiled reference is not present."""
    import numpy as np
    from oracle import Oracle, available
    kind = "reference" if available("reference") else "port"
    orc = Oracle(kind)
    cores = len(os.sched_getaffinity(0))
    orc.set_num_threads(cores)
    Ah = A.cpu().numpy()
    Bh = B.cpu().numpy()
    best, total, reps = float("inf"), 0.0, 0
    while reps < 5 and total < 20.0:
        t0 = time.perf_counter()
        orc.posv(Ah, Bh)
        dt = time.perf_counter() - t0
        best, total, reps = min(best, dt), total + dt, reps + 1
    batch = Ah.shape[0]
    return {"value": batch / best, "unit": "solves/s", "cores": cores, "kind": kind, "cpu_model": cpu_model(),
            "sample": f"full workload ({batch} matrices, nrhs={nrhs}), best of {reps} reps"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--batch", type=int, default=131072)
    ap.add_argument("--factor", default="fp32", choices=["fp32", "bf16", "fp16", "fp64"],
                    help="low precision of the refinement tier")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device copy (reference kernel for scaling rocprof byte counters)")
    ap.add_argument("--workload", default="posv", choices=["posv", "potrf", "getrf", "ragged", "router"],
                    help="posv: the headline (default); potrf: fp32 n=32 Cholesky of 65536 matrices; getrf: LU "
                         "+ gesv with 4 right-hand sides on ~1 GiB of A (--n, --dtype); ragged: 1,000,000 "
                         "matrices of Zipf sizes 2..64, half Cholesky half LU, one graph launch (--dtype); router: "
                         "2^20 SPD matrices, n in {16,32,64}, kappa up to 1e12, router trained here, routed vs "
                         "un-routed posv")
    ap.add_argument("--n", type=int, default=32, help="matrix size of the getrf workload")
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"], help="type of the getrf / ragged workloads")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the batched solvers shard by "
              "splitting the batch, with no exchange between GPUs", file=sys.stderr)
        return 2
    run = {"posv": run_posv, "potrf": run_potrf, "getrf": run_getrf, "ragged": run_ragged,
           "router": run_router}[a.workload]
    out = run(a)
    print(json.dumps(out), flush=True)
    return 0


def run_posv(a):
    import numpy as np
    import torch
    import tinysolverdx  # noqa: F401  (binds to torch's HIP runtime)
    from tinysolverdx import device

    n, batch, nrhs = 64, a.batch, a.nrhs
    log(f"[bench] generating {batch} SPD fp64 n={n} matrices on {torch.cuda.get_device_name(0)}")
    t0 = time.perf_counter()
    A, kappa = make_spd(n, batch, 1e1, 1e6, seed=20261205)
    g = torch.Generator(device="cuda").manual_seed(7)
    B = torch.randn((batch, n, nrhs), generator=g, device="cuda", dtype=torch.float64)
    X = torch.empty_like(B)
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f}s")

    # Buffers of the three paths, all allocated before anything is timed; the
    # three timed phases run back to back and every answer is judged afterwards,
    # so that no path is timed after the (large) error computations and all
    # three see the device in the same state.
    Xm = torch.empty_like(B)
    info_m = torch.zeros(batch, dtype=torch.int32, device="cuda")
    stat = torch.zeros(batch, dtype=torch.int32, device="cuda")
    work = device.posv_ir_workspace(batch)
    from tinysolverdx import router as rt
    R = rt.default()  # the shipped router (trained on the GPU, tinysolverdx/router_default.npy)
    net = rt.Router(R.params.cpu().numpy())  # the same network at every size, for comparison
    net_used = rt.bucket(n) not in (R.direct | R.fp32_direct)
    Xn = torch.empty_like(B)
    stat_n = torch.zeros(batch, dtype=torch.int32, device="cuda")
    tier_n = torch.zeros(batch, dtype=torch.int32, device="cuda")
    Xr = torch.empty_like(B)
    info_r = torch.zeros(batch, dtype=torch.int32, device="cuda")
    stat_r = torch.zeros(batch, dtype=torch.int32, device="cuda")
    tier = torch.zeros(batch, dtype=torch.int32, device="cuda")
    work_r = torch.empty(tinysolverdx._lib.lib().tsdx_hip_posv_routed_workspace_size(batch), dtype=torch.uint8,
                         device="cuda")

    # ---- plain FP64 posv (factor + solve in one launch, A and B read, X written)
    def plain():
        device.posv_solve(A, B, X, info)

    # ---- mixed precision + float64 refinement, float64 fallback
    def mixed():
        device.posv_ir(A, B, Xm, factor=a.factor, max_iter=10, tol=1e-14, info=info_m, stat=stat, work=work)

    # ---- routed: the shipped router's policy for this size (the network
    # picking each matrix's tier inside the solve kernel, or one tier for every
    # matrix where the router's timing found that faster); float64 residual
    # and fallback as above
    def routed():
        R.posv(A, B, Xr, info=info_r, stat=stat_r, tier=tier, work=work_r)

    # ---- the network at this size whatever the policy (for comparison)
    def network():
        net.posv(A, B, Xn, stat=stat_n, tier=tier_n, work=work_r)

    for fn in (plain, mixed, routed, network):  # every path once before anything is timed (clocks, code objects)
        fn()
    wall, ev = time_steps(plain, a.steps, a.warmup)
    wall_m, ev_m = time_steps(mixed, a.steps, a.warmup)
    wall_r, ev_r = time_steps(routed, a.steps, a.warmup)
    wall_n, ev_n = time_steps(network, a.steps, a.warmup)

    bw = backward_error(A, X, B)
    fails = int((info != 0).sum().item())
    ok = int((bw <= 1e-14).sum().item())
    plain_rate = batch * a.steps / wall
    plain_kern = ev / a.steps
    log(f"[bench] plain fp64 posv: {plain_kern * 1e3:.3f} ms/launch, {plain_rate / 1e6:.2f} M solves/s, "
        f"info>0: {fails}, bwd err <= 1e-14: {ok}/{batch}, max {bw.max().item():.2e}")

    bw_m = backward_error(A, Xm, B)
    ok_m = int((bw_m <= 1e-14).sum().item())
    mixed_rate = batch * a.steps / wall_m
    mixed_kern = ev_m / a.steps
    st = stat.cpu().numpy()
    fallback = int((st < 0).sum())
    iters = np.bincount(st[st >= 0], minlength=11)[:11].tolist()
    log(f"[bench] mixed {a.factor}+IR posv: {mixed_kern * 1e3:.3f} ms/call, {mixed_rate / 1e6:.2f} M solves/s, "
        f"fp64 fallback {fallback}, corrections histogram {iters}, bwd err <= 1e-14: {ok_m}/{batch}, "
        f"max {bw_m.max().item():.2e}")
    if ok_m != batch:
        log("[bench] WARNING: not every mixed answer meets the 1e-14 bar")

    bw_r = backward_error(A, Xr, B)
    ok_r = int((bw_r <= 1e-14).sum().item())
    routed_rate = batch * a.steps / wall_r
    routed_kern = ev_r / a.steps
    st_r, tr = stat_r.cpu().numpy(), tier.cpu().numpy()
    tiers = np.bincount(tr, minlength=3)[:3].tolist()
    fallback_r = int(((st_r < 0) & (tr != 2)).sum())
    log(f"[bench] routed posv: {routed_kern * 1e3:.3f} ms/call, {routed_rate / 1e6:.2f} M solves/s, tiers "
        f"(bf16, fp32, fp64) {tiers}, refinement fallbacks {fallback_r}, bwd err <= 1e-14: {ok_r}/{batch}, "
        f"max {bw_r.max().item():.2e}")
    if ok_r != batch:
        log("[bench] WARNING: not every routed answer meets the 1e-14 bar")
    policy = "network" if net_used else ("fp32+IR for every matrix" if rt.bucket(n) in R.fp32_direct
                                          else "fp64 for every matrix")
    bw_n = backward_error(A, Xn, B)
    ok_n = int((bw_n <= 1e-14).sum().item())
    net_rate = batch * a.steps / wall_n
    tiers_n = np.bincount(tier_n.cpu().numpy(), minlength=3)[:3].tolist()
    log(f"[bench] shipped policy at n={n}: {policy}; the network at n={n}: {ev_n / a.steps * 1e3:.3f} ms/call, "
        f"{net_rate / 1e6:.2f} M solves/s, tiers {tiers_n}, bwd err <= 1e-14: {ok_n}/{batch}")

    # error against the float64 solve and failure rate per kappa decade
    decades = kappa_decades(kappa, X, {"fixed_tier": (Xm, A, B, stat, None),
                                       "routed": (Xr, A, B, stat_r, tier),
                                       "network": (Xn, A, B, stat_n, tier_n)})
    log(f"[bench] per kappa decade: {json.dumps(decades)}")

    # algorithmic bytes per matrix: read A (8n^2), read B and write X (8n*nrhs each), info (4)
    bytes_per = 8 * n * n + 2 * 8 * n * nrhs + 4
    achieved = bytes_per * batch / routed_kern / 1e9
    traffic = None
    if net_used:
        kname = "k_posv_mx_routed<4>"
        kdesc = ("k_posv_mx_routed<4> (features + network + bf16/fp32 factor on MFMA + float64 refinement in one "
                 "kernel; the event window also holds the fallback-count reset and the float64 tier's launch over "
                 "the matrices routed to it, k_posv_mx64<4>)")
    else:
        kname = "k_posv_mx<4,F32>" if rt.bucket(n) in R.fp32_direct else "k_posv_mx64<4,1>"
        kdesc = (f"{kname} (the shipped router's policy at n={n}: {policy}; the event window also holds the "
                 "fallback-count reset and the float64 tier's launch over the fallback list)")
    try:  # HBM bytes per launch measured by rocprofv3 for this workload (profiles/traffic.json)
        if batch == 131072 and nrhs == 1:
            traffic = json.loads((ROOT / "profiles" / "traffic.json").read_text())["posv_n64_b131072_nrhs1"][kname]
    except (OSError, KeyError, ValueError):
        traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kdesc,
                "bytes_per_matrix": bytes_per,
                "fixed_tier_kernel": {"kernel": f"k_posv_mx<4> ({a.factor})", "ms": round(mixed_kern * 1e3, 4),
                                      "achieved": round(bytes_per * batch / mixed_kern / 1e9, 1)},
                "plain_fp64_kernel": {"kernel": "k_posv<double,64,1>", "ms": round(plain_kern * 1e3, 4),
                                      "achieved": round(bytes_per * batch / plain_kern / 1e9, 1)}}

    out = {
        "metric": METRIC, "value": round(routed_rate, 1), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wall_r / a.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"posv n={n} SPD kappa log-uniform [1e1,1e6], batch={batch}, nrhs={nrhs}, fp64 inputs",
                   "tier": "routed: the shipped router (tinysolverdx.router.default): per size bucket the network "
                           "over {bf16+IR, fp32+IR, fp64} per matrix, or one tier for every matrix where its timing "
                           "on the training distribution found that faster; refinement <=10 corrections to tol "
                           "1e-14 with fp64 fallback",
                   "routed_policy_at_n": policy,
                   "routed_policy_ns_per_matrix": R.policy_times.get(str(rt.bucket(n))),
                   "routed_tiers_bf16_fp32_fp64": tiers, "routed_refinement_fallbacks": fallback_r,
                   "routed_bwd_err_le_1e-14": f"{ok_r}/{batch}",
                   "routed_over_plain": round(routed_rate / plain_rate, 3),
                   "network_solves_per_s": round(net_rate, 1),
                   "network_over_plain": round(net_rate / plain_rate, 3),
                   "network_tiers_bf16_fp32_fp64": tiers_n,
                   "network_bwd_err_le_1e-14": f"{ok_n}/{batch}",
                   "fixed_tier": a.factor, "fixed_tier_solves_per_s": round(mixed_rate, 1),
                   "bwd_err_le_1e-14": f"{ok_m}/{batch}", "fp64_fallback": fallback,
                   "plain_fp64_solves_per_s": round(plain_rate, 1),
                   "plain_fp64_bwd_err_le_1e-14": f"{ok}/{batch}",
                   "fixed_tier_over_plain": round(mixed_rate / plain_rate, 3),
                   "per_kappa_decade": decades,
                   # the reference's convention (its bench.py): posv = n^3/3 + 2 n^2 nrhs
                   "gflops_routed_fixed_plain": [round(r * (n ** 3 / 3 + 2 * n * n * nrhs) / 1e9, 1)
                                                 for r in (routed_rate, mixed_rate, plain_rate)]},
        "roofline": roofline,
    }
    _copy_probe(a)
    if not a.no_cpu:
        log("[bench] CPU baseline (reference posv on the host cores)")
        out["cpu_baseline"] = cpu_baseline(A, B, nrhs)
        log(f"[bench] CPU baseline {out['cpu_baseline']['value']:.3g} solves/s on {out['cpu_baseline']['cores']} cores")
    else:
        out["cpu_baseline"] = None
    return out



# --------------------------------------------------------------------------
# The other workloads.  The factorizations work in place, so every step first
# restores A from a pristine copy and pushes A out of the 256 MB Infinity Cache
# (an untimed read-modify-write of 1 GiB); the step's time is the HIP-event
# time of the factorization alone, summed over the K steps.

def _copy_probe(a):
    """With --copy-probe: three plain 1 GiB device copies, the reference
    kernel for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes."""
    if not a.copy_probe:
        return
    import torch
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("[bench] copy probe: 3 x 1 GiB device copies done")


def _inplace_steps(fn, restore, steps, warmup):
    import torch
    flush = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    for _ in range(warmup):
        restore()
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        restore()
        flush.add_(1.0)
        e0.record(s)
        fn()
        e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, sum(e0.elapsed_time(e1) for e0, e1 in ev) * 1e-3


def _line(metric, value, unit, a, ms, dtype, config, roofline, cpu):
    return {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config, "roofline": roofline,
            "cpu_baseline": cpu}


def _roof(bytes_total, sec, kernel, per_unit, traffic_key=None):
    ach = bytes_total / sec / 1e9
    traffic = None
    try:  # HBM bytes per launch measured by rocprofv3 for this workload (profiles/traffic.json)
        if traffic_key:
            traffic = json.loads((ROOT / "profiles" / "traffic.json").read_text())[traffic_key][kernel]
    except (OSError, KeyError, ValueError):
        traffic = None
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel, "bytes_per_matrix": per_unit}


def _cpu_timed(fn, units, unit, sample):
    """Best of up to 5 runs (<= ~20 s) of a reference CPU call on the host cores."""
    from oracle import available
    best, total, reps = float("inf"), 0.0, 0
    while reps < 5 and total < 20.0:
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        best, total, reps = min(best, dt), total + dt, reps + 1
    return {"value": units / best, "unit": unit, "cores": len(os.sched_getaffinity(0)),
            "kind": "reference" if available("reference") else "port", "cpu_model": cpu_model(),
            "sample": f"{sample}, best of {reps} reps"}


def _cpu_oracle():
    from oracle import Oracle, available
    orc = Oracle("reference" if available("reference") else "port")
    orc.set_num_threads(len(os.sched_getaffinity(0)))
    return orc


def make_spd_iid(n, batch, lmin, seed, chunk=16384):
    """A = Q diag(lambda) Q^T symmetrised, lambda i.i.d. log-uniform on [lmin, 1]."""
    import math
    import torch
    from tinysolverdx import matrices
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.empty((batch, n, n), dtype=torch.float64, device="cuda")
    for s in range(0, batch, chunk):
        e = min(batch, s + chunk)
        q = matrices.haar_device(torch.randn((e - s, n, n), generator=g, device="cuda", dtype=torch.float64))
        lam = 10.0 ** (math.log10(lmin) * torch.rand((e - s, n), generator=g, device="cuda", dtype=torch.float64))
        a = torch.matmul(q * lam[:, None, :], q.transpose(1, 2).contiguous())
        A[s:e] = 0.5 * (a + a.transpose(1, 2))
    return A


def run_potrf(a):
    """Plain fp32 Cholesky, n = 32, 65536 matrices (lambda log-uniform on [1e-3, 1])."""
    import torch
    import tinysolverdx  # noqa: F401
    from tinysolverdx import device
    n, batch = 32, 65536
    log(f"[bench] potrf: generating {batch} SPD n={n} on {torch.cuda.get_device_name(0)}")
    A0 = make_spd_iid(n, batch, 1e-3, seed=20261203).float()
    A = A0.clone()
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    wall, ev = _inplace_steps(lambda: device.potrf(A, info), lambda: A.copy_(A0), a.steps, a.warmup)
    bad = int((info != 0).sum().item())
    L = A[:256].double()
    res = (torch.linalg.matrix_norm(L @ L.transpose(1, 2) - A0[:256].double()) /
           torch.linalg.matrix_norm(A0[:256].double())).max().item()
    ms = ev / a.steps * 1e3
    per = 2 * 4 * n * n + 4
    log(f"[bench] potrf fp32 n={n}: {ms:.3f} ms/launch, info>0: {bad}, max ||A-LL^T||/||A|| (256 sampled) {res:.2e}")
    # parity over the whole batch: the timed run's factors and info against
    # the reference's CPU backend (compiled from its sources, or the oracle)
    orc = _cpu_oracle()
    Ah = A0.cpu().numpy()
    L0, i0 = orc.cholesky(Ah)
    parity = {"checked": batch, "factor_bit_mismatches": bit_mismatches(A.cpu().numpy(), L0),
              "info_mismatches": int((info.cpu().numpy() != i0).sum()), "against": orc.kind}
    log(f"[bench] potrf parity vs {orc.kind}: {parity}")
    cpu = None
    if not a.no_cpu:
        cpu = _cpu_timed(lambda: orc.cholesky(Ah), batch, "matrices/s", f"full workload ({batch} matrices)")
    return _line("potrf matrices/s fp32 n=32 (65536 SPD); HBM GB/s", batch / (ms * 1e-3), "matrices/s", a, ms,
                 "f32", {"workload": f"potrf fp32 n={n}, batch={batch}, Q diag(lambda) Q^T, lambda log-uniform "
                                     "[1e-3,1]", "info_gt0": bad, "max_rel_residual_sampled": res,
                         "parity": parity,
                         "gflops": round(batch / (ms * 1e-3) * n ** 3 / 3 / 1e9, 1),
                         "wall_ms_per_step_incl_restore": round(wall / a.steps * 1e3, 4)},
                 _roof(per * batch, ms * 1e-3, "k_potrf_rows<float,8,4,1,1,false,true,false>", per,
                       f"potrf_f32_n{n}_b{batch}"), cpu)


def run_getrf(a):
    """LU with partial pivoting + gesv (4 right-hand sides), ~1 GiB of A, 1% zero columns."""
    import torch
    import tinysolverdx  # noqa: F401
    from tinysolverdx import device
    n = a.n
    dt = torch.float32 if a.dtype == "f32" else torch.float64
    es = 4 if a.dtype == "f32" else 8
    batch = (1 << 30) // (es * n * n)
    g = torch.Generator(device="cuda").manual_seed(20261204)
    log(f"[bench] getrf: {batch} N(0,1) {a.dtype} n={n}")
    A0 = torch.randn((batch, n, n), generator=g, device="cuda", dtype=dt)
    zc = torch.nonzero(torch.rand(batch, generator=g, device="cuda") < 0.01)[:, 0]
    cols = torch.randint(0, n, (zc.numel(),), generator=g, device="cuda")
    A0[zc, :, cols] = 0.0
    A = A0.clone()
    piv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    wall, ev = _inplace_steps(lambda: device.getrf(A, piv, info), lambda: A.copy_(A0), a.steps, a.warmup)
    _copy_probe(a)
    ms = ev / a.steps * 1e3
    bad = int((info != 0).sum().item())
    B = torch.randn((batch, n, 4), generator=g, device="cuda", dtype=dt)
    X = torch.empty_like(B)
    wall_s, ev_s = time_steps(lambda: device.gesv_solve(A0, B, X, piv, info), a.steps, a.warmup)
    ms_s = ev_s / a.steps * 1e3
    per = 2 * es * n * n + 4 * n + 4
    per_s = es * n * n + 2 * es * n * 4 + 4 * n + 4
    log(f"[bench] getrf {a.dtype} n={n}: {ms:.3f} ms, gesv(4) {ms_s:.3f} ms, info>0: {bad} "
        f"(zeroed columns {zc.numel()})")
    # parity over the whole batch: pivots, info and factors of the timed run
    # against the reference's CPU backend, bit for bit
    orc = _cpu_oracle()
    Afull = A0.cpu().numpy()
    LU0, p0, i0 = orc.lu(Afull)
    parity = {"checked": batch, "pivot_mismatches": int((piv.cpu().numpy() != p0).any(axis=1).sum()),
              "info_mismatches": int((info.cpu().numpy() != i0).sum()),
              "factor_bit_mismatches": bit_mismatches(A.cpu().numpy(), LU0), "against": orc.kind}
    del Afull, LU0
    log(f"[bench] getrf parity vs {orc.kind}: {parity}")
    cpu = None
    if not a.no_cpu:
        take = min(batch, max(1024, int(4e8 // (n ** 3))))  # a bounded sample
        Ah = A0[:take].cpu().numpy()
        cpu = _cpu_timed(lambda: orc.lu(Ah), take, "matrices/s", f"first {take} of the {batch} matrices")
    cfg = {"workload": f"getrf {a.dtype} n={n}, batch={batch} (~1 GiB of A), N(0,1), 1% zeroed columns",
           "info_gt0": bad, "zeroed_columns": int(zc.numel()), "parity": parity,
           "gflops": round(batch / (ms * 1e-3) * 2 * n ** 3 / 3 / 1e9, 1),
           "gesv4_matrices_per_s": round(batch / (ms_s * 1e-3), 1),
           "gesv4_gflops": round(batch / (ms_s * 1e-3) * (2 * n ** 3 / 3 + 2 * n * n * 4) / 1e9, 1),
           "gesv4_GBps": round(per_s * batch / (ms_s * 1e-3) / 1e9, 1),
           "wall_ms_per_step_incl_restore": round(wall / a.steps * 1e3, 4)}
    return _line(f"getrf matrices/s {a.dtype} n={n} (~1 GiB of A); HBM GB/s", batch / (ms * 1e-3), "matrices/s",
                 a, ms, a.dtype, cfg, _roof(per * batch, ms * 1e-3, _getrf_kernel(es, n, batch), per,
                       f"getrf_{a.dtype}_n{n}_b{batch}"), cpu)


def _getrf_kernel(es, n, batch):
    """Name of the kernel getrf launches for this element size, n and batch."""
    nb = max(4, 1 << (n - 1).bit_length())
    ztile = "true" if n <= 24 else "false"  # the reference's zero-pivot rule by n (tile kernel up to 24)
    if es == 4 and nb == 64 and batch >= 16384:  # two rows per lane (csrc/rows_kernels.hpp, GetrfShape)
        return f"k_getrf_rows<float,32,2,1,{ztile}>"
    return f"k_getrf<{'float' if es == 4 else 'double'},{nb},{ztile}>"


def ragged_inputs(count, dtype, seed):
    """Zipf sizes on 2..64 (P(n) ~ 1/n), half SPD (kappa log-uniform [1, 1e4]),
    half N(0,1), packed with every matrix starting on a 16-byte boundary."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    ns = np.arange(2, 65)
    p = 1.0 / ns
    sizes = rng.choice(ns, size=count, p=p / p.sum()).astype(np.int32)
    kinds = (rng.random(count) < 0.5).astype(np.int32)
    al = 4
    span = ((sizes.astype(np.int64) ** 2 + al - 1) // al) * al
    offs = np.zeros(count, np.int64)
    offs[1:] = np.cumsum(span)[:-1]
    poffs = np.zeros(count, np.int64)
    poffs[1:] = np.cumsum(sizes.astype(np.int64))[:-1]
    tdt = torch.float32 if dtype == "f32" else torch.float64
    A = torch.zeros(int(span.sum()), dtype=tdt, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    d_off = torch.from_numpy(offs).cuda()
    for kind in (0, 1):
        for n in range(2, 65):
            ids = np.nonzero((sizes == n) & (kinds == kind))[0]
            if ids.size == 0:
                continue
            if kind == 0:
                M, _ = make_spd(n, ids.size, 1.0 + 1e-9, 1e4, seed=seed + n)
            else:
                M = torch.randn((ids.size, n, n), generator=g, device="cuda", dtype=torch.float64)
            idx = d_off[torch.from_numpy(ids).cuda()][:, None] + torch.arange(n * n, device="cuda")[None, :]
            A[idx.reshape(-1)] = M.reshape(-1).to(tdt)
    return A, sizes, kinds, offs, poffs


def run_ragged(a):
    """1,000,000 matrices, Zipf sizes 2..64, half Cholesky half LU, one graph launch."""
    import numpy as np
    import torch
    import tinysolverdx  # noqa: F401
    from tinysolverdx import device
    count = a.batch if a.batch != 131072 else 1_000_000
    t0 = time.perf_counter()
    A0, sizes, kinds, offs, poffs = ragged_inputs(count, a.dtype, 20261206)
    log(f"[bench] ragged {a.dtype}: {count} matrices generated in {time.perf_counter() - t0:.1f}s")
    A = A0.clone()
    d_off, d_poff = torch.from_numpy(offs).cuda(), torch.from_numpy(poffs).cuda()
    d_size, d_kind = torch.from_numpy(sizes).cuda(), torch.from_numpy(kinds).cuda()
    piv = torch.zeros(int(sizes.sum()), dtype=torch.int32, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    t0 = time.perf_counter()
    plan = device.RaggedPlan(A, d_off, d_size, d_kind, piv, d_poff, info, tune=True)
    log(f"[bench] plan + autotune {time.perf_counter() - t0:.1f}s: {plan.report}")
    wall, ev = _inplace_steps(plan.run, lambda: A.copy_(A0), a.steps, a.warmup)
    _copy_probe(a)
    ms = ev / a.steps * 1e3
    es = 4 if a.dtype == "f32" else 8
    n64 = sizes.astype(np.int64)
    total = int((2 * es * n64 * n64 + 4 * n64 * kinds + 4).sum())
    bad = int((info != 0).sum().item())
    # the same matrices as per-size launches (one launch per kind and n), for
    # comparison of time and, bit for bit, of every factor, pivot and info;
    # and every (kind, n) group against the reference's CPU backend
    per_ms = 0.0
    s = torch.cuda.current_stream()
    got_A, got_piv, got_info = A.cpu().numpy(), piv.cpu().numpy(), info.cpu().numpy()
    orc = _cpu_oracle()
    Ah0 = A0.cpu().numpy()
    mism = {"per_size": 0, "oracle": 0}
    outs = []
    for kind in (0, 1):
        for n in range(2, 65):
            ids = np.nonzero((sizes == n) & (kinds == kind))[0]
            if ids.size == 0:
                continue
            idx = d_off[torch.from_numpy(ids).cuda()][:, None] + torch.arange(n * n, device="cuda")[None, :]
            D = A0[idx.reshape(-1)].reshape(ids.size, n, n)
            I = torch.zeros(ids.size, dtype=torch.int32, device="cuda")
            P = torch.zeros((ids.size, n), dtype=torch.int32, device="cuda")
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            if kind == 0:
                device.potrf(D, I)
            else:
                device.getrf(D, P, I)
            e1.record(s)
            torch.cuda.synchronize()
            per_ms += e0.elapsed_time(e1)
            outs.append((kind, n, ids, D, I, P))
    # parity after all launches are timed (long host work between timed
    # launches would let the clocks drop)
    for kind, n, ids, D, I, P in outs:
        gi = (offs[ids][:, None] + np.arange(n * n)[None, :])
        pi = (poffs[ids][:, None] + np.arange(n)[None, :])
        mine = (got_A[gi], got_info[ids], got_piv[pi] if kind == 1 else None)
        ref_dev = (D.cpu().numpy().reshape(ids.size, -1), I.cpu().numpy(), P.cpu().numpy() if kind == 1 else None)
        if kind == 0:
            Lo, Io = orc.cholesky(Ah0[gi].reshape(ids.size, n, n))
            ref_cpu = (Lo.reshape(ids.size, -1), Io, None)
        else:
            Lo, Po, Io = orc.lu(Ah0[gi].reshape(ids.size, n, n))
            ref_cpu = (Lo.reshape(ids.size, -1), Io, Po)
        for key, ref in (("per_size", ref_dev), ("oracle", ref_cpu)):
            bad_m = np.zeros(ids.size, bool)
            g_u = mine[0].view({4: np.uint32, 8: np.uint64}[mine[0].dtype.itemsize]).copy()
            r_u = ref[0].view(g_u.dtype).copy()
            g_u[np.isnan(mine[0])] = 0
            r_u[np.isnan(ref[0])] = 0
            bad_m |= ((g_u != r_u) | (np.isnan(mine[0]) != np.isnan(ref[0]))).any(axis=1)
            bad_m |= mine[1] != ref[1]
            if kind == 1:
                bad_m |= (mine[2] != ref[2]).any(axis=1)
            mism[key] += int(bad_m.sum())
    del outs
    log(f"[bench] ragged {a.dtype}: {ms:.3f} ms per graph launch, per-size launches {per_ms:.3f} ms, info>0: {bad}, "
        f"matrices differing from the per-size launches / the {orc.kind} CPU backend: {mism}")
    cpu = None
    if not a.no_cpu:
        take = min(count, 100_000)
        Ah = Ah0
        groups = []
        for kind in (0, 1):
            for n in range(2, 65):
                ids = np.nonzero((sizes[:take] == n) & (kinds[:take] == kind))[0]
                if ids.size:
                    groups.append((kind, np.stack([Ah[offs[m]:offs[m] + n * n].reshape(n, n) for m in ids])))

        def cpu_run():
            for kind, M in groups:
                (orc.cholesky if kind == 0 else orc.lu)(M)
        cpu = _cpu_timed(cpu_run, take, "matrices/s", f"first {take} of the {count} matrices, one call per kind and n")
    plan.close()
    cfg = {"workload": f"ragged {a.dtype}: {count} matrices, n Zipf on 2..64 (P(n) ~ 1/n), half SPD Cholesky "
                       "(kappa log-uniform [1,1e4]) half N(0,1) LU, one graph launch",
           "plan": {k: v for k, v in plan.report.items()}, "info_gt0": bad,
           "parity": {"checked": count, "mismatches_vs_per_size_launches": mism["per_size"],
                      "mismatches_vs_cpu_reference": mism["oracle"], "cpu_reference": orc.kind},
           "per_size_launches_ms": round(per_ms, 4), "graph_over_per_size": round(per_ms / ms, 3),
           "gflops": round(float(((n64.astype(np.float64) ** 3) * np.where(kinds == 1, 2.0 / 3.0, 1.0 / 3.0)).sum())
                           / (ms * 1e-3) / 1e9, 1),
           "wall_ms_per_step_incl_restore": round(wall / a.steps * 1e3, 4)}
    return _line(f"ragged factor matrices/s {a.dtype} (1M, n 2..64, half Cholesky half LU); HBM GB/s",
                 count / (ms * 1e-3), "matrices/s", a, ms, a.dtype, cfg,
                 _roof(total, ms * 1e-3, "k_ragged<T,class,variant> x3 (one graph)", round(total / count, 1),
                       f"ragged_{a.dtype}_{count}"), cpu)



def run_router(a):
    """Learned router: trained here on separately seeded matrices (labels from
    the refinement pipeline), then the routed posv against the un-routed fp32
    refinement and the plain float64 posv on 2^20 SPD matrices, n uniform on
    {16, 32, 64}, kappa log-uniform on [1, 1e12]; every answer judged at 1e-14."""
    import numpy as np
    import torch
    import tinysolverdx  # noqa: F401
    from tinysolverdx import device, router as rt
    total = a.batch if a.batch != 131072 else 1 << 20
    t0 = time.perf_counter()
    feat, label = rt.training_set(seed=20261208)
    R = rt.Router(seed=3)
    hist = R.fit(feat, label, epochs=60, batch=4096, lr=3e-3, seed=3)
    train_acc = float((R.predict(feat) == label).float().mean().item())
    shift = R.calibrate(rt.training_set.costs[64])  # expected-cost decision between fp32 and float64
    # per-size policy: the network or one tier for every matrix, whichever is fastest (timed, fresh seed)
    direct, direct32, dtimes = rt.choose_direct(R)
    R.direct, R.fp32_direct = frozenset(direct), frozenset(direct32)
    net = rt.Router(R.params.cpu().numpy())  # the same network at every size, for comparison
    log(f"[bench] router trained in {time.perf_counter() - t0:.1f}s: loss {hist[0]:.3f} -> {hist[-1]:.3f}, "
        f"labels {np.bincount(label.cpu().numpy(), minlength=3).tolist()}, train accuracy {train_acc:.3f}, "
        f"fp64 logit shift {shift:.3f}, direct buckets float64 {direct} fp32 {direct32} (ns/matrix "
        f"{ {n: {k: round(v * 1e9, 2) for k, v in t.items()} for n, t in dtimes.items()} })")
    sizes = (16, 32, 64)
    per = total // 3
    ms = {"routed": 0.0, "network": 0.0, "unrouted": 0.0, "fp64": 0.0}
    ok = {"routed": 0, "network": 0, "unrouted": 0, "fp64": 0}
    tiers = np.zeros(3, np.int64)
    wasted = {"routed": 0, "unrouted": 0}
    count = 0
    for k, n in enumerate(sizes):
        A, kappa = make_spd(n, per, 1.0, 1e12, seed=20261209 + n)
        g = torch.Generator(device="cuda").manual_seed(n)
        B = torch.randn((per, n, 1), generator=g, device="cuda", dtype=torch.float64)
        X = torch.empty_like(B)
        info = torch.zeros(per, dtype=torch.int32, device="cuda")
        stat = torch.zeros(per, dtype=torch.int32, device="cuda")
        tier = torch.zeros(per, dtype=torch.int32, device="cuda")
        wr = torch.empty(tinysolverdx._lib.lib().tsdx_hip_posv_routed_workspace_size(per), dtype=torch.uint8,
                         device="cuda")
        wi = device.posv_ir_workspace(per)
        runs = {"routed": lambda: R.posv(A, B, X, info=info, stat=stat, tier=tier, work=wr),
                "network": lambda: net.posv(A, B, X, info=info, stat=stat, work=wr),
                "unrouted": lambda: device.posv_ir(A, B, X, factor="fp32", info=info, stat=stat, work=wi),
                "fp64": lambda: device.posv_solve(A, B, X, info)}
        for name, fn in runs.items():
            _, ev = time_steps(fn, a.steps, a.warmup)
            ms[name] += ev / a.steps * 1e3
            ok[name] += int((backward_error(A, X, B) <= 1e-14).sum().item())
            st = stat.cpu().numpy()
            if name == "routed":
                tr = tier.cpu().numpy()
                tiers += np.bincount(tr, minlength=3)[:3]
                wasted["routed"] += int(((st < 0) & (tr != 2)).sum())
            elif name == "unrouted":
                wasted["unrouted"] += int((st < 0).sum())
        count += per
        del A, B, X
        log(f"[bench] router n={n}: cumulative ms {ms}")
    rate = {k: count / (v * 1e-3) for k, v in ms.items()}
    es_bytes = sum(8 * n * n + 2 * 8 * n + 4 for n in sizes) * per
    log(f"[bench] router: routed {rate['routed'] / 1e6:.2f} M/s, network at every size {rate['network'] / 1e6:.2f} M/s, "
        f"un-routed fp32 {rate['unrouted'] / 1e6:.2f} M/s, fp64 {rate['fp64'] / 1e6:.2f} M/s, tiers {tiers.tolist()}, "
        f"wasted attempts {wasted}")
    cfg = {"workload": f"routed posv, {count} SPD fp64 matrices, n uniform on {{16,32,64}}, kappa log-uniform "
                       "[1,1e12], nrhs=1; router trained in this run on the shipped training mix (tinysolverdx.router.TRAINING_MIX), separately seeded",
           "train_loss_first_last": [round(hist[0], 4), round(hist[-1], 4)], "train_accuracy": round(train_acc, 4),
           "direct_float64_buckets": direct, "direct_fp32_buckets": direct32, "fp64_logit_shift": round(shift, 4),
           "policy_ns_per_matrix": {n: {k: round(v * 1e9, 2) for k, v in t.items()} for n, t in dtimes.items()},
           "network_at_every_size_solves_per_s": round(rate["network"], 1),
           "train_labels_bf16_fp32_fp64": np.bincount(label.cpu().numpy(), minlength=3).tolist(),
           "routed_tiers_bf16_fp32_fp64": tiers.tolist(),
           "unrouted_fp32_solves_per_s": round(rate["unrouted"], 1), "plain_fp64_solves_per_s": round(rate["fp64"], 1),
           "routed_over_unrouted": round(rate["routed"] / rate["unrouted"], 3),
           "routed_over_plain_fp64": round(rate["routed"] / rate["fp64"], 3),
           "wasted_refinement_attempts": wasted, "bwd_err_le_1e-14": {k: f"{v}/{count}" for k, v in ok.items()}}
    return _line("routed posv solves/s (2^20, n in {16,32,64}, kappa up to 1e12) at bwd err<=1e-14; HBM GB/s",
                 rate["routed"], "solves/s", a, ms["routed"], "f64", cfg,
                 _roof(es_bytes, ms["routed"] * 1e-3, "k_posv_mx_routed<4> or the float64 tier per size bucket",
                       round(es_bytes / count, 1)), None)


if __name__ == "__main__":
    sys.exit(main())
