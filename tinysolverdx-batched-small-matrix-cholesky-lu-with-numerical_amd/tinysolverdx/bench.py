"""Per-size throughput sweep of the GPU factorizations, with the roofline of
the measured machine ceilings (the reference's benchmark module,
python/tinysolverdx/bench.py:1-137, measures its CPU backends the same way).

For every kind (Cholesky, LU) and size n the sweep times, on one GPU:
  "hip device (kernel)"      the stream-ordered entry point on device-resident
                             matrices, HIP-event time of the factorization alone
                             (A restored from a copy before every timed launch);
  "hip host API (with PCIe)" tinysolverdx.cholesky / lu on NumPy arrays: the
                             synchronous host-pointer path, copies included.
Rows go to bench_<precision>.csv with the reference's columns (kind, n,
precision, series, gflops, batch, threads; threads is the number of compute
units for the GPU series) plus the algorithmic bandwidth and its fraction of
the measured HBM bandwidth; the chart goes to roofline_<precision>.png.

    python -m tinysolverdx.bench --out DIR [--quick] [--precision fp32|fp64]
"""
import argparse
import os
import pathlib
import time

import numpy as np

import tinysolverdx as tsdx
from tinysolverdx import matrices

SIZES = [2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64]
SERIES = ("hip device (kernel)", "hip host API (with PCIe)")


def flops(kind, n):
    """The reference's flop convention: n^3/3 (Cholesky), 2n^3/3 (LU)."""
    return n ** 3 / 3 if kind == "chol" else 2 * n ** 3 / 3


def alg_bytes(kind, n, item):
    """Read and write A, info; pivots for LU."""
    return 2 * n * n * item + 4 + (4 * n if kind == "lu" else 0)


def _device_seconds(kind, A0, reps):
    """Smallest HIP-event time of the in-place factorization of a copy of A0."""
    import torch
    from tinysolverdx import device
    A = A0.clone()
    info = torch.empty(A0.shape[0], dtype=torch.int32, device=A0.device)
    piv = torch.empty(A0.shape[:2], dtype=torch.int32, device=A0.device)
    run = (lambda: device.potrf(A, info)) if kind == "chol" else (lambda: device.getrf(A, piv, info))
    best = float("inf")
    s = torch.cuda.current_stream()
    for _ in range(reps + 1):
        A.copy_(A0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        run()
        e1.record(s)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    return best


def _host_seconds(kind, A, prec, reps):
    f = tsdx.cholesky if kind == "chol" else tsdx.lu
    f(A, precision=prec)  # staging buffers grow once
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        f(A, precision=prec)
        best = min(best, time.perf_counter() - t0)
    return best


def sweep(prec, sizes, quick):
    import torch
    dtype = np.float32 if prec == "fp32" else np.float64
    item = np.dtype(dtype).itemsize
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    rows = []
    for kind in ("chol", "lu"):
        for n in sizes:
            # ~256 MiB of matrices on the device (beyond the Infinity Cache),
            # a smaller batch through the host API
            batch = max(4096, (1 << 28) // (item * n * n))
            if quick:
                batch = min(batch, 65536)
            rng = np.random.default_rng(n)
            make = matrices.spd if kind == "chol" else matrices.general
            base = make(rng, 1024, n, 10.0).astype(dtype)
            A = np.concatenate([base] * (batch // 1024 + 1))[:batch]
            t_dev = _device_seconds(kind, torch.from_numpy(A).cuda(), 3 if quick else 10)
            hb = min(batch, 16384 if quick else 65536)
            t_host = _host_seconds(kind, A[:hb], prec, 2 if quick else 5)
            for series, t, b in zip(SERIES, (t_dev, t_host), (batch, hb)):
                gbs = alg_bytes(kind, n, item) * b / t / 1e9
                rows.append(dict(kind=kind, n=n, precision=prec, series=series, gflops=flops(kind, n) * b / t * 1e-9,
                                 batch=b, threads=cus, gbs=gbs))
            r = {x["series"]: x for x in rows if x["kind"] == kind and x["n"] == n}
            print(f"{kind:4s} n={n:2d} device {r[SERIES[0]]['gflops']:9.1f} GFLOP/s {r[SERIES[0]]['gbs']:7.1f} GB/s  "
                  f"host API {r[SERIES[1]]['gflops']:8.1f} GFLOP/s", flush=True)
    return rows


def write_csv(path, rows):
    cols = list(rows[0])
    text = [",".join(cols)] + [",".join(str(r[c]) for c in cols) for r in rows]
    pathlib.Path(path).write_text("\n".join(text) + "\n")


def plot(rows, out, prec, peak_gflops, bw_gbs):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    item = 4 if prec == "fp32" else 8
    fig, axes = plt.subplots(1, 2, figsize=(12.5, 4.6), sharey=True)
    marks = {SERIES[0]: ("o", "-"), SERIES[1]: ("s", "--")}
    for ax, kind, title in zip(axes, ("chol", "lu"), ("batched Cholesky", "batched LU, partial pivoting")):
        ns = sorted({r["n"] for r in rows if r["kind"] == kind})
        for name, (mk, ls) in marks.items():
            ys = [next(r["gflops"] for r in rows if r["kind"] == kind and r["n"] == n and r["series"] == name)
                  for n in ns]
            ax.loglog(ns, ys, ls, marker=mk, label=name)
        nn = np.geomspace(min(ns), max(ns), 100)
        ai = np.array([flops(kind, v) / alg_bytes(kind, v, item) for v in nn])  # flop per byte moved
        ax.loglog(nn, np.minimum(peak_gflops, bw_gbs * ai), "k-", lw=1.0, alpha=0.7,
                  label=f"roofline: {bw_gbs:.0f} GB/s, {peak_gflops / 1e3:.0f} TFLOP/s (measured)")
        ax.set_title(f"{title} ({prec}, one MI355X)")
        ax.set_xlabel("n")
        ax.set_xticks(ns)
        ax.set_xticklabels([str(v) for v in ns])
        ax.grid(alpha=0.3, which="both")
    axes[0].set_ylabel("GFLOP/s")
    axes[1].legend(fontsize=8, loc="lower right")
    fig.tight_layout()
    fig.savefig(out / f"roofline_{prec}.png", dpi=120)
    plt.close(fig)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=os.environ.get("TSDX_OUT_DIR", "bench_report"))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--no-plot", action="store_true")
    a = ap.parse_args(argv)
    out = pathlib.Path(a.out)
    out.mkdir(parents=True, exist_ok=True)
    peak = tsdx.measure_peak_gflops(a.precision)
    bw = tsdx.measure_bandwidth_gbs()
    print(f"measured: {a.precision} vector peak {peak:.0f} GFLOP/s, HBM triad {bw:.0f} GB/s", flush=True)
    rows = sweep(a.precision, [2, 8, 16, 32, 64] if a.quick else SIZES, a.quick)
    for r in rows:
        r["frac_of_measured_hbm"] = r["gbs"] / bw
    write_csv(out / f"bench_{a.precision}.csv", rows)
    if not a.no_plot:
        plot(rows, out, a.precision, peak, bw)
    print(f"wrote {out}/bench_{a.precision}.csv", flush=True)
    return rows


if __name__ == "__main__":
    main()
