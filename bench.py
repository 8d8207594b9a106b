"""Benchmark of the RBF-FD weight build and apply on one MI355X.

Workload: the learned-surrogate evaluation configuration: a fresh 4,000,000-point 2-D
cloud whose density varies x10 across [0,1]^2, Laplacian with PHS(3), polynomial
degree 2, k = 20 (n = 26 per stencil). A step is one weight build of all 4M stencils
from the device-resident adjacency (the work of update_weights!); the k-NN and the
apply are timed alongside and reported as extra fields.

Prints one JSON line (see DESIGN.md for every field).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def density_cloud(n, seed):
    """x from the density 1 + 9x (x10 variation), y uniform."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    # inverse CDF of p(x) = (1 + 9x) / 5.5
    x = (-1.0 + np.sqrt(1.0 + 9.0 * 11.0 * u)) / 9.0
    y = rng.random(n)
    return np.stack([x, y], 1)


def flops_per_stencil(n, r):
    return n ** 3 / 3.0 + 2.0 * n * n * r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--cpu-sample", type=int, default=200_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU figures are not produced "
                                   "(see DESIGN.md for how the path would shard)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import rbfx_loader

    R = rbfx_loader.load()
    dev = R._lib.device_for("cuda:0")
    k, deg = 20, 2
    m = 6
    n_sys = k + m
    pts = density_cloud(args.n, seed=20240404)
    d_pts = torch.as_tensor(pts, device=dev)
    stream = torch.cuda.current_stream(dev)

    # k-NN (reported separately)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    adj = R._lib.knn(d_pts, d_pts, k)
    torch.cuda.synchronize()
    knn_s = time.perf_counter() - t0

    f = np.sin(4 * pts[:, 0]) * np.cos(3 * pts[:, 1])
    d_f = torch.as_tensor(f, device=dev)
    exact = -25.0 * f

    results = {}
    ops = {}
    for method in ("fp64", "fp32_refined"):
        op = R.RadialBasisOperator(R.Laplacian(), d_pts, basis=R.PHS(3, poly_deg=deg), k=k, adjl=adj,
                                   precision=method)
        for _ in range(args.warmup):
            op.update_weights()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            op.update_weights()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1) / args.steps
        y = op(d_f).cpu().numpy()
        nrmse = float(np.sqrt(np.mean((y - exact) ** 2)) / np.sqrt(np.mean(exact ** 2)))
        results[method] = dict(ms_per_build=gpu_ms, wall_ms_per_build=1e3 * wall / args.steps,
                               stencils_per_s=args.n / (gpu_ms * 1e-3), nrmse=nrmse,
                               status=op.status_counts[:5])
        ops[method] = op

    # apply: y = W f (Laplacian, 1 field), HBM traffic per SURVEY formula
    op = ops["fp64"]
    X = d_f[:, None].contiguous()
    for _ in range(3):
        op._spmv(X)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    ev0.record(stream)
    for _ in range(reps):
        op._spmv(X)
    ev1.record(stream)
    torch.cuda.synchronize()
    apply_ms = ev0.elapsed_time(ev1) / reps
    nnz = args.n * k
    apply_bytes = 4 * (args.n + 1) * 0 + nnz * (4 + 8) + 8 * (args.n + args.n)
    apply_gbs = apply_bytes / (apply_ms * 1e-3) / 1e9

    plain = results["fp64"]
    best_name = max(results, key=lambda kk: results[kk]["stencils_per_s"])
    best = results[best_name]
    fl = flops_per_stencil(n_sys, 1)
    achieved_tf = best["stencils_per_s"] * fl / 1e12
    peak = FP64_PEAK_TFLOPS if best_name == "fp64" else FP32_PEAK_TFLOPS

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        ns = min(args.cpu_sample, args.n)
        sample_adj = adj[:ns].cpu().numpy()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        lap = O.abi.make_operator(1, [(0, O.abi.OP_LAP, 0, 0, 1.0)])
        t0 = time.perf_counter()
        O.build(pts, pts[:ns], sample_adj, O.basis_struct(3, 2), lap, scaled=False, threads=threads)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": ns / cpu_s, "unit": "stencils/s", "cores": threads, "kind": "port",
               "sample": f"{ns} stencils (the first {ns} eval points of the same 4M cloud), "
                         f"C restatement of the reference arithmetic, OpenMP"}

    out = {
        "metric": "weight-build stencils/s, learned+FP32-refined vs plain FP64, at equal operator err",
        "value": best["stencils_per_s"],
        "unit": "stencils/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_build"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best_name == "fp64" else "f32+f64-refine",
        "data": "synthetic (seeded 2-D cloud, density x10 across the domain)",
        "config": {"workload": "laplacian-2d-4M-phs3-deg2-k20", "n_points": args.n, "k": k, "poly_deg": deg,
                   "basis": "PHS3", "method": best_name},
        "methods": results,
        "speedup_vs_plain_fp64": best["stencils_per_s"] / plain["stencils_per_s"],
        "error_ratio_vs_plain": best["nrmse"] / plain["nrmse"],
        "learned": None,
        "knn_ms": knn_s * 1e3,
        "apply": {"ms": apply_ms, "GB_s": apply_gbs, "frac_hbm_peak": apply_gbs / HBM_PEAK_GBS,
                  "bytes": apply_bytes},
        "roofline": {"bound": "valu-fp64" if best_name == "fp64" else "valu-fp32", "achieved": achieved_tf,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved_tf / peak, "traffic": None,
                     "flops_per_stencil": fl},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
