#!/usr/bin/env python3
"""Benchmark of the MI355X regularisers (one JSON line on stdout).

Headline workload (the configuration the metric is quoted on): FGP-TV 3D on
a 256^3 float32 phantom of 40 random ellipsoids plus N(0, 0.05^2) noise,
lambda = 0.02, 400 iterations, isotropic TV, nonnegativity on, tolerance 1e-6
checked every iteration.  One step = one full solve with the input already in
HBM; value = voxel-iterations per second over the timed steps.

Also reported:
* roofline: the fused FGP iteration kernel's algorithmic bytes per launch
  (48 B/voxel with the every-iteration residual check) over its mean launch
  time, measured with HIP events on the solve's stream around every
  iteration's kernel;
* cpu_baseline: the reference C-OMP library (oracle/_ref, built from the
  reference sources) or, where it is absent, the C restatement (oracle/), on a
  bounded sample of the same workload, on the host's cores;
* secondary: TGV 3D 192x512x512 (500 iterations, FP32), PD-TV 3D 256^3
  (400 iterations, L = 12) and ROF-TV 2D 512^2 (2000 iterations) solves.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
       python bench.py --workload {tgv3d,pd3d,copy} --steps K   (single-workload runs for
       profiler passes; 'copy' copies a 256 MiB buffer K times, the reference kernel of
       known size that calibrates the FETCH_SIZE / WRITE_SIZE counters)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ccpi-regularisation-toolkit_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "voxel-iters/s & HBM GB/s (% peak) FGP/TGV 3D; served p50/p99 latency"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fgp_tv_3d_256_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fgp_workload():
    from ccpi.supp import phantoms
    clean = phantoms.ellipsoids_3d((256, 256, 256), 40, seed=2)
    return phantoms.noisy(clean, 0.05, seed=102)


FGP_PARAMS = dict(lam=0.02, iters=400, tol=1e-6, method=0, nonneg=1, check_every=1)


def run_fgp(D, x, profile=False):
    p = FGP_PARAMS
    return D.fgp(x, p["lam"], p["iters"], p["tol"], p["method"], p["nonneg"], check_every=p["check_every"],
                 profile=profile)


def read_traffic(n_voxels):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            line = line.strip()
            if line.startswith("#") or "=" not in line:
                continue
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(x_np, iters):
    import numpy as np
    import oracle
    lib = oracle.reference()
    kind = "reference"
    if lib is None:
        lib = oracle.cpu_ref()
        kind = "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p = FGP_PARAMS
    t0 = time.perf_counter()
    lib.fgp(x_np, p["lam"], iters, 0.0, p["method"], p["nonneg"])
    dt = time.perf_counter() - t0
    vi = x_np.size * iters / dt
    return {"value": vi, "unit": "voxel-iters/s", "cores": threads, "kind": kind,
            "sample": f"FGP-TV 3D 256^3, {iters} of 400 iterations (fixed count), {dt:.1f} s",
            "seconds": round(dt, 2)}


def secondary(D, torch):
    """TGV 3D, PD 3D and ROF 2D solves (one timed run each after one warm-up)."""
    from ccpi.supp import phantoms
    out = {}
    # TGV 3D precision study: FP32 / BF16 / FP16 storage of P, Q, v_bar against an FP64 solve
    log("bench: TGV 3D 192x512x512 phantom")
    clean = phantoms.ramp_ellipsoids_3d((192, 512, 512), 30, seed=3)
    x = torch.from_numpy(phantoms.noisy(clean, 0.03, seed=103)).cuda()
    del clean
    n = x.numel()
    tgv_args = (0.02, 1.0, 2.0)   # lambda (reference tests' value), alpha1, alpha0; L2 = 12

    def objective(u):
        # E = sum 2 lambda |grad+ u| + sum (u - f)^2 in FP64, replicate boundary (utils.c:89-117)
        ud = u.double()
        gx = torch.diff(ud, dim=2, append=ud[:, :, -1:])
        gy = torch.diff(ud, dim=1, append=ud[:, -1:, :])
        gz = torch.diff(ud, dim=0, append=ud[-1:, :, :])
        e = (2 * tgv_args[0] * torch.sqrt(gx * gx + gy * gy + gz * gz)).sum() + ((ud - x.double()) ** 2).sum()
        return float(e)

    D.tgv(x, *tgv_args, 20, 12.0, 0.0)
    results = {}
    for prec in ("fp32", "bf16", "fp16"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y, info = D.tgv(x, *tgv_args, 500, 12.0, 0.0, precision=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        D.tgv(x, *tgv_args, 50, 12.0, 0.0, precision=prec, profile=True)
        m, _, _ = D.last_profile()
        b = 108 if prec == "fp32" else 60
        results[prec] = {"y": y, "voxel_iters_per_s": n * info[0] / dt, "seconds": round(dt, 4),
                         "iterations": int(info[0]), "kernel_ms_mean": m, "alg_bytes_per_voxel_iter": b,
                         "kernel_alg_GBps": b * n / (m * 1e-3) / 1e9 if m else None}
    log("bench: TGV FP64 reference solve")
    x64 = x.double()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref, _ = D.tgv_f64(x64, *tgv_args, 500, 12.0, 0.0)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    del x64
    ref_norm = float(torch.linalg.vector_norm(ref))
    for prec, r in results.items():
        y = r.pop("y")
        r["rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(y.double() - ref)) / ref_norm
        r["objective_fp64"] = objective(y)
        out[f"tgv_3d_192x512x512_{prec}"] = r
    out["tgv_3d_192x512x512_fp64_reference"] = {"seconds": round(t64, 4), "objective_fp64": objective(ref)}
    out["tgv_3d_192x512x512_fp32"]["kernel_frac_of_8TBps"] = (
        out["tgv_3d_192x512x512_fp32"]["kernel_alg_GBps"] / HBM_PEAK_GBS)
    del ref, x
    # PD 3D
    log("bench: PD-TV 3D 256^3")
    x = torch.from_numpy(fgp_workload()).cuda()
    n = x.numel()
    D.pd(x, 0.02, 20, 1e-6, 12.0, 0, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = D.pd(x, 0.02, 400, 1e-6, 12.0, 0, 1, check_every=1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    D.pd(x, 0.02, 50, 0.0, 12.0, 0, 1, profile=True)
    m, _, _ = D.last_profile()
    out["pd_tv_3d_256_fp32"] = {
        "voxel_iters_per_s": n * info[0] / dt, "iterations": int(info[0]), "seconds": round(dt, 4),
        "kernel_ms_mean": m, "alg_bytes_per_voxel_iter": 36,
        "kernel_alg_GBps": 36 * n / (m * 1e-3) / 1e9 if m else None}
    del x
    # ROF 2D
    log("bench: ROF-TV 2D 512^2")
    x = torch.from_numpy(phantoms.noisy(phantoms.shapes_2d((512, 512), 20, seed=1), 0.05, seed=101)).cuda()
    n = x.numel()
    D.rof(x, 0.02, 100, 0.001, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = D.rof(x, 0.02, 2000, 0.001, 0.0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["rof_tv_2d_512_fp32"] = {"voxel_iters_per_s": n * info[0] / dt, "iterations": int(info[0]),
                                 "seconds": round(dt, 4)}
    return out


def single_workload(name, steps, D, torch):
    """One workload on its own (for profiler passes); prints a JSON line."""
    from ccpi.supp import phantoms
    if name == "copy":
        n = 64 * 1024 * 1024   # 256 MiB of float32
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.copy_(a)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / max(steps, 1)
        return {"workload": "copy 256 MiB f32", "bytes_read": 4 * n, "bytes_written": 4 * n,
                "GBps": 8 * n / dt / 1e9}
    if name == "tgv3d":
        x = torch.from_numpy(phantoms.noisy(phantoms.ramp_ellipsoids_3d((192, 512, 512), 30, seed=3), 0.03,
                                            seed=103)).cuda()
        out = {}
        for prec in ("fp32", "bf16", "fp16"):
            D.tgv(x, 0.02, 1.0, 2.0, steps, 12.0, 0.0, precision=prec, profile=True)
            m, _, c = D.last_profile()
            out[prec] = {"kernel_ms_mean": m, "launches": c}
        return {"workload": "TGV 3D 192x512x512", **out}
    if name == "pd3d":
        x = torch.from_numpy(fgp_workload()).cuda()
        D.pd(x, 0.02, steps, 0.0, 12.0, 0, 1, profile=True)
        m, _, c = D.last_profile()
        return {"workload": "PD-TV 3D 256^3", "kernel_ms_mean": m, "launches": c}
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=16)
    ap.add_argument("--workload", default="fgp3d", choices=["fgp3d", "tgv3d", "pd3d", "copy"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the volumes and requests shard across GPUs without communication, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from ccpi.filters import device as D
    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    if args.workload != "fgp3d":
        print(json.dumps(single_workload(args.workload, args.steps, D, torch), default=float), flush=True)
        return

    log("bench: FGP-TV 3D 256^3 phantom")
    x_np = fgp_workload()
    x = torch.from_numpy(x_np).cuda()
    n = x.numel()
    for _ in range(args.warmup):
        run_fgp(D, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    total_vi = 0
    iters = []
    for _ in range(args.steps):
        _, info = run_fgp(D, x)
        total_vi += n * int(info[0])
        iters.append(int(info[0]))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = total_vi / dt
    log(f"bench: FGP {args.steps} steps in {dt:.4f} s, iterations {iters}, {value:.4e} voxel-iters/s")

    # per-kernel timing with HIP events on the solve's stream
    run_fgp(D, x, profile=True)
    mean_ms, min_ms, cnt = D.last_profile()
    bytes_per_launch = 48 * n
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9 if mean_ms else None
    traffic = read_traffic(n)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "kernel": "fgp_iter_kernel<3,8,float,float,iso,uint32>", "kernel_ms_mean": mean_ms,
                "kernel_ms_min": min_ms, "launches_timed": cnt, "alg_bytes_per_voxel": 48}

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline, {args.cpu_iters} FGP iterations")
        cpu = cpu_baseline(x_np, args.cpu_iters)
    extras = None
    if not args.no_extras:
        del x
        torch.cuda.empty_cache()
        extras = secondary(D, torch)

    line = {
        "metric": METRIC, "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "FGP-TV 3D 256^3 f32, 40-ellipsoid phantom + N(0,0.05^2), lambda 0.02, "
                               "400 iterations, iso, nonneg, tol 1e-6 checked every iteration",
                   "voxels": n, "iterations_per_step": iters},
        "hbm_GBps_alg": achieved, "roofline": roofline, "cpu_baseline": cpu, "secondary": extras,
    }
    print(json.dumps(line, default=float), flush=True)


if __name__ == "__main__":
    main()
