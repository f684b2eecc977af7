#!/usr/bin/env python3
"""Benchmark of the MI355X regularisers (one JSON line on stdout).

Headline workload (the configuration the metric is quoted on): FGP-TV 3D on
a 256^3 float32 phantom of 40 random ellipsoids plus N(0, 0.05^2) noise,
lambda = 0.02, 400 iterations, isotropic TV, nonnegativity on, tolerance 1e-6
checked every iteration.  One step = one full solve with the input already in
HBM; value = voxel-iterations per second over the timed steps.

Also reported:
* roofline: the FGP kernel's algorithmic bytes (48 B per voxel-iteration with
  the every-iteration residual check; a launch advances two iterations) over
  its launch time, measured with HIP events on the solve's stream around every
  launch;
* cpu_baseline: the reference C-OMP library (oracle/_ref, built from the
  reference sources) or, where it is absent, the C restatement (oracle/), on a
  bounded sample of the same workload, on the host's cores; with the extras,
  also on bounded samples of every other measured workload
  (cpu_baseline.other_workloads);
* secondary: TGV 3D 192x512x512 (500 iterations; FP32, BF16 and FP16 dual
  storage against an FP64 solve), PD-TV 3D 256^3 (400 iterations, L = 12; FP32,
  BF16 and FP16 storage of P against an FP64 solve),
  ROF-TV 2D 512^2 (2000 iterations); 64 uneven volumes (half FGP, half ROF)
  solved one after another vs in one batched hipGraph launch (FP32 and BF16
  duals); a Poisson request stream (200 req/s, ROF/FGP/PD/TGV, 2D and 3D)
  replayed in real time through the continuous-batching server (p50/p99
  latency, aggregate voxel-iterations/s): the full 60 s stream once with BF16
  dual storage, and its first 2 s with FP32, BF16 and FP16 storage and
  solved plainly one after another (the offered load exceeds one device's
  throughput several times over, so the full stream takes minutes to drain).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu] [--no-batch]
                       [--no-serve] [--serve-window SECONDS] [--serve-full-window SECONDS]
       python bench.py --profile-pass --steps K --warmup 0   (the headline solves only)
       python bench.py --workload {tgv3d,pd3d,sb3d,rof3d,dtv3d,llt3d,ndf3d,d43d,copy,batch,serve} --steps K   (single-workload runs for
       profiler passes; 'copy' copies a 256 MiB buffer K times, the reference kernel of
       known size that calibrates the FETCH_SIZE / WRITE_SIZE counters)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

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
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), None)
    except OSError:
        pass
    p = FGP_PARAMS
    t0 = time.perf_counter()
    lib.fgp(x_np, p["lam"], iters, 0.0, p["method"], p["nonneg"])
    dt = time.perf_counter() - t0
    vi = x_np.size * iters / dt
    return {"value": vi, "unit": "voxel-iters/s", "cores": threads, "kind": kind, "cpu_model": model,
            "host_logical_cpus": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "unset"),
            "sample": f"FGP-TV 3D 256^3, {iters} of 400 iterations (fixed count), {dt:.1f} s",
            "seconds": round(dt, 2)}


def cpu_rows(x_np):
    """The reference CPU path on a bounded sample of every other measured
    workload (same volumes and parameters as the GPU rows, fewer iterations
    except ROF 2D, which runs its full 2000): voxel-iterations per second."""
    import numpy as np
    import oracle
    from ccpi.supp import phantoms
    lib = oracle.reference()
    kind = "reference"
    if lib is None:
        lib = oracle.cpu_ref()
        kind = "port"
    rows = {}

    def timed(name, method, x, params, label):
        t0 = time.perf_counter()
        _, info = lib.run(method, x, params)
        dt = time.perf_counter() - t0
        rows[name] = {"voxel_iters_per_s": x.size * int(info[0]) / dt, "seconds": round(dt, 3),
                      "sample": label}

    n256 = "256^3"
    timed("rof_tv_2d_512", "rof", phantoms.noisy(phantoms.shapes_2d((512, 512), 20, seed=1), 0.05, seed=101),
          {"lambda": 0.02, "iters": 2000, "tau": 0.001, "tol": 0.0}, "ROF-TV 2D 512^2, all 2000 iterations")
    timed("rof_tv_3d_256", "rof", x_np, {"lambda": 0.02, "iters": 16, "tau": 0.001, "tol": 0.0},
          f"ROF-TV 3D {n256}, 16 iterations")
    timed("pd_tv_3d_256", "pd", x_np, {"lambda": 0.02, "iters": 16, "tol": 0.0, "L": 12.0, "method": 0, "nonneg": 1},
          f"PD-TV 3D {n256}, 16 of 400 iterations")
    timed("sb_tv_3d_256", "sb", x_np, {"mu": 0.02, "iters": 8, "tol": 0.0, "method": 0},
          f"SB-TV 3D {n256}, 8 of 150 iterations")
    timed("fgp_dtv_3d_256", "dtv", x_np, {"ref": dtv_reference(), "lambda": 0.02, "iters": 8, "tol": 0.0,
                                          "eta": 0.01, "method": 0, "nonneg": 1},
          f"FGP-dTV 3D {n256}, 8 of 400 iterations")
    for name in ("llt", "ndf", "d4"):
        (a, b), kw, _ = DIFFUSION[name]
        if name == "llt":
            prm = {"lambda_rof": a, "lambda_llt": b, "iters": 16, "tau": kw["tau"], "tol": 0.0}
        elif name == "ndf":   # the reference's own update order (in place)
            prm = {"lambda": a, "sigma": b, "iters": 8, "tau": kw["tau"], "penalty": kw["penalty"], "tol": 0.0,
                   "in_place": True}
        else:
            prm = {"lambda": a, "sigma": b, "iters": 16, "tau": kw["tau"], "tol": 0.0}
        timed(f"{name}_3d_256", name, x_np, prm, f"{name} 3D {n256}, {prm['iters']} of 500 iterations")
    xt = phantoms.noisy(phantoms.ramp_ellipsoids_3d((192, 512, 512), 30, seed=3), 0.03, seed=103)
    timed("tgv_3d_192x512x512", "tgv", xt, {"lambda": 0.02, "alpha1": 1.0, "alpha0": 2.0, "iters": 4, "L2": 12.0,
                                            "tol": 0.0}, "TGV 3D 192x512x512, 4 of 500 iterations")
    del xt
    return {"kind": kind, "rows": rows}


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
        # E = sum 2 lambda |grad+ u| + sum (u - f)^2 in FP64, replicate boundary: the library's
        # cilreg_tv_energy_dev(_f64), TV_energy3D of utils.c:89-117 summed in double
        if u.dtype == torch.float64:
            return D.tv_energy(u, x.double(), tgv_args[0], 1)
        return D.tv_energy(u, x, tgv_args[0], 1)

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
    # PD precision study: P stored in FP32 / BF16 / FP16 against an FP64 solve (400 iterations, no tolerance)
    log("bench: PD-TV 3D 256^3 precision study")
    ref, _ = D.pd_f64(x.double(), 0.02, 400, 0.0, 12.0, 0, 1)
    ref_norm = float(torch.linalg.vector_norm(ref))
    for prec in ("fp32", "bf16", "fp16"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y, info = D.pd(x, 0.02, 400, 0.0, 12.0, 0, 1, precision=prec)
        torch.cuda.synchronize()
        dtp = time.perf_counter() - t0
        D.pd(x, 0.02, 50, 0.0, 12.0, 0, 1, precision=prec, profile=True)
        mp, _, _ = D.last_profile()
        b = 36 if prec == "fp32" else 24   # P at 2 B: read u_bar, f (8) + P (6); write u_bar (4) + P (6)
        out[f"pd_tv_3d_256_{prec}_duals"] = {
            "voxel_iters_per_s": n * info[0] / dtp, "iterations": int(info[0]), "seconds": round(dtp, 4),
            "kernel_ms_mean": mp, "alg_bytes_per_voxel_iter": b,
            "kernel_alg_GBps": b * n / (mp * 1e-3) / 1e9 if mp else None,
            "rel_l2_vs_fp64": float(torch.linalg.vector_norm(y.double() - ref)) / ref_norm}
        del y
    del ref
    # SB 3D (two kernels per iteration)
    log("bench: SB-TV 3D 256^3")
    D.sb(x, 0.02, 10, 0.0, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = D.sb(x, 0.02, 150, 1e-6, 0, check_every=1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    D.sb(x, 0.02, 30, 0.0, 0, profile=True)
    m, _, _ = D.last_profile()
    out["sb_tv_3d_256_fp32"] = {
        "voxel_iters_per_s": n * info[0] / dt, "iterations": int(info[0]), "seconds": round(dt, 4),
        "kernel": "sb_fused_kernel<3,8,float,iso,uint32>", "kernel_ms_mean": m, "alg_bytes_per_voxel_iter": 60,
        "kernel_alg_GBps": 60 * n / (m * 1e-3) / 1e9 if m else None}
    # FGP-dTV 3D, guided by the clean phantom with 1% noise
    log("bench: FGP-dTV 3D 256^3")
    out["fgp_dtv_3d_256_fp32"] = dtv_section(D, torch, x)
    log("bench: explicit diffusion 3D 256^3")
    out.update(diffusion_section(D, torch, x))
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


def dtv_reference():
    from ccpi.supp import phantoms
    return phantoms.noisy(phantoms.ellipsoids_3d((256, 256, 256), 40, seed=2), 0.01, seed=202)


def dtv_section(D, torch, x):
    """FGP-dTV 3D 256^3: 400 iterations, tol 1e-6 checked every iteration; kernel time per iteration."""
    n = x.numel()
    r = torch.from_numpy(dtv_reference()).cuda()
    D.dtv(x, r, 0.02, 20, 1e-6, 0.01, 0, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = D.dtv(x, r, 0.02, 400, 1e-6, 0.01, 0, 1, check_every=1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    D.dtv(x, r, 0.02, 50, 0.0, 0.01, 0, 1, profile=True)
    m, _, c = D.last_profile()
    return {"voxel_iters_per_s": n * info[0] / dt, "iterations": int(info[0]), "seconds": round(dt, 4),
            "kernel": "dtv2_iter_kernel<8,float,iso,uint32> (two iterations per launch)", "kernel_ms_mean": m,
            "iterations_timed": c,
            "alg_bytes_per_voxel_iter": 52,
            "kernel_alg_GBps": 52 * n / (m * 1e-3) / 1e9 if m else None,
            "kernel_frac_of_8TBps": 52 * n / (m * 1e-3) / 1e9 / HBM_PEAK_GBS if m else None}


DIFFUSION = {   # the reference tests' parameters (test_2d_cpu_vs_gpu.py:370-377, :431-438, :496-502)
    "llt": ((0.01, 0.0085), {"tau": 1e-4}, "llt_iter_kernel<3,8,uint32>"),
    "ndf": ((0.02, 0.017), {"tau": 0.01, "penalty": 1}, "ndf_iter_kernel<3,8,uint32>"),
    "d4": ((0.8, 0.02), {"tau": 1e-4}, "d4_iter_kernel<3,8,uint32>"),
}


def run_diffusion(D, name, x, iters, tol=0.0, profile=False, check_every=5):
    (a, b), kw, _ = DIFFUSION[name]
    if name == "llt":
        return D.llt(x, a, b, iters, kw["tau"], tol, profile=profile, check_every=check_every)
    if name == "ndf":
        return D.ndf(x, a, b, iters, kw["tau"], kw["penalty"], tol, profile=profile, check_every=check_every)
    return D.d4(x, a, b, iters, kw["tau"], tol, profile=profile, check_every=check_every)


def diffusion_section(D, torch, x):
    """LLT-ROF, NDF (Huber) and Diff4th on 256^3, 500 iterations each: whole-solve
    rate and kernel time per iteration.  12 B/voxel-iteration of HBM traffic;
    the divisions and square roots of every flux bound these kernels."""
    n = x.numel()
    out = {}
    for name in ("llt", "ndf", "d4"):
        run_diffusion(D, name, x, 10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, info = run_diffusion(D, name, x, 500)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        run_diffusion(D, name, x, 50, profile=True)
        m, _, c = D.last_profile()
        out[f"{name}_3d_256_fp32"] = {
            "voxel_iters_per_s": n * info[0] / dt, "iterations": int(info[0]), "seconds": round(dt, 4),
            "kernel": DIFFUSION[name][2], "kernel_ms_mean": m, "iterations_timed": c,
            "alg_bytes_per_voxel_iter": 12, "kernel_alg_GBps": 12 * n / (m * 1e-3) / 1e9 if m else None}
    return out


def batched_section(D, B, torch):
    """Plain (one solve after another) vs one autotuned batched hipGraph launch, FP32 and BF16 / FP16 duals."""
    from ccpi.supp.workloads import batched_jobs
    jobs = batched_jobs()
    xs = [B.phantom(shape, *ph) for shape, ph, _ in jobs]
    specs = [dict(sp, input=x) for (_, _, sp), x in zip(jobs, xs)]
    total_vi = sum(x.numel() * sp["iters"] for x, sp in zip(xs, specs))

    def plain(specs, prec):
        outs = []
        for sp in specs:
            x = sp["input"]
            if sp["method"] == "fgp":
                outs.append(D.fgp(x, sp["lambda"], sp["iters"], 0.0, 0, 0, precision=prec)[0])
            else:
                outs.append(D.rof(x, sp["lambda"], sp["iters"], sp["tau"], 0.0)[0])
        return outs

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    short = [dict(sp, iters=2) for sp in specs]
    plain(short, "fp32")   # warm-up: workspaces, code objects
    # the autotuner picks each job's pass width once per size class (grouped
    # jobs march 32-plane chunks); tuning happens here, outside the timed runs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for prec in ("fp32", "bf16", "fp16"):
        B.batch_run(short, precision=prec, autotune=True)
    t_tune = time.perf_counter() - t0
    log("bench: batched section, plain FP32")
    ref, t_plain = timed(lambda: plain(specs, "fp32"))
    log("bench: batched section, one graph FP32")
    got32, t_b32 = timed(lambda: B.batch_run(specs, autotune=True))
    identical = all(bool(torch.equal(a, b)) for a, (b, _) in zip(ref, got32))
    del got32

    out16 = {}
    for prec in ("bf16", "fp16"):
        log(f"bench: batched section, one graph {prec.upper()} duals")
        got16, t16 = timed(lambda: B.batch_run(specs, precision=prec, autotune=True))
        # error against FP64 solves of the first four FGP volumes (ROF has no dual field)
        errs = []
        for sp, (y, _) in list(zip(specs, got16))[:8]:
            if sp["method"] != "fgp":
                continue
            r64, _ = D.fgp_f64(sp["input"].double(), sp["lambda"], sp["iters"], 0.0, 0, 0)
            errs.append(float(torch.linalg.vector_norm(y.double() - r64) / torch.linalg.vector_norm(r64)))
            del r64
        del got16
        out16[prec] = {"seconds": round(t16, 4), "voxel_iters_per_s": total_vi / t16, "speedup_vs_plain": t_plain / t16,
                       "fgp_rel_l2_vs_fp64_max_of_4": max(errs) if errs else None}
    del ref
    # the same split by method: what batching and 16-bit duals can do for
    # each half, and the bound the ROF half sets on the whole batch (ROF has
    # no dual field, so its batched time is a floor for any dual storage)
    split = {}
    for m in ("fgp", "rof"):
        sub = [sp for sp in specs if sp["method"] == m]
        vi = sum(sp["input"].numel() * sp["iters"] for sp in sub)
        _, tp = timed(lambda: plain(sub, "fp32"))
        _, tb = timed(lambda: B.batch_run(sub, autotune=True))
        _, tb16 = timed(lambda: B.batch_run(sub, precision="bf16", autotune=True))
        split[m] = {"voxel_iters": vi, "plain_fp32_s": round(tp, 4), "batched_fp32_s": round(tb, 4),
                    "batched_bf16_s": round(tb16, 4), "speedup_fp32": tp / tb, "speedup_bf16": tp / tb16}
    split["ceiling_speedup_if_fgp_took_no_time"] = t_plain / split["rof"]["batched_bf16_s"]
    return {
        "workload": "64 volumes, extents U{64..320}^3, half FGP-TV 300 iterations, half ROF-TV 1000 iterations",
        "per_method": split,
        "voxel_iters": total_vi,
        "autotune_seconds": round(t_tune, 3),
        "plain_fp32": {"seconds": round(t_plain, 4), "voxel_iters_per_s": total_vi / t_plain},
        "batched_graph_fp32": {"seconds": round(t_b32, 4), "voxel_iters_per_s": total_vi / t_b32,
                               "speedup_vs_plain": t_plain / t_b32, "bit_identical_to_plain": identical},
        "batched_graph_bf16": out16["bf16"],
        "batched_graph_fp16": out16["fp16"],
    }


def served_run(B, reqs, prec, window, label):
    """One replay of `reqs` in real time through a server with `prec` dual storage."""
    import numpy as np
    log(f"bench: served stream {label}, {prec}, {len(reqs)} requests")
    t_start = time.perf_counter()

    def progress(done, total):
        log(f"bench:   {done}/{total} requests done after {time.perf_counter() - t_start:.0f} s")

    with B.Server(0, hbm_budget=200e9, precision=prec) as srv:
        r = B.serve_stream(srv, reqs, levels=8, progress=progress)
    lat = np.array(r["latency_s"])
    vis = np.array(r["voxels"], dtype=np.float64) * np.array(r["iterations"], dtype=np.float64)
    vi = float(vis.sum())
    methods = np.array([sp["method"] for _, sp in reqs])
    per_method = {}
    for m in ("rof", "fgp", "pd", "tgv"):
        sel = methods == m
        if sel.any():
            per_method[m] = {"requests": int(sel.sum()), "p50_latency_s": float(np.percentile(lat[sel], 50)),
                             "p99_latency_s": float(np.percentile(lat[sel], 99)),
                             "voxel_iters": float(vis[sel].sum())}
    return {"window_s": window, "requests": len(reqs), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "makespan_s": r["makespan_s"],
            "voxel_iters_per_s": vi / r["makespan_s"], "offered_voxel_iters_per_s": vi / window,
            "mean_jobs_on_device": r["mean_active_per_step"], "server_steps": r["steps"],
            "per_method": per_method}


def served_section(D, B, torch, window, full_window, budget_s=420.0):
    """The Poisson request stream replayed in real time through the server.

    The configuration's full stream (200 req/s for `full_window` seconds)
    runs once, with BF16 dual storage (the served form of the metric); the
    first `window` seconds of the same stream run with FP32, BF16 and FP16
    storage and plainly (one single solve after another), for the
    precision comparison and the served-vs-plain ratio."""
    from ccpi.supp.workloads import request_stream, solve_alone
    reqs = request_stream(seconds=window)
    out = {"workload": "Poisson 200 req/s, ROF/FGP/PD/TGV, 2D 256^2..2048^2 or 3D 64^3..384^3, 100..500 iterations, "
                       "HBM budget 200 GB",
           "short_window_s": window}
    # warm-up of code objects and workspaces
    with B.Server(0, hbm_budget=200e9) as srv:
        B.serve_stream(srv, [(0.0, dict(sp, iters=2)) for _, sp in reqs[:8]])
    for prec in ("fp32", "bf16", "fp16"):
        out[f"served_{prec}"] = served_run(B, reqs, prec, window, f"first {window:g} s")
    # plain form: the same requests, one single solve after another (input
    # generation included, as the server generates it at admission)
    log(f"bench: served section, the {len(reqs)} requests solved plainly")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vi = 0
    for _, sp in reqs:
        x = B.phantom(sp["shape"], *sp["phantom"])
        _, info = solve_alone(D, sp, x)
        vi += x.numel() * int(info[0])
        del x
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    out["plain_fp32"] = {"window_s": window, "requests": len(reqs), "seconds": round(tp, 4),
                         "voxel_iters_per_s": vi / tp}
    for prec in ("fp32", "bf16", "fp16"):
        out[f"served_{prec}"]["vs_plain_voxel_iters"] = (out[f"served_{prec}"]["voxel_iters_per_s"]
                                                          / out["plain_fp32"]["voxel_iters_per_s"])
    if full_window > 0:
        import numpy as np
        full = request_stream(seconds=full_window)
        key = f"served_bf16_full_{full_window:g}s"
        # predicted from the short window's served rate; a run that would not
        # drain within the budget is skipped and says so
        vi_full = float(sum(np.prod(sp["shape"]) * sp["iters"] for _, sp in full))
        predicted = vi_full / out["served_bf16"]["voxel_iters_per_s"]
        if predicted > budget_s:
            out[key] = {"skipped": f"predicted {predicted:.0f} s to drain, above the {budget_s:g} s budget; "
                                   f"bench.py --workload serve --serve-full-window {full_window:g} "
                                   "--serve-budget 1e9 runs it"}
        else:
            r = served_run(B, full, "bf16", full_window, f"full {full_window:g} s")
            r["vs_plain_voxel_iters"] = r["voxel_iters_per_s"] / out["plain_fp32"]["voxel_iters_per_s"]
            r["predicted_drain_s"] = predicted
            out[key] = r
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
            out[prec] = {"kernel_ms_mean_per_iteration": m, "iterations_timed": c}
        return {"workload": "TGV 3D 192x512x512", **out}
    if name == "pd3d":
        x = torch.from_numpy(fgp_workload()).cuda()
        D.pd(x, 0.02, steps, 0.0, 12.0, 0, 1, profile=True)
        m, _, c = D.last_profile()
        return {"workload": "PD-TV 3D 256^3", "kernel_ms_mean_per_iteration": m, "iterations_timed": c}
    if name == "rof3d":
        x = torch.from_numpy(fgp_workload()).cuda()
        D.rof(x, 0.02, 10, 0.001, 0.0)
        D.rof(x, 0.02, steps, 0.001, 0.0, profile=True)
        m, _, c = D.last_profile()
        return {"workload": "ROF-TV 3D 256^3", "kernel_ms_mean_per_iteration": m, "iterations_timed": c,
                "voxel_iters_per_s_kernel": x.numel() / (m * 1e-3) if m else None}
    if name == "sb3d":
        x = torch.from_numpy(fgp_workload()).cuda()
        D.sb(x, 0.02, steps, 0.0, 0, profile=True)
        m, _, c = D.last_profile()
        return {"workload": "SB-TV 3D 256^3", "kernel_ms_mean_per_iteration": m, "iterations_timed": c,
                "kernel_alg_GBps": 60 * x.numel() / (m * 1e-3) / 1e9 if m else None}
    if name in ("llt3d", "ndf3d", "d43d"):
        x = torch.from_numpy(fgp_workload()).cuda()
        run_diffusion(D, name[:-2], x, steps, profile=True)
        m, _, c = D.last_profile()
        return {"workload": f"{name[:-2]} 3D 256^3", "kernel_ms_mean_per_iteration": m, "iterations_timed": c,
                "voxel_iters_per_s_kernel": x.numel() / (m * 1e-3) if m else None}
    if name == "dtv3d":
        x = torch.from_numpy(fgp_workload()).cuda()
        r = torch.from_numpy(dtv_reference()).cuda()
        D.dtv(x, r, 0.02, steps, 0.0, 0.01, 0, 1, profile=True)
        m, _, c = D.last_profile()
        return {"workload": "FGP-dTV 3D 256^3", "kernel_ms_mean_per_iteration": m, "iterations_timed": c,
                "kernel_alg_GBps": 52 * x.numel() / (m * 1e-3) / 1e9 if m else None}
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=240)
    ap.add_argument("--profile-pass", action="store_true",
                    help="only the timed headline solves (for rocprofv3 passes)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-window", type=float, default=2.0,
                    help="seconds of Poisson arrivals replayed per precision and plainly by the served section")
    ap.add_argument("--serve-full-window", type=float, default=60.0,
                    help="seconds of the full stream replayed once with BF16 duals (0 skips it)")
    ap.add_argument("--serve-budget", type=float, default=420.0,
                    help="skip the full stream when its predicted drain time exceeds this many seconds")
    ap.add_argument("--workload", default="fgp3d", choices=["fgp3d", "tgv3d", "pd3d", "sb3d", "rof3d", "dtv3d", "llt3d",
                                                                "ndf3d", "d43d", "copy", "batch", "serve"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(volumes and requests shard across GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from ccpi.filters import device as D
    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    if args.workload in ("batch", "serve"):
        from ccpi.filters import batch as B
        res = (batched_section(D, B, torch) if args.workload == "batch"
               else served_section(D, B, torch, args.serve_window, args.serve_full_window, args.serve_budget))
        print(json.dumps(res, default=float), flush=True)
        return
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

    if args.profile_pass:
        print(json.dumps({"workload": "FGP-TV 3D 256^3 headline solves only", "steps": args.steps,
                          "iterations_per_step": iters, "voxel_iters_per_s": value}), flush=True)
        return

    # per-kernel timing with HIP events on the solve's stream
    run_fgp(D, x, profile=True)
    mean_ms, min_ms, cnt = D.last_profile()   # per iteration (launch time / 2)
    per_launch = 2
    bytes_per_launch = 48 * n * per_launch
    launch_ms = mean_ms * per_launch if mean_ms else None
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if mean_ms else None
    traffic = read_traffic(n)
    achievable = D.stream_bench(1 << 30, 20)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "kernel": "fgp2_iter_kernel<8,float,float,iso,uint32>", "iterations_per_launch": per_launch,
                "kernel_ms_mean": launch_ms, "kernel_ms_min": min_ms * per_launch if min_ms else None,
                "kernel_ms_mean_per_iteration": mean_ms, "iterations_timed": cnt,
                "alg_bytes_per_voxel_iteration": 48, "alg_bytes_per_launch": bytes_per_launch,
                "achievable_GBps": achievable,
                "frac_of_achievable_sum2": achieved / achievable["sum2"] if achieved else None}

    # the same solve through the reference C-ABI on host buffers (PCIe copies
    # included; the reference's every-5 check cadence)
    from ccpi.filters import regularisers as R
    R.FGP_TV(x_np, 0.02, 400, 1e-6, 0, 1, device="gpu")
    capi_info = np.zeros(2, dtype=np.float32)
    t0 = time.perf_counter()
    R.FGP_TV(x_np, 0.02, 400, 1e-6, 0, 1, device="gpu", infovector=capi_info)
    t_capi = time.perf_counter() - t0
    capi = {"seconds": round(t_capi, 4), "iterations": int(capi_info[0]),
            "voxel_iters_per_s": n * float(capi_info[0]) / t_capi,
            "note": "TV_FGP_GPU_main on host arrays, H2D/D2H included, tolerance checked every 5 iterations"}

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline, {args.cpu_iters} FGP iterations")
        cpu = cpu_baseline(x_np, args.cpu_iters)
        if not args.no_extras:
            log("bench: CPU baseline of the other workloads")
            cpu["other_workloads"] = cpu_rows(x_np)
    extras = None
    if not args.no_extras:
        del x
        torch.cuda.empty_cache()
        extras = secondary(D, torch)
        from ccpi.filters import batch as B
        torch.cuda.empty_cache()
        if not args.no_batch:
            extras["batched_uneven_volumes"] = batched_section(D, B, torch)
            torch.cuda.empty_cache()
            D.release_cache()
        if not args.no_serve:
            extras["served_stream"] = served_section(D, B, torch, args.serve_window, args.serve_full_window,
                                                     args.serve_budget)
            D.release_cache()

    line = {
        "metric": METRIC, "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "FGP-TV 3D 256^3 f32, 40-ellipsoid phantom + N(0,0.05^2), lambda 0.02, "
                               "400 iterations, iso, nonneg, tol 1e-6 checked every iteration",
                   "voxels": n, "iterations_per_step": iters},
        "hbm_GBps_alg": achieved, "roofline": roofline, "cpu_baseline": cpu, "capi_end_to_end": capi,
        "secondary": extras,
    }
    print(json.dumps(line, default=float), flush=True)


if __name__ == "__main__":
    main()
