"""Every benchmark configuration run at its stated size on the GPU.

* TGV 3D 192 x 512 x 512 (30 ramp ellipsoids + N(0, 0.03^2), lambda 0.02,
  alpha1 1, alpha0 2, L2 12): the FP32 solve against the oracle bit for bit
  (the reference's own CPU-vs-GPU tests compare whole volumes the same way,
  test/test_3d_cpu_vs_gpu.py:109-168), for TGV_ITERS iterations of the 500
  (the oracle runs them on the host's cores).
* PD-TV 3D 256^3 (the FGP volume, lambda 0.02, L 12, iso, nonneg) for its
  stated 400 iterations with tolerance 1e-6 under the reference's
  every-5 rule, through the drop-in TV_PD_GPU_main: the stop iteration and the
  output against the oracle's, bit for bit.
* The 64-volume uneven batch (half FGP 300 iterations, half ROF 1000): every
  job's batched result bit-identical to its single solve, with FP32 and with
  BF16 dual storage (single solves then also store P in BF16).
* The served Poisson stream: its first SERVE_N requests replayed in real time
  through the server; every result identical to the same request solved
  alone, with the iteration count.
"""
import os
import time

import numpy as np
import pytest

import oracle
from conftest import bits_equal, rel_l2

pytestmark = pytest.mark.gpu

TGV_ITERS = int(os.environ.get("CILREG_TEST_TGV_ITERS", "500"))
SERVE_N = 50


def _gpu(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_tgv_192x512x512_matches_oracle():
    from ccpi.supp import phantoms
    from ccpi.filters import regularisers as R
    x = phantoms.noisy(phantoms.ramp_ellipsoids_3d((192, 512, 512), 30, seed=3), 0.03, seed=103)
    info = np.zeros(2, dtype=np.float32)
    t0 = time.perf_counter()
    y = R.TGV(x, 0.02, 1.0, 2.0, TGV_ITERS, 12.0, 0.0, device="gpu", infovector=info)
    t1 = time.perf_counter()
    print(f"TGV 192x512x512, {TGV_ITERS} iterations: GPU (C ABI, host arrays) {t1 - t0:.2f} s", flush=True)
    y_ref, info_ref = oracle.cpu_ref().tgv(x, 0.02, 1.0, 2.0, TGV_ITERS, 12.0, 0.0)
    print(f"  oracle {time.perf_counter() - t1:.1f} s", flush=True)
    assert info[0] == info_ref[0] == TGV_ITERS
    assert bits_equal(y, y_ref), f"rel L2 {rel_l2(y, y_ref):.3e}"


def test_pd_256_full_solve_matches_oracle():
    from ccpi.supp import phantoms
    from ccpi.filters import regularisers as R
    x = phantoms.noisy(phantoms.ellipsoids_3d((256, 256, 256), 40, seed=2), 0.05, seed=102)
    info = np.zeros(2, dtype=np.float32)
    y = R.PD_TV(x, 0.02, 400, 1e-6, 12.0, 0, 1, device="gpu", infovector=info)
    t0 = time.perf_counter()
    y_ref, info_ref = oracle.cpu_ref().pd(x, 0.02, 400, 1e-6, 12.0, 0, 1)
    print(f"PD 256^3: GPU {info}, oracle {info_ref} ({time.perf_counter() - t0:.1f} s)", flush=True)
    if info[0] != info_ref[0]:
        # a stop at another check (FP64 vs serial FP32 residual sums): the
        # output must still be the oracle's iterate at the GPU's stop
        with oracle.forced_stop(int(info[0])):
            y_ref, info_f = oracle.cpu_ref().pd(x, 0.02, 400, 1e-6, 12.0, 0, 1)
        assert info_f[0] == info[0]
    assert bits_equal(y, y_ref), f"rel L2 {rel_l2(y, y_ref):.3e}"
    assert info[0] == info_ref[0], "stopped at another check than the reference's rule"


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_64_volume_batch_matches_single_solves(prec):
    import torch
    from ccpi.filters import batch as B, device as D
    from ccpi.supp.workloads import batched_jobs, solve_alone
    jobs = batched_jobs()
    xs = [B.phantom(shape, *ph) for shape, ph, _ in jobs]
    specs = [dict(sp, input=x) for (_, _, sp), x in zip(jobs, xs)]
    t0 = time.perf_counter()
    got = B.batch_run(specs, precision=prec, autotune=True)
    torch.cuda.synchronize()
    print(f"batch of 64 ({prec}): {time.perf_counter() - t0:.2f} s", flush=True)
    bad = []
    for i, (sp, (y, info)) in enumerate(zip(specs, got)):
        kw = {"precision": prec} if sp["method"] == "fgp" else {}
        r, rinfo = solve_alone(D, sp, sp["input"], **kw)
        if not (torch.equal(y, r) and info[0] == rinfo[0] == sp["iters"]):
            bad.append((i, sp["method"], tuple(sp["input"].shape)))
    assert not bad, f"{len(bad)} of 64 jobs differ from their single solves: {bad[:5]}"


def test_served_stream_results_match_single_solves():
    import torch
    from ccpi.filters import batch as B, device as D
    from ccpi.supp.workloads import request_stream, solve_alone
    reqs = request_stream()[:SERVE_N]
    with B.Server(0, hbm_budget=200e9) as srv:
        r = B.serve_stream(srv, reqs, levels=8, keep_results=True)
    lat = np.array(r["latency_s"])
    print(f"served {SERVE_N} requests: makespan {r['makespan_s']:.2f} s, p50 {np.percentile(lat, 50):.2f} s, "
          f"p99 {np.percentile(lat, 99):.2f} s, mean jobs per step {r['mean_active_per_step']:.1f}", flush=True)
    bad = []
    for i, ((_, sp), (y, info)) in enumerate(zip(reqs, r["results"])):
        x = B.phantom(sp["shape"], *sp["phantom"])
        ref, rinfo = solve_alone(D, sp, x)
        if not (torch.equal(y, ref) and info[0] == rinfo[0] == sp["iters"]):
            bad.append((i, sp["method"], sp["shape"]))
        del x, ref
    assert not bad, f"{len(bad)} of {SERVE_N} requests differ from their single solves: {bad[:5]}"
