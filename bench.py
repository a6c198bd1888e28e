"""Benchmark: vegas+ iterations of the 8-D double-peak workload on one MI355X.

A step is one vegas+ iteration (allocation, sampling, map, integrand,
per-hcube sums, training histogram, adapt, running average) over
neval = 1e7 evaluations of the 8-D double Gaussian, all on the device.
`value` is integrand evaluations per second over the K timed steps (points
already resident: the whole iteration is generated on the device).

Also reported:
  roofline      the fused sampler kernel's algorithmic FP32 work per launch
                (evaluations x flops/eval) over its average HIP-event time
  cpu_baseline  the FP64 CPU oracle (reference-structured C restatement,
                1 thread) on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "integrand evals/s; GPU wall-time to 1e-4 rel err (vegas grid vs learned map)"
PEAK_FP32_TFLOPS = 157.3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=15)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-iters", type=int, default=5)
    p.add_argument("--copy-calibration", action="store_true",
                   help="only copy 512 MiB device to device 4 times (HBM counter calibration)")
    p.add_argument("--workload", default="double_peak_8d",
                   choices=["double_peak_8d", "ridge_8d", "batched_4096", "flow_ridge_9d", "moments_8d"],
                   help="double_peak_8d (default): the headline; ridge_8d: 1000-Gaussian diagonal "
                        "ridge, neval 2e7 (diagonal-ridge and generic mixture functors); batched_4096: "
                        "4096 independent Gaussian integrals in one graph, nitn 8 per call (--steps "
                        "calls timed); flow_ridge_9d: time to --rtol, learned map against the grid")
    p.add_argument("--no-tune", action="store_true", help="batched: skip the work-item autotuner")
    p.add_argument("--rtol", type=float, default=1e-4, help="flow_ridge_9d: target relative error")
    return p.parse_args()


def batched(a):
    """4096 independent single-Gaussian integrals (dim 2-10, sigma 0.01-0.1,
    neval 1e4-1e6 log-uniform), 8 iterations per call, one hipGraph."""
    import numpy as np
    import torch
    from vegas_amd import workloads
    from vegas_amd.batch import BatchIntegrator, gaussian_params
    probs = workloads.batched_gaussians(4096)
    dims = [p["dim"] for p in probs]
    nev = [p["neval"] for p in probs]
    cs = [p["center"] for p in probs]
    sig = [p["sigma"] for p in probs]
    exact = np.array([p["exact"] for p in probs])
    A, B = gaussian_params(dims, sig)
    bi = BatchIntegrator(dims, nev)
    tuned = None
    if not a.no_tune:
        tuned = bi.autotune(A, B, cs, nitn=2)
    for _ in range(max(a.warmup, 1)):
        bi(A, B, cs, nitn=8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        bi.launch(A, B, cs, nitn=8)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    r = bi(A, B, cs, nitn=8)
    evals = float(r.sum_neval)  # per call (the timed calls draw the same counts to within the allocation)
    pull = (r.mean - exact) / np.where(r.sdev > 0, r.sdev, np.inf)
    # plain vegas+ misses some narrow high-dimensional peaks at 1e4 samples
    # (as the reference does): those return ~0 with a tiny error
    resolved = r.mean > 0.5 * exact
    out = {
        "metric": METRIC, "value": evals * a.steps / dt, "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (device Philox samples)",
        "config": {"workload": "4096 independent Gaussian integrals, dim U{2..10}, sigma log-U[0.01,0.1], "
                               "neval log-U[1e4,1e6], ninc <= 500, nitn 8 per step (one hipGraph launch)",
                   "integrals": 4096, "evals_per_step": int(evals), "tiles_per_item": bi.tiles_per_item,
                   "autotune_ms": getattr(bi, "tune_times", None), "parallelism": "single GPU"},
        "result": {"integrals": len(exact), "resolved": int(resolved.sum()),
                   "resolved_within_5sigma": int((np.abs(pull[resolved]) < 5).sum()),
                   "median_rel_sdev_resolved": float(np.median(r.sdev[resolved] / exact[resolved]))},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def copy_calibration():
    """Plain device copy of a known size: FETCH_SIZE / WRITE_SIZE per dispatch
    of this copy, against 512 MiB, scale the fused kernel's counters to bytes."""
    import torch
    n = 512 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "dispatches": 4}))


def cpu_baseline(wl, iters):
    """Reference-structured FP64 oracle on `iters` iterations of the workload
    (integrand inlined in C): one thread alone, then one independent run per
    host core (up to 16, the box's CPU share) side by side; the C call runs
    without the GIL.  value is the all-core rate."""
    import threading
    from oracle import oracle as orc
    f = wl["f"]
    g = orc.GaussMix(f.A, f.B, f.centers)

    errors = []

    def one(seed, out):
        try:
            I = orc.OracleIntegrator([(0, 1)] * wl["dim"], neval=wl["neval"], seed=seed)
            _, _, nev = I.run(g, iters)
            out.append(float(nev.sum()))
        except Exception as e:  # reported below: a partial count would misstate the rate
            errors.append(e)

    solo = []
    t0 = time.perf_counter()
    one(1, solo)
    dt1 = time.perf_counter() - t0
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(ncores, 16))
    outs = []
    ths = [threading.Thread(target=one, args=(s + 1, outs)) for s in range(ncores)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dtn = time.perf_counter() - t0
    if errors or len(outs) != ncores:
        raise RuntimeError("CPU baseline run failed: %r" % (errors[:1],))
    return dict(value=sum(outs) / dtn, unit="evals/s", cores=ncores, kind="port",
                value_1core=solo[0] / dt1,
                sample="%d iteration(s) of the 8-D double peak, neval 1e7, through the FP64 C oracle: "
                       "%d independent runs on %d threads in %.1f s (%.3g evals); one thread alone "
                       "%.3g evals/s" % (iters, ncores, ncores, dtn, sum(outs), solo[0] / dt1))


def ridge(a):
    """8-D ridge: mean of 1000 isotropic Gaussians (sigma 0.005) centred on
    the diagonal, neval 2e7 per iteration.  The same integrand through the
    diagonal-ridge functor (|x - t 1|^2 = P + d (xbar - t)^2) and through the
    generic mixture functor; value = the diagonal functor's evals/s."""
    import torch
    import vegas_amd as va
    from vegas_amd import _lib, workloads
    out_r = {}
    for name, diag in (("diagonal", True), ("generic", False)):
        wl = workloads.ridge_8d(diagonal=diag)
        integ = va.Integrator([[0, 1]] * 8, neval=wl["neval"], seed=2026, precision=wl["precision"])
        f = wl["f"]
        if a.warmup > 0:
            integ(f, nitn=a.warmup)
        torch.cuda.synchronize()
        L, cx = _lib.lib(), _lib.ctx()
        L.vgx_set_timing(cx, 1)
        t0 = time.perf_counter()
        r = integ(f, nitn=a.steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms = (C.c_float * max(a.steps, 1))()
        got = L.vgx_fused_times(cx, ms, max(a.steps, 1))
        L.vgx_set_timing(cx, 0)
        kms = [ms[i] for i in range(got)]
        out_r[name] = {"evals_per_s": float(r.sum_neval) / dt, "ms_per_iteration": dt * 1e3 / a.steps,
                       "kernel_ms_avg": sum(kms) / len(kms) if kms else None,
                       "evals_per_iteration": float(r.sum_neval) / a.steps,
                       "mean": r.mean, "sdev": r.sdev, "chi2_dof": r.chi2 / max(r.dof, 1),
                       "nstrat": [int(s) for s in integ.nstrat], "nhcube": int(integ.nhcube)}
    d = out_r["diagonal"]
    fpe = workloads.flops_per_eval(8, 1000)
    achieved = d["evals_per_iteration"] * fpe / (d["kernel_ms_avg"] * 1e-3) / 1e12 if d["kernel_ms_avg"] else None
    out = {
        "metric": METRIC, "value": d["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": d["ms_per_iteration"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (device Philox samples)",
        "config": {"workload": "vegas+ 8-D ridge: 1000 Gaussians sigma 0.005 on the diagonal, [0,1]^8, "
                               "neval 2e7/iteration", "dim": 8, "neval": 20000000,
                   "nstrat": d["nstrat"], "nhcube": d["nhcube"], "parallelism": "single GPU"},
        "result": out_r,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None, "traffic": None,
                     "kernel": "fused_kernel<8, true, DiagRidge32>", "kernel_ms_avg": d["kernel_ms_avg"],
                     "flops_per_eval": fpe,
                     "note": "algorithmic flops of the 1000-term mixture (the diagonal functor evaluates "
                             "it in about a third of them); peak is the FP32 vector peak"},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def moments(a):
    """Multi-component integrand through the staged path: the 8-D double peak
    g and its moments x0 g, x0^2 g, x0 x1 g (4 components, full covariance)
    evaluated by a torch integrand on the device, neval 1e7 per iteration.
    value = evaluations/s of the whole staged iteration (device sampling,
    the torch integrand, the scalar and covariance accumulations, adapt)."""
    import torch
    import vegas_amd as va
    from vegas_amd import workloads
    wl = workloads.double_peak_8d()
    f = wl["f"]
    A = torch.as_tensor(f.A, device="cuda")
    B = torch.as_tensor(f.B, device="cuda")
    c = torch.as_tensor(f.centers, device="cuda")

    @va.gpu_lbatchintegrand
    def g(x):
        r2 = ((x[:, None, :] - c[None, :, :]) ** 2).sum(-1)
        v = (A[None, :] * torch.exp(-B[None, :] * r2)).sum(-1)
        return torch.stack([v, x[:, 0] * v, x[:, 0] ** 2 * v, x[:, 0] * x[:, 1] * v], dim=1)

    integ = va.Integrator([[0, 1]] * 8, neval=wl["neval"], seed=2026, min_neval_batch=2000000)
    if a.warmup > 0:
        integ(g, nitn=a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = integ(g, nitn=a.steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {
        "metric": METRIC, "value": float(r.sum_neval) / dt, "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (device Philox samples)",
        "config": {"workload": "8-D double peak and 3 moments (4-component integrand, correlated), torch "
                               "integrand on the device, staged batches of >= 2e6, neval 1e7/iteration",
                   "parallelism": "single GPU"},
        "result": {"mean": [float(v) for v in r.mean], "sdev": [float(v) for v in r.sdev],
                   "chi2_dof": r.chi2 / max(r.dof, 1), "Q": r.Q},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def flow_ridge(a):
    """GPU wall-time to a relative error rtol on the 9-D ridge (500 diagonal
    Gaussians, sigma 0.01): the learned map (vegas warm-up, 2e7 vegas
    samples, flow training, flow sampling) against the plain vegas+ grid
    (neval 2e7 iterations), which gets 3x the flow's time.  With --warmup > 0
    one untimed run of the whole pipeline (other seeds) comes first."""
    import torch
    import vegas_amd as va
    from vegas_amd import workloads
    from vegas_amd.flow import FlowMap
    wl = workloads.ridge_9d(diagonal=True)
    f, exact = wl["f"], wl["exact"]
    # warm-up: one whole untimed run (vegas warm-up, training, sampling) with
    # other seeds, so the timed run excludes first-use costs (module loads,
    # allocator growth) as the iteration benches do
    nwarm = min(max(a.warmup, 0), 1)
    for _ in range(nwarm):
        wi = va.Integrator([[0, 1]] * 9, neval=wl["neval"], seed=11)
        wi(f, nitn=10)
        wm = FlowMap(9, seed=7)
        wm.fit(f, wi, ntrain=2e7)
        wm.integrate(f, rtol=a.rtol, batch=20000000, max_samples=2e10)
        del wi, wm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    integ = va.Integrator([[0, 1]] * 9, neval=wl["neval"], seed=5)
    integ(f, nitn=10)
    fm = FlowMap(9, seed=1)
    fm.fit(f, integ, ntrain=2e7)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    est, nflow = fm.integrate(f, rtol=a.rtol, batch=20000000, max_samples=2e10)
    torch.cuda.synchronize()
    t_flow = time.perf_counter() - t0
    flow_ok = est.sdev < a.rtol * abs(est.mean) and abs(est.mean - exact) < 5 * est.sdev
    # plain vegas+: iterations of 2e7 until sdev < rtol |mean| or 3 t_flow
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    vi = va.Integrator([[0, 1]] * 9, neval=wl["neval"], seed=5)
    r = None
    nit = 0
    while time.perf_counter() - t1 < 3 * t_flow:
        r = vi(f, nitn=10)
        nit += 10
        if r.sdev < a.rtol * abs(r.mean):
            break
    torch.cuda.synchronize()
    t_vegas = time.perf_counter() - t1
    vegas_ok = r.sdev < a.rtol * abs(r.mean) and abs(r.mean - exact) < 5 * r.sdev
    out = {
        "metric": METRIC, "value": t_flow, "unit": "s (GPU wall-time to rel err %g, learned map)" % a.rtol,
        "n_gpus": 1, "steps": 1, "warmup": nwarm, "ms_per_step": t_flow * 1e3, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "bf16 (flow MLP), f32 (integrand), f64 (sums)",
        "data": "synthetic (device Philox samples)",
        "config": {"workload": "9-D ridge: 500 diagonal Gaussians sigma 0.01, [0,1]^9; 6-layer affine coupling "
                               "flow, MLP 9-128-128 per layer, trained on 2e7 vegas samples in 2.5e5 batches",
                   "parallelism": "single GPU"},
        "result": {"flow": {"mean": est.mean, "sdev": est.sdev, "samples": int(nflow), "train_s": t_train,
                            "reached_target": bool(flow_ok)},
                   "vegas": {"mean": r.mean, "sdev": r.sdev, "iterations": nit, "seconds": t_vegas,
                             "chi2_dof": r.chi2 / max(r.dof, 1), "reached_target": bool(vegas_ok)},
                   "exact": exact,
                   "speedup_vs_vegas": (t_vegas / t_flow) if vegas_ok else ">= %.1f (vegas did not reach the "
                                                                          "target in 3x the flow's time)" % (
                                                                              t_vegas / t_flow)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.workload == "batched_4096":
        batched(a)
        return
    if a.workload == "flow_ridge_9d":
        flow_ridge(a)
        return
    if a.workload == "ridge_8d":
        ridge(a)
        return
    if a.workload == "moments_8d":
        moments(a)
        return
    import torch

    import vegas_amd as va
    from vegas_amd import _lib, workloads

    wl = workloads.double_peak_8d()
    integ = va.Integrator([[0, 1]] * wl["dim"], neval=wl["neval"], seed=2026,
                          precision=wl["precision"])
    f = wl["f"]
    if a.warmup > 0:
        integ(f, nitn=a.warmup)
    torch.cuda.synchronize()
    L, cx = _lib.lib(), _lib.ctx()
    L.vgx_set_timing(cx, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = integ(f, nitn=a.steps)  # reads back once, at the end
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nt = max(a.steps, 1)
    ms = (C.c_float * nt)()
    got = L.vgx_fused_times(cx, ms, nt)
    L.vgx_set_timing(cx, 0)
    evals = [int(g) for g in getattr(integ, "_last_itn_neval", [])] or [integ.last_neval] * a.steps
    total = float(r.sum_neval)
    value = total / dt
    fpe = workloads.flops_per_eval(wl["dim"], wl["ncomp"])
    kern_ms = [ms[i] for i in range(got)]
    avg_ms = sum(kern_ms) / len(kern_ms) if kern_ms else float("nan")
    avg_evals = total / a.steps
    achieved = avg_evals * fpe / (avg_ms * 1e-3) / 1e12 if kern_ms else None
    traffic = None
    prof = os.path.join(ROOT, "profiles", "fused_double_peak_8d_traffic.json")
    if os.path.exists(prof):
        with open(prof) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (device Philox samples)",
        "config": {"workload": "vegas+ 8-D double peak, sigma 0.02, [0,1]^8, neval 1e7/iteration",
                   "dim": 8, "neval": int(wl["neval"]), "nstrat": [int(s) for s in integ.nstrat],
                   "nhcube": int(integ.nhcube), "ninc": [int(n) for n in integ.map.ninc],
                   "precision": "integrand/RNG f32; map, weights, sums f64",
                   "parallelism": "single GPU"},
        "result": {"mean": r.mean, "sdev": r.sdev, "chi2_dof": r.chi2 / max(r.dof, 1), "Q": r.Q,
                   "exact": wl["exact"]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
                     "traffic": traffic, "kernel": "fused_kernel<8, true, GaussMix32>",
                     "kernel_ms_avg": avg_ms, "flops_per_eval": fpe,
                     "hbm_GBps": (traffic / (avg_ms * 1e-3) / 1e9) if (traffic and kern_ms) else None,
                     "hbm_frac": (traffic / (avg_ms * 1e-3) / 8.0e12) if (traffic and kern_ms) else None,
                     "note": "VALU-bound (no MFMA in this kernel); peak is the FP32 vector peak; "
                             "hbm_GBps = measured HBM bytes per launch / launch time, against 8 TB/s"},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_iters)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
