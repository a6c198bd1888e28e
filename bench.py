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
PEAK_BF16_TFLOPS = 2500.0   # dense MFMA
PEAK_HBM_GBPS = 8000.0
TRAIN_REC_BYTES_8D = 24     # training record per evaluation at 8-D: 8 u16 bins + (w f n)^2
# algorithmic flops per sample and coupling layer of the flow (MLP D -> 128 -> 128 -> 2D)
def flow_layer_flops(d):
    return 2 * (d * 128 + 128 * 128 + 128 * 2 * d)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=15)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-iters", type=int, default=5)
    p.add_argument("--no-unordered", action="store_true",
                   help="double_peak_8d: skip the unordered-histogram comparison run (profiling)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only copy 512 MiB device to device 4 times (HBM counter calibration)")
    p.add_argument("--workload", default="double_peak_8d",
                   choices=["double_peak_8d", "ridge_8d", "batched_4096", "flow_ridge_9d", "flow_ridge_8d",
                            "moments_8d"],
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
    # the sampler's launch time: one more call of the same kernels launched
    # directly (not through the graph) with HIP events around each sampler
    # launch; algorithmic work per launch = sum_i n_i (12 d_i + 13) flops
    from vegas_amd import _lib, workloads as wls
    L, cx = _lib.lib(), _lib.ctx()
    L.vgx_set_timing(cx, 1)
    rt = bi(A, B, cs, nitn=8)
    bms = kernel_ms("b_fused_kernel")
    L.vgx_set_timing(cx, 0)
    fpe_i = np.array([wls.flops_per_eval(d, 1) for d in dims], dtype=np.float64)
    flops_it = (np.asarray(rt.itn_neval, np.float64) * fpe_i[None, :]).sum(axis=1)
    achieved = (float(flops_it.sum()) / (sum(bms) * 1e-3) / 1e12) if len(bms) == len(flops_it) else None
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None, "traffic": None,
                     "kernel": "b_fused_kernel<true>", "kernel_ms_avg": (sum(bms) / len(bms)) if bms else None,
                     "flops_per_eval": "12 d + 13 (d = the integral's dimension)",
                     "note": "sampler launches of one extra call with the kernels launched directly and "
                             "event-timed (8 launches, one per iteration); peak is the FP32 vector peak"},
        "cpu_baseline": None if a.no_cpu else batched_cpu(probs, A, B),
    }
    print(json.dumps(out))


def batched_cpu(probs, A, B, stride=64, nitn=8):
    """The FP64 C oracle on a sample of the batched workload: every
    stride-th integral (64 of 4096) for nitn iterations at its own neval
    (maxinc_axis 500), the sample's integrals spread over the host cores."""
    from oracle import oracle as orc
    pick = list(range(0, len(probs), stride))
    ncores = _ncores()
    groups = [pick[k::ncores] for k in range(ncores)]

    def job(ids):
        def run():
            tot = 0.0
            for i in ids:
                p = probs[i]
                g = orc.GaussMix([A[i]], [B[i]], [p["center"]])
                I = orc.OracleIntegrator([(0, 1)] * p["dim"], neval=p["neval"], seed=i + 1, maxinc_axis=500)
                _, _, nev = I.run(g, nitn)
                tot += float(nev.sum())
            return tot
        return run

    tot, dt = _cpu_runs([job(gr) for gr in groups if gr])
    return dict(value=tot / dt, unit="evals/s", cores=ncores, kind="port",
                sample="%d of the 4096 integrals (every %dth), %d iterations each at their own neval, "
                       "through the FP64 C oracle on %d threads: %.3g evals in %.1f s"
                       % (len(pick), stride, nitn, ncores, tot, dt))


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


def _ncores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))  # the box's CPU share


def _cpu_runs(jobs):
    """Run the oracle jobs (callables returning an evaluation count) one per
    host core side by side (the C calls run without the GIL); returns the
    total count and the wall time.  Any failure fails the baseline loudly:
    a partial count would misstate the rate."""
    import threading
    errors, outs = [], []

    def run(job):
        try:
            outs.append(float(job()))
        except Exception as e:
            errors.append(e)

    ths = [threading.Thread(target=run, args=(j,)) for j in jobs]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    if errors or len(outs) != len(jobs):
        raise RuntimeError("CPU baseline run failed: %r" % (errors[:1],))
    return sum(outs), dt


def cpu_oracle_rate(f, dim, neval, iters, what):
    """The reference-structured FP64 oracle (integrand inlined in C) on
    `iters` iterations of f at neval: one thread alone, then one independent
    run per host core (seeds 1..n) side by side.  value is the all-core rate."""
    from oracle import oracle as orc
    g = orc.GaussMix(f.A, f.B, f.centers)

    def job(seed):
        def run():
            I = orc.OracleIntegrator([(0, 1)] * dim, neval=neval, seed=seed)
            _, _, nev = I.run(g, iters)
            return nev.sum()
        return run

    n1, dt1 = _cpu_runs([job(1)])
    ncores = _ncores()
    tot, dtn = _cpu_runs([job(s + 1) for s in range(ncores)])
    return dict(value=tot / dtn, unit="evals/s", cores=ncores, kind="port", value_1core=n1 / dt1,
                sample="%d iteration(s) of %s, neval %.3g, through the FP64 C oracle: %d independent runs on "
                       "%d threads in %.1f s (%.3g evals); one thread alone %.3g evals/s"
                       % (iters, what, neval, ncores, ncores, dtn, tot, n1 / dt1))


def cpu_baseline(wl, iters):
    return cpu_oracle_rate(wl["f"], wl["dim"], wl["neval"], iters, "the 8-D double peak")


def kernel_ms(name):
    """Durations (ms) of the recorded launches of one kernel."""
    from vegas_amd import _lib
    return [ms for k, ms in _lib.timed_launches() if k == name]


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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None, "traffic": None,
                     "kernel": "fused_kernel<8, true, DiagRidge32>", "kernel_ms_avg": d["kernel_ms_avg"],
                     "flops_per_eval": fpe,
                     "note": "algorithmic flops of the 1000-term mixture (the diagonal functor evaluates "
                             "it in about a third of them); peak is the FP32 vector peak"},
        "cpu_baseline": None if a.no_cpu else cpu_oracle_rate(
            workloads.ridge_8d(diagonal=False)["f"], 8, 1e5, 1,
            "the 8-D ridge (1000 Gaussians; the per-evaluation rate on a smaller neval than the GPU's 2e7)"),
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
    from vegas_amd import _lib
    L, cx = _lib.lib(), _lib.ctx()
    L.vgx_set_timing(cx, 1)
    t0 = time.perf_counter()
    r = integ(g, nitn=a.steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cov_ms = kernel_ms("cov")
    L.vgx_set_timing(cx, 0)
    # the covariance pass reads each point's 4 values and weight once
    # (algorithmic 40 B per point; the second pass over an hcube is served
    # from cache)
    cov_bytes = float(r.sum_neval) * 40.0
    cov_gbps = cov_bytes / (sum(cov_ms) * 1e-3) / 1e9 if cov_ms else None
    out = {
        "metric": METRIC, "value": float(r.sum_neval) / dt, "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (device Philox samples)",
        "config": {"workload": "8-D double peak and 3 moments (4-component integrand, correlated), torch "
                               "integrand on the device, staged batches of >= 2e6, neval 1e7/iteration",
                   "parallelism": "single GPU"},
        "result": {"mean": [float(v) for v in r.mean], "sdev": [float(v) for v in r.sdev],
                   "chi2_dof": r.chi2 / max(r.dof, 1), "Q": r.Q},
        "roofline": {"bound": "hbm", "achieved": cov_gbps, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": (cov_gbps / PEAK_HBM_GBPS) if cov_gbps else None, "traffic": None,
                     "kernel": "cov_hcube_all_kernel<4> + cov_reduce_kernel",
                     "kernel_ms_avg": (sum(cov_ms) / len(cov_ms)) if cov_ms else None,
                     "launches": len(cov_ms),
                     "note": "algorithmic bytes 40 per point (4 FP64 values + weight); the torch "
                             "integrand's own kernels take most of the iteration"},
        "cpu_baseline": None if a.no_cpu else cpu_oracle_rate(
            wl["f"], 8, 1e6, 1, "the 8-D double peak, component 0 only (a lower bound on the reference's "
                                "4-component cost)"),
    }
    print(json.dumps(out))


def flow_ridge(a):
    """GPU wall-time to a relative error rtol on a non-separable ridge: the
    9-D ridge (500 diagonal Gaussians, sigma 0.01; flow_ridge_9d) or the 8-D
    ridge (1000 diagonal Gaussians, sigma 0.005; flow_ridge_8d).  The learned
    map (vegas warm-up, 2e7 vegas samples, flow training, flow sampling)
    against the plain vegas+ grid (neval 2e7 iterations), which gets 3x the
    flow's time.  With --warmup > 0 one untimed run of the whole pipeline
    (other seeds) comes first."""
    import torch
    import vegas_amd as va
    from vegas_amd import workloads
    from vegas_amd.flow import FlowMap
    wl = workloads.ridge_8d(diagonal=True) if a.workload == "flow_ridge_8d" else workloads.ridge_9d(diagonal=True)
    # the 9-D ridge (config 5) uses the fixed recipe swept on it; on the
    # narrower 8-D ridge the variance phase runs until more training stops
    # shortening the projected time to rtol (FlowMap.fit with rtol)
    adaptive = a.rtol if a.workload == "flow_ridge_8d" else None
    f, exact = wl["f"], wl["exact"]
    D = wl["dim"]
    # warm-up: one whole untimed run (vegas warm-up, training, sampling) with
    # other seeds, so the timed run excludes first-use costs (module loads,
    # allocator growth) as the iteration benches do
    nwarm = min(max(a.warmup, 0), 1)
    for _ in range(nwarm):
        wi = va.Integrator([[0, 1]] * D, neval=wl["neval"], seed=11)
        wi(f, nitn=10)
        wm = FlowMap(D, seed=7)
        wm.fit(f, wi, ntrain=2e7, rtol=adaptive)
        wm.integrate(f, rtol=a.rtol, batch=20000000, max_samples=2e10)
        del wi, wm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    integ = va.Integrator([[0, 1]] * D, neval=wl["neval"], seed=5)
    integ(f, nitn=10)
    fm = FlowMap(D, seed=1)
    fm.fit(f, integ, ntrain=2e7, rtol=adaptive)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    est, nflow = fm.integrate(f, rtol=a.rtol, batch=20000000, max_samples=2e10)
    torch.cuda.synchronize()
    t_flow = time.perf_counter() - t0
    flow_ok = est.sdev < a.rtol * abs(est.mean) and abs(est.mean - exact) < 5 * est.sdev
    # plain vegas+: iterations of 2e7 until sdev < rtol |mean| or 3 t_flow
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    vi = va.Integrator([[0, 1]] * D, neval=wl["neval"], seed=5)
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
    # kernel times after the timed region: two more training steps on the
    # flow's own samples (batches of 2.5e5) with the flow kernels event-timed
    from vegas_amd import _lib
    L, cx = _lib.lib(), _lib.ctx()
    nb = 250000
    L.vgx_set_timing(cx, 1)
    fm.train(f, None, None, None, epochs=0, self_batches=2, self_batch=nb, lr=1e-5, loss='variance')
    torch.cuda.synchronize()
    tl = _lib.timed_launches()
    L.vgx_set_timing(cx, 0)
    kt = {}
    for k, ms in tl:
        kt.setdefault(k, []).append(ms)
    lf = flow_layer_flops(D)
    # per layer: recomputed MLP, backpropagation and the weight gradients
    work = {"flow_backward_kernel": nb * 3 * lf,
            "flow_sample_kernel": nb * 6 * lf}     # 6 layers forward (the integrand's flops aside)
    kern = {k: {"ms_avg": sum(v) / len(v), "launches": len(v),
                "TFLOPs": work[k] / (sum(v) / len(v) * 1e-3) / 1e12} for k, v in kt.items() if k in work}
    bwd = kern.get("flow_backward_kernel")
    out = {
        "metric": METRIC, "value": t_flow, "unit": "s (GPU wall-time to rel err %g, learned map)" % a.rtol,
        "n_gpus": 1, "steps": 1, "warmup": nwarm, "ms_per_step": t_flow * 1e3, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "bf16 (flow MLP), f32 (integrand), f64 (sums)",
        "data": "synthetic (device Philox samples)",
        "config": {"workload": "%s: %d diagonal Gaussians sigma %g, [0,1]^%d; 6-layer affine coupling "
                               "flow, MLP %d-128-128 per layer, trained on 2e7 vegas samples in 2.5e5 batches"
                               % (wl["name"], wl["ncomp"], wl["sigma"], D, D),
                   "parallelism": "single GPU"},
        "result": {"flow": {"mean": est.mean, "sdev": est.sdev, "samples": int(nflow), "train_s": t_train,
                            "variance_loss_steps": int(fm.history[-1][0]) if fm.history else 200,
                            "reached_target": bool(flow_ok)},
                   "vegas": {"mean": r.mean, "sdev": r.sdev, "iterations": nit, "seconds": t_vegas,
                             "chi2_dof": r.chi2 / max(r.dof, 1), "reached_target": bool(vegas_ok)},
                   "exact": exact,
                   "speedup_vs_vegas": (t_vegas / t_flow) if vegas_ok else ">= %.1f (vegas did not reach the "
                                                                          "target in 3x the flow's time)" % (
                                                                              t_vegas / t_flow)},
        "roofline": {"bound": "mfma", "achieved": bwd["TFLOPs"] if bwd else None, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": (bwd["TFLOPs"] / PEAK_BF16_TFLOPS) if bwd else None,
                     "traffic": None, "kernel": "flow_backward_kernel", "kernel_ms_avg": bwd["ms_avg"] if bwd else None,
                     "kernels": kern,
                     "note": "BF16 MFMA kernels of a training step on 2.5e5 samples per layer; algorithmic "
                             "flops 2(D*128 + 128*128 + 128*2D) per sample and layer for the forward MLP "
                             "(backward kernel, which also forms the weight gradients: three times that); peak is the "
                             "BF16 dense MFMA peak"},
        "cpu_baseline": None if a.no_cpu else cpu_oracle_rate(
            (workloads.ridge_8d if D == 8 else workloads.ridge_9d)(diagonal=False)["f"], D, 1e5, 1,
            "the %s (per-evaluation rate: plain vegas+ does not reach 1e-4 on this integrand in the "
            "GPU comparison either, see result.vegas)" % wl["name"]),
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
    if a.workload in ("flow_ridge_9d", "flow_ridge_8d"):
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
    rows_ms = kernel_ms("train_rows_kernel")
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
                     "traffic": traffic, "kernel": "fused_kernel<8, true, GaussMix32>",
                     "kernel_ms_avg": avg_ms, "flops_per_eval": fpe,
                     "hbm_GBps": (traffic / (avg_ms * 1e-3) / 1e9) if (traffic and kern_ms) else None,
                     "hbm_frac": (traffic / (avg_ms * 1e-3) / 8.0e12) if (traffic and kern_ms) else None,
                     "train_rows_ms_avg": (sum(rows_ms) / len(rows_ms)) if rows_ms else None,
                     "train_rows_GBps": (avg_evals * TRAIN_REC_BYTES_8D / (sum(rows_ms) / len(rows_ms) * 1e-3) / 1e9)
                                        if rows_ms else None,
                     "note": "VALU-bound FP32 kernel (no MFMA in it); peak is the FP32 vector peak; "
                             "hbm_GBps = measured HBM bytes per launch / launch time, against 8 TB/s; "
                             "train_rows: the training histogram built from the sampler's records "
                             "(24 B per evaluation read)"},
    }
    # the same iterations with the unordered training histogram (LDS atomics
    # in timing order, Integrator(deterministic=False)): not part of value
    if a.no_unordered:
        out["cpu_baseline"] = None if a.no_cpu else cpu_baseline(wl, a.cpu_iters)
        print(json.dumps(out))
        return
    ui = va.Integrator([[0, 1]] * wl["dim"], neval=wl["neval"], seed=2026, precision=wl["precision"],
                       deterministic=False)
    if a.warmup > 0:
        ui(f, nitn=a.warmup)
    torch.cuda.synchronize()
    L.vgx_set_timing(cx, 1)
    t0 = time.perf_counter()
    ru = ui(f, nitn=a.steps)
    torch.cuda.synchronize()
    dtu = time.perf_counter() - t0
    ums = kernel_ms("fused_kernel")
    L.vgx_set_timing(cx, 0)
    out["result"]["unordered_histogram"] = {
        "evals_per_s": float(ru.sum_neval) / dtu, "ms_per_step": dtu * 1e3 / a.steps,
        "fused_kernel_ms_avg": (sum(ums) / len(ums)) if ums else None, "mean": ru.mean, "sdev": ru.sdev,
        "note": "Integrator(deterministic=False): training histogram by LDS atomics in timing order"}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_iters)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
