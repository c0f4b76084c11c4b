#!/usr/bin/env python3
"""Benchmark: 2D2V two-stream instability at 128^4, float64, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is one Strang-split Vlasov-Poisson step (X, Y drifts dt/2, rho,
FFT Poisson, Vx, Vy kicks dt, Y, X drifts dt/2) over the whole
f[128,128,128,128] (2.15 GB, resident in HBM).  Prints ONE JSON line:
value = phase-space cell-updates/s, plus the HBM roofline of the dominant
kernel (timed with events on the stream it runs on) and a CPU baseline
(a PyTorch-CPU restatement of the reference step on a bounded sample of the
same workload, all usable cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "jax-pytorch-vlasov_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
N = 128
L = 10.0 * math.pi      # k = 0.2
VMAX = 8.0
DT = 0.05


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N, help="points per axis (default 128)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-storage comparison")
    ap.add_argument("--fp32-error-steps", type=int, default=300,
                    help="steps of the FP32-storage vs float64 error comparison (default 300)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 512-instance 1D1V ensemble")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-surrogate", action="store_true", help="skip training the Fourier-operator surrogate")
    ap.add_argument("--no-reference-configs", action="store_true",
                    help="skip the per-step times on the configurations the reference published")
    ap.add_argument("--surrogate-iters", type=int, default=18000, help="training iterations per member network")
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="write profiles/traffic_2d2v_fp64.json from the four rocprofv3 --pmc runs under DIR "
                         "and exit (no GPU needed)")
    ap.add_argument("--only", choices=("ensemble", "surrogate"),
                    help="run one side workload alone (for profiling) and print its record")
    ap.add_argument("--stream-only", action="store_true",
                    help="only run the HBM copy kernel on two f-sized arrays (counter calibration) and exit")
    return ap.parse_args()


def stream_copy_gbs(n: int, reps: int = 10) -> float:
    """Measured HBM copy bandwidth (GB/s, read + write bytes) of the library's
    STREAM copy kernel on two float64 arrays of n^4 elements, in its fastest
    measured form: one 16-byte element per thread, one workgroup per 256
    elements (no grid-stride loop), non-temporal loads and stores."""
    import torch

    from vlasov_amd import _lib

    lib = _lib.lib()
    cells = n ** 4
    a = torch.ones(cells, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    s = torch.cuda.current_stream().cuda_stream
    grid = cells // 2 // 256

    def copy():
        _lib.check(lib.vp_stream_f64(0, a.data_ptr(), b.data_ptr(), cells, 1, 1, grid, s), "vp_stream_f64")

    copy()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        copy()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * 8 * cells / (ms * 1e-3) / 1e9


def fp32_storage(grid, f0, steps: int, warmup: int, err_steps: int) -> dict:
    """FP32-storage variant on the same workload: speed (wall time per step,
    no per-kernel events) and the error of f, phi, Ex, Ey and the
    field-energy history after `err_steps` steps against the float64 path
    (which matches the float64 oracle to <= 1e-12 in the parity tests).
    Two storage forms: "plain" (float32 f) and "deviation" (float32
    g = f - feq, feq = fx(vx) fy(vy) the separable part of the (x, y) mean of
    f0 with float64 profiles: rounding scales with the perturbation g
    instead of f).  The record's top level reports the form that meets the
    target (speed-up >= 1.8, masked pointwise field-energy error <= 1e-5)."""
    import torch

    from vlasov_amd import runner

    forms = {"fp64": (torch.float64, "plain"), "deviation": (torch.float64, "deviation"),
             "plain": (torch.float32, "plain")}

    def make(form):
        dtype, storage = forms[form]
        return runner.Run2D(grid, f0.to(dtype), DT, physics_mode=True, storage=storage)

    wall = {}
    for form in forms:
        run = make(form)
        run.advance(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.advance(steps)
        torch.cuda.synchronize()
        wall[form] = (time.perf_counter() - t0) / steps
        del run
        torch.cuda.empty_cache()
    hist = {}
    for form in forms:
        run = make(form)
        energy = torch.zeros(err_steps, dtype=torch.float64, device=f0.device)
        run.advance(err_steps, energy=energy)
        torch.cuda.synchronize()
        hist[form] = (run.state_f64(), run.phi.clone(), run.Ex.clone(), run.Ey.clone(), energy)
        del run
        torch.cuda.empty_cache()
    a = hist["fp64"]

    def rel(x, y):
        return float((y.double() - x).abs().max() / x.abs().max())

    def errors(form):
        b = hist[form]
        w64, w32 = a[4], b[4]
        rel_w = (w32 - w64).abs() / w64.abs()
        mask = w64.abs() > 1e-6 * w64.abs().max()  # pointwise error where W_E is not near zero
        return {
            "max_rel_err_f": rel(a[0], b[0]),
            "max_rel_err_phi": rel(a[1], b[1]),
            "max_rel_err_Ex": rel(a[2], b[2]),
            "max_rel_err_Ey": rel(a[3], b[3]),
            "max_rel_err_field_energy": float(((w32 - w64).abs() / w64.abs()).max()),
            "max_err_field_energy_over_max": float((w32 - w64).abs().max() / w64.abs().max()),
            "steps_within_1e-5_field_energy": _steps_within(rel_w, 1e-5),
            "max_rel_err_field_energy_masked": float(rel_w[mask].max()),
            "field_energy_mask": "steps where W_E(fp64) > 1e-6 max W_E(fp64)",
            "steps_masked": int(mask.sum()),
        }

    plain = {"storage": "float32 f (weights rounded to float32 with a unit sum); rho, Poisson, fields in float64",
             "ms_per_step": wall["plain"] * 1e3, "speedup_vs_fp64": wall["fp64"] / wall["plain"], **errors("plain")}
    dev = {"storage": "float32 g = f - feq(vx, vy), feq = fx(vx) fy(vy) the separable part of the (x, y) mean of "
                      "f0 (float64 profiles); drifts and kick of g in float32, kick correction (Kvx fx)(Kvy fy) - "
                      "fx fy in float64",
           "ms_per_step": wall["deviation"] * 1e3, "speedup_vs_fp64": wall["fp64"] / wall["deviation"],
           **errors("deviation")}
    for form in (plain, dev):
        form["meets_target"] = bool(form["speedup_vs_fp64"] >= 1.8 and form["max_rel_err_field_energy_masked"] <= 1e-5)
    # the reported form: one that meets the target (the faster if both do), else the more accurate
    best = max((f for f in (dev, plain) if f["meets_target"]), key=lambda f: f["speedup_vs_fp64"], default=dev)
    name = "deviation" if best is dev else "plain"
    res = {
        "form": name,
        "target": "speed-up >= 1.8 over the FP64 step with <= 1e-5 field-energy error, pointwise at every step "
                  "where W_E(fp64) > 1e-6 max W_E(fp64), over the error steps",
        "ms_per_step": best["ms_per_step"],
        "value": grid.Nx * grid.Ny * grid.Nvx * grid.Nvy / (best["ms_per_step"] * 1e-3),
        "fp64_ms_per_step_same_timing": wall["fp64"] * 1e3,
        "speedup_vs_fp64": best["speedup_vs_fp64"],
        "max_rel_err_field_energy_masked": best["max_rel_err_field_energy_masked"],
        "meets_target": best["meets_target"],
        "error_steps": err_steps,
        "error_reference": "float64 path of this library, same initial f, same steps",
        "field_energy_last_fp64": float(a[4][-1]),
        "forms": {"deviation": dev, "plain": plain},
    }
    return res


def ensemble_bench(steps: int, n: int = 512, seed: int = 0, sample: int = 8, sample_steps: int = 10) -> dict:
    """Batched 1D1V ensemble: n seeded instances (configs.ensemble_specs),
    dt = 0.1, `steps` steps, launch shapes autotuned, run from one captured
    HIP graph.  Baseline: the plain per-instance loop (one single-instance
    library step per instance per step) timed on `sample` instances spread
    over the size range and extrapolated to the whole ensemble."""
    import torch

    from vlasov_amd import configs
    from vlasov_amd import vlasov1D1V as v1
    from vlasov_amd.ensemble import Ensemble

    specs = configs.ensemble_specs(n, seed)
    ens = Ensemble(specs, 0.1, physics_mode=True)
    choice = ens.autotune()
    ens.prepare_graph(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    energy = ens.run(steps, graph=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    finite = bool(torch.isfinite(energy).all())
    # plain per-instance loop on a sample, extrapolated per instance-step
    order = sorted(range(n), key=lambda i: specs[i]["nx"] * specs[i]["nv"])
    pick = [order[(k * (n - 1)) // max(sample - 1, 1)] for k in range(sample)]
    per_step = []
    for i in pick:
        g = ens.grids[i]
        f = configs.ensemble_f0(g, specs[i]["k"], specs[i]["amplitude"], specs[i]["drift"])
        f, *_ = v1.step_with_field(f, 0.1, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, True)  # warm
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(sample_steps):
            f, *_ = v1.step_with_field(f, 0.1, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, True)
        torch.cuda.synchronize()
        per_step.append((time.perf_counter() - t1) / sample_steps)
    base_t = float(np.mean(per_step)) * n * steps
    return {
        "metric": "ensemble solves/s (512 seeded 1D1V instances, 500 steps each, FP64)",
        "value": n / t,
        "unit": "solves/s",
        "seconds": t,
        "instances": n,
        "steps": steps,
        "cells": ens.cells,
        "cell_updates_per_s": ens.cells * steps / t,
        "energy_finite": finite,
        "launch_shapes": {k: choice[k] for k in ("ens_xtile", "ens_ktile", "ens_wgs")},
        "graph": "one HIP graph of the whole run (a work-list set-up launch, then 3 launches per step: "
                 "persistent X drift pair + rho, field solve, persistent kick)",
        "per_instance_loop": {"value": n / base_t, "unit": "solves/s",
                              "sample": f"{sample} instances across the size range x {sample_steps} steps, "
                                        "extrapolated per instance-step",
                              "ms_per_instance_step": float(np.mean(per_step)) * 1e3},
        "speedup_vs_per_instance_loop": base_t / t,
        "seed": seed,
    }


def surrogate_bench(iters: int, ntrain: int = 20000, ncal: int = 2000, ntest: int = 1000, steps: int = 500,
                    tol: float = 0.02, quantile: float = 0.95) -> dict:
    """Fourier-operator surrogate of log W_E(t): data from the batched solver
    (128 x 256 grid, 500 steps), two member networks trained on the GPU
    kernels (`iters` iterations each), routing calibrated on held-out solver
    runs, then evaluated on other held-out runs against the solver: accuracy
    of the mean prediction, and the routed query at `tol` (served fraction,
    error of the served answers, throughput against solving every query)."""
    import torch

    from vlasov_amd import surrogate as sg

    ptr, pcal, pte = sg.sample_params(ntrain, 1), sg.sample_params(ncal, 3), sg.sample_params(ntest, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Wtr, Mtr = sg.solve(ptr, steps, return_moments=True)
    Wcal, Mcal = sg.solve(pcal, steps, return_moments=True)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    Wte = sg.solve(pte, steps)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    sur = sg.Surrogate(steps, members=2)
    t0 = time.perf_counter()
    _progress(f"surrogate: data {t_data:.1f} s; training 2 x {iters} iterations")
    losses = sur.fit(ptr, Wtr, Mtr, iters=iters, uniform=0.1,
                     log=lambda x: _progress(f"surrogate: member {x[0]} iteration {x[1]} loss {x[2]:.3g}")
                     if x[1] % 3000 == 0 else None)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    sur.calibrate(pcal, Wcal, Mcal)
    sur.predict(pte[:8])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Wp = sur.predict(pte)
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    e = sg.rel_l2(Wp, Wte).cpu().numpy()
    elog = (torch.log(Wp) - torch.log(Wte)).abs().mean(dim=1).cpu().numpy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Wq, info = sur.query(pte, tol=tol, quantile=quantile)
    torch.cuda.synchronize()
    t_query = time.perf_counter() - t0
    served = info.pop("served_mask")
    eq = sg.rel_l2(Wq, Wte).cpu().numpy()
    es = eq[served]
    return {
        "model": "1D Fourier neural operator over time, 2 members: 4 layers, width 64, 32 modes, "
                 "BF16 MFMA / FP32 accumulate",
        "input": "(k, amplitude, drift, sigma) + f0 on the 128 x 256 grid, encoded by vp_nn_f0_moments "
                 "(density mode amplitude, <v^2>, <v^4>/<v^2>^2)",
        "loss": "log W_E error weighted 0.1 / T + 0.9 W_t^2 / sum W^2 per run (the relative L2 error to first order)",
        "train_runs": ntrain, "calibration_runs": ncal, "test_runs": ntest, "steps": steps, "grid": [128, 256],
        "data_seconds": t_data, "train_iters_per_member": iters, "train_seconds": t_train,
        "final_loss": [l[2] for l in losses if l[1] == iters - 1],
        "test_rel_l2_median": float(np.median(e)), "test_rel_l2_p90": float(np.percentile(e, 90)),
        "test_frac_within_2pct": float(np.mean(e <= 0.02)), "test_frac_within_10pct": float(np.mean(e <= 0.10)),
        "test_mean_abs_log_err_median": float(np.median(elog)),
        "predict_queries_per_s": ntest / t_pred, "solver_queries_per_s": ntest / t_solve,
        "speedup_vs_solver": t_solve / t_pred,
        "routing": {
            "rule": f"served when the members' disagreement <= the largest value at which <= {1 - quantile:.0%} of "
                    f"the {ncal} calibration runs miss the tolerance; else the solver; {5}% of served queries "
                    "judged by the solver",
            "tol": tol, "quantile": quantile, "threshold": info["threshold"],
            "served_frac": float(served.mean()),
            "served_frac_within_tol": float(np.mean(es <= tol)) if len(es) else None,
            "served_err_median": float(np.median(es)) if len(es) else None,
            "served_err_p95": float(np.percentile(es, 95)) if len(es) else None,
            "served_err_max": float(es.max()) if len(es) else None,
            "judged": info["judged"], "judged_misses": info["judged_misses"], "judge_max_err": info["judge_max_err"],
            "fallback": info["fallback"],
            "query_seconds": t_query, "solver_seconds": t_solve,
            "throughput_gain_vs_solver": t_solve / t_query,
        },
    }


# The reference's published per-step times for the configurations below
# (elapsed time / steps from the cited logs; Landau damping, dt 0.05, L 4 pi,
# V_max 5, eps 1e-3).
REFERENCE_TIMES = {
    "2d2v_128_fp64": [
        ("Kokkos s1", "H200", 17.16, "artifact/vlasov2d_2v/kokkos/d875542/vlp2d_2v_H200_fp64_Nx128_Nvx128_solver1.txt:29"),
        ("PyTorch s2", "H200", 17.23, "artifact/vlasov2d_2v/pytorch/d875542/vlp2d_2v_H200_float64_Nx128_Nvx128_solver2.txt:6"),
        ("Kokkos s1", "MI250X GCD", 48.15,
         "artifact/vlasov2d_2v/kokkos/d1904d7/vlp2d_2v_MI250X_fp64_Nx128_Nvx128_solver1.txt:96")],
    "2d2v_128_fp32": [
        ("PyTorch s2", "H200", 9.49, "artifact/vlasov2d_2v/pytorch/d875542/vlp2d_2v_H200_float32_Nx128_Nvx128_solver2.txt:6"),
        ("Kokkos s1", "MI250X GCD", 39.93,
         "artifact/vlasov2d_2v/kokkos/d1904d7/vlp2d_2v_MI250X_fp32_Nx128_Nvx128_solver1.txt:96")],
    "2d2v_64_fp64": [
        ("PyTorch s2", "H200", 0.841, "artifact/vlasov2d_2v/pytorch/d875542/vlp2d_2v_H200_float64_Nx64_Nvx64_solver2.txt:6"),
        ("Kokkos s1", "MI250X GCD", 3.20, "artifact/vlasov2d_2v/kokkos/d1904d7/vlp2d_2v_MI250X_fp64_Nx64_Nvx64_solver1.txt:96")],
    "1d1v_128_fp64": [
        ("PyTorch s2", "H200", 0.363, "artifact/vlasov1d_1v/pytorch/d875542/vlp1d_1v_H200_float64_N128_solver2.txt:6"),
        ("Kokkos s1", "MI250X GCD", 0.135, "artifact/vlasov1d_1v/kokkos/d1904d7/vlp1d_1v_MI250X_fp64_N128_solver1.txt:96")],
    "1d1v_4096_fp64": [
        ("PyTorch s2", "H200", 0.518, "artifact/vlasov1d_1v/pytorch/d875542/vlp1d_1v_H200_float64_N4096_solver2.txt:6"),
        ("Kokkos s1", "MI250X GCD", 1.379,
         "artifact/vlasov1d_1v/kokkos/d1904d7/vlp1d_1v_MI250X_fp64_N4096_solver1.txt:96")],
}


def reference_configs(ms_2d_fp64: float, ms_2d_fp32) -> dict:
    """This library's per-step time on the configurations the reference
    published, next to those published times (other hardware).  2D 128^4
    comes from the main and FP32 legs; 2D 64^4 runs the fused runner, 1D1V
    the drop-in steps from Python, one call per step, as the reference's own
    loop does: solver 0 (vlasov1D1V.step_with_field) as `ms_per_step`, and
    solvers 1 and 2, the ones the reference's times were measured with, in
    `ms_per_step_by_solver` / `speedup_same_solver`."""
    import torch

    from vlasov_amd import configs, runner
    from vlasov_amd import vlasov1D1V as v1
    from vlasov_amd import vlasov2D2V as v2

    def timed(fn, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(steps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    ours = {"2d2v_128_fp64": ms_2d_fp64, "2d2v_128_fp32": ms_2d_fp32}
    g = v2.Grid.create(64, 64, 64, 64, 4 * math.pi, 4 * math.pi, 5.0, 5.0, device="cuda")
    run = runner.Run2D(g, configs.landau_2d(g, eps=1e-3), 0.05, True)
    run.advance(5)
    ours["2d2v_64_fp64"] = timed(run.advance, 128)
    del run
    by_solver = {}
    for n, steps in ((128, 1000), (4096, 50)):
        g1 = v1.Grid.create(n, n, 4 * math.pi, 5.0, device="cuda")
        state = {"f": configs.landau_1d(g1, eps=1e-3, k=0.5)}
        args = (g1.v, g1.inv_kx, g1.inv_kx_sq, g1.x_indices, g1.v_indices, g1.x_batch_offsets,
                g1.v_batch_offsets, g1.stencil_offsets, g1.dx, g1.dv, g1.V_max, True)
        blocks, W_x = v1.precompute_x_drift(g1.v, 0.025, g1.dx, torch.float64)
        so2 = v1.static_stencil_offsets(g1.stencil_offsets)

        def steps_s0(k):
            f = state["f"]
            for _ in range(k):
                f, _, _, _ = v1.step_with_field(f, 0.05, g1.v, g1.inv_kx, g1.inv_kx_sq, g1.dx, g1.dv, True)
            state["f"] = f

        def steps_s1(k):
            f = state["f"]
            for _ in range(k):
                f, _, _ = v1.vlasov_step_matrix_diag(f, 0.05, *args)
            state["f"] = f

        def steps_s2(k):
            # the driver's solver-2 loop: tables precomputed once, the core step per iteration
            f = state["f"]
            for _ in range(k):
                f, _, _ = v1.vlasov_step_gather_free_impl(f, 0.05, W_x, blocks, so2, g1.inv_kx, g1.inv_kx_sq, g1.dv,
                                                          True, 1)
            state["f"] = f

        times = {}
        for sid, fn in ((0, steps_s0), (1, steps_s1), (2, steps_s2)):
            fn(5)
            times[sid] = timed(fn, steps)
        ours[f"1d1v_{n}_fp64"] = times[0]
        by_solver[f"1d1v_{n}_fp64"] = {f"solver{k}": v for k, v in times.items()}
    out = {}
    for key, refs in REFERENCE_TIMES.items():
        if ours.get(key) is None:
            continue
        best = min(r[2] for r in refs)
        out[key] = {"ms_per_step": ours[key],
                    "reference": [{"impl": a, "hardware": b, "ms_per_step": c, "source": d} for a, b, c, d in refs],
                    "speedup_vs_fastest_reference": best / ours[key]}
        if key in by_solver:
            # the reference's times are its solver 1 (Kokkos) and solver 2 (PyTorch) steps
            out[key]["ms_per_step_by_solver"] = by_solver[key]
            out[key]["speedup_same_solver"] = {
                f"{a} {b}": c / by_solver[key][f"solver{1 if 's1' in a else 2}"] for a, b, c, _ in refs}
    return out


def _progress(msg: str) -> None:
    """Progress line on stderr (stdout carries only the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _steps_within(rel_err, tol: float) -> int:
    """Number of leading steps whose relative error stays <= tol."""
    bad = (rel_err > tol).nonzero()
    return int(bad[0, 0]) if bad.numel() else int(rel_err.numel())


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic_2d2v_fp64.json")

# Kernel instantiations the fused float64 schedule launches at 128^4 with the
# default launch knobs, by bench key (rocprofv3 kernel names without arguments).
FP64_KERNELS = {
    "x_drift": "void vp::drift_dma_kernel<double, 4, 1, false, false, 2, true, 1>",
    "xx_drift": "void vp::drift_dma_kernel<double, 4, 2, false, false, 2, true, 1>",
    "y_drift_rho": "void vp::drift_dma_kernel<double, 4, 1, true, true, 2, true, 1>",
    "y_drift": "void vp::drift_dma_kernel<double, 4, 1, false, true, 2, true, 1>",
    "kick": "void vp::kick_rows_kernel<double, 128, 8, 4, true, false, 128>",
}


# Sources the profiled kernels (the fused float64 sweeps and the copy
# calibration kernel) are built from, launch knobs included.
PROFILED_SOURCES = ("jax-pytorch-vlasov_amd/csrc/fast2d.hip", "jax-pytorch-vlasov_amd/csrc/fast2d.hpp",
                    "jax-pytorch-vlasov_amd/csrc/vp_common.hpp", "jax-pytorch-vlasov_amd/csrc/stream.hip",
                    "include/vlasov_amd.h")


def kernel_sources_sha() -> str:
    """sha256 over the sources of the profiled kernels: a counter profile is
    valid only for the sources it was taken from."""
    import hashlib

    h = hashlib.sha256()
    for path in sorted(os.path.join(ROOT, p) for p in PROFILED_SOURCES):
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def traffic_from_profile(kernel: str, n: int):
    """(HBM bytes per launch of `kernel` at n^4, note) from the committed
    counter summary (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by the copy
    kernel's calibration).  None when no summary covers this kernel and size,
    or when the summary was taken from other kernel sources than these."""
    try:
        with open(TRAFFIC_JSON) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None, "no counter summary"
    if data.get("n") != n:
        return None, f"counter summary is for n = {data.get('n')}"
    if data.get("kernel_sources_sha256") != kernel_sources_sha():
        return None, "counter summary was taken from other kernel sources (stale); not reported"
    k = data.get("kernels", {}).get(kernel)
    if k is None:
        return None, f"counter summary has no {kernel}"
    return k.get("hbm_bytes"), f"rocprofv3 FETCH_SIZE + WRITE_SIZE of {k.get('kernel')}, " \
                               f"kernel sources sha256 {data['kernel_sources_sha256'][:12]}"


def summarize_counters(root_dir: str, n: int) -> dict:
    """Build the counter summary from four rocprofv3 --pmc runs (csv) under
    root_dir: pmc_FETCH_SIZE, pmc_WRITE_SIZE (the bench command) and
    pmc_copy_FETCH_SIZE, pmc_copy_WRITE_SIZE (``--stream-only``).  Counters
    are KiB per dispatch; each is scaled by what it reads for the copy
    kernel, which moves exactly 8 n^4 bytes each way."""
    import csv
    import glob

    def load(sub, counter):
        rows = {}
        for path in glob.glob(os.path.join(root_dir, sub, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as fh:
                for r in csv.DictReader(fh):
                    if r.get("Counter_Name") != counter:
                        continue
                    rows.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
        return rows

    copy_bytes = 8.0 * n ** 4
    fetch, write = load("pmc_FETCH_SIZE", "FETCH_SIZE"), load("pmc_WRITE_SIZE", "WRITE_SIZE")
    cf, cw = load("pmc_copy_FETCH_SIZE", "FETCH_SIZE"), load("pmc_copy_WRITE_SIZE", "WRITE_SIZE")

    def copy_mean(rows):
        vals = [v for k, lst in rows.items() if "stream_copy_kernel" in k for v in lst]
        return float(np.mean(vals))

    ff, wf = copy_bytes / copy_mean(cf), copy_bytes / copy_mean(cw)
    names = FP64_KERNELS
    try:
        import subprocess
        commit = subprocess.run(["git", "-C", ROOT, "rev-parse", "--short", "HEAD"], capture_output=True,
                                text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        commit = None
    out = {"n": n, "kernel_sources_sha256": kernel_sources_sha(), "kernel_sources_commit": commit,
           "fetch_factor": ff, "write_factor": wf,
           "algorithmic_bytes_per_launch": 16 * n ** 4, "kernels": {}}
    for key, kname in names.items():
        fr = [v for k, lst in fetch.items() if k.startswith(kname + "(") or k == kname for v in lst]
        wr = [v for k, lst in write.items() if k.startswith(kname + "(") or k == kname for v in lst]
        if not fr or not wr:
            continue
        fb, wb = float(np.mean(fr)) * ff, float(np.mean(wr)) * wf
        out["kernels"][key] = {"kernel": kname, "fetch_bytes": fb, "write_bytes": wb, "hbm_bytes": fb + wb,
                               "calls": len(fr)}
    return out


def cpu_baseline(n: int, frac: int = 8) -> dict:
    """The reference's CPU path, restated in PyTorch-CPU eager float64 in its
    op order (oracle/torch_cpu.py: solver 0, vlasov2D2V.py:943-998 and
    vlasov1D1V.py:574-622), on every CPU this process may use (capped by
    OMP_NUM_THREADS, the CPU share of the machine).

    * 2D2V n^4 (the bench workload): one step on a bounded sample, every
      sweep on 1/frac of its lines at full line length (X, Y drifts and the
      Vx kick on f[:, :, :, :n/frac], the Vy kick and rho on f[:n/frac]),
      the Poisson solve on the full [n, n] grid; the step time is the sum
      with the sweeps scaled by frac.
    * 1D1V 128^2 Landau damping, 600 steps (the 1D1V reference
      configuration), timed in full after one warm-up step.
    """
    import torch

    from oracle import torch_cpu as tc
    from oracle import vlasov_oracle as vo

    threads = tc.use_all_cores()
    T = torch.from_numpy
    # 1D1V, 600 steps in full
    g1 = vo.grid_1d(128, 128, 4 * math.pi, 6.0)
    f = T((1 + 0.01 * np.cos(0.5 * g1.x))[:, None] * np.exp(-0.5 * g1.v ** 2)[None, :] / math.sqrt(2 * math.pi))
    a1 = (T(g1.v), T(g1.inv_kx), T(g1.inv_kx_sq), g1.dx, g1.dv)
    f, *_ = tc.vlasov_step_1d(f, 0.1, *a1)
    t0 = time.perf_counter()
    for _ in range(600):
        f, *_ = tc.vlasov_step_1d(f, 0.1, *a1)
    t_1d = time.perf_counter() - t0
    # 2D2V n^4 sample
    g = vo.grid_2d(n, n, n, n, L, L, VMAX, VMAX)
    m = max(n // frac, 1)
    gen = torch.Generator().manual_seed(0)
    a = torch.rand((n, n, n, m), dtype=torch.float64, generator=gen)
    b = torch.rand((m, n, n, n), dtype=torch.float64, generator=gen)
    E = torch.randn((n, n), dtype=torch.float64, generator=gen) * 0.01
    vx, vy = T(g.vx), T(g.vy)
    t = {}

    def timed(key, fn):
        t0 = time.perf_counter()
        fn()
        t[key] = time.perf_counter() - t0

    tc.advect_axis(a[:, :, :, :1], vx[None, :, None], 0, DT / 2, g.dx, True)  # warm-up
    timed("x", lambda: tc.advect_axis(a, vx[None, :, None], 0, DT / 2, g.dx, True))
    timed("y", lambda: tc.advect_axis(a, vy[None, None, :m], 1, DT / 2, g.dy, True))
    timed("vx", lambda: tc.advect_axis(a, E[:, :, None], 2, DT, g.dvx, False))
    timed("vy", lambda: tc.advect_axis(b, E[:m, :, None], 3, DT, g.dvy, False))
    timed("rho", lambda: torch.sum(b, dim=(2, 3)) * g.dvx * g.dvy)
    rho = 1.0 + 0.001 * torch.randn((n, n), dtype=torch.float64, generator=gen)
    timed("poisson", lambda: tc.compute_poisson_real(rho, T(g.kx), T(g.ky), T(g.inv_ksq)))
    sample_s = sum(t.values())
    step_s = (n / m) * (2 * t["x"] + 2 * t["y"] + t["vx"] + t["vy"] + t["rho"]) + t["poisson"]
    return {
        "value": n ** 4 / step_s,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"PyTorch-CPU eager float64 restatement of the reference solver-0 step (oracle/torch_cpu.py), "
                   f"{threads} threads: every sweep of one {n}^4 step on 1/{n // m} of its lines at full line "
                   f"length, full-size rho and Poisson; {sample_s:.1f} s of CPU work, scaled to one step "
                   f"({step_s:.1f} s/step)"),
        "cpu": _cpu_model(),
        "ms_per_step": step_s * 1e3,
        "landau_1d1v_128": {"steps": 600, "seconds": t_1d, "ms_per_step": t_1d / 600 * 1e3,
                            "value": 128 * 128 * 600 / t_1d, "unit": "cell-updates/s",
                            "sample": "the full 600-step 1D1V Landau run (nx = nv = 128, dt = 0.1) after one "
                                      "warm-up step"},
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the solver runs on a single "
              "MI355X and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.summarize_counters:
        data = summarize_counters(args.summarize_counters, args.n)
        with open(TRAFFIC_JSON, "w") as fh:
            json.dump(data, fh, indent=1)
        print(json.dumps(data))
        return
    import torch

    from vlasov_amd import configs, runner
    from vlasov_amd import vlasov2D2V as v2

    n = args.n
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_bench(args.ensemble_steps)}), flush=True)
        return
    if args.only == "surrogate":
        print(json.dumps({"surrogate": surrogate_bench(args.surrogate_iters)}), flush=True)
        return
    copy_gbs = stream_copy_gbs(n)
    if args.stream_only:
        print(json.dumps({"stream_copy_GBps": copy_gbs, "bytes_per_launch": 2 * 8 * n ** 4}), flush=True)
        return
    grid = v2.Grid.create(n, n, n, n, L, L, VMAX, VMAX, device=dev, dtype=torch.float64)
    f0 = configs.two_stream_2d(grid)
    run = runner.Run2D(grid, f0, DT, physics_mode=True)
    run.advance(args.warmup)
    torch.cuda.synchronize()

    names = runner.FUSED_KERNELS if run.fused else runner.GENERIC_KERNELS
    energy = torch.zeros(max(args.steps, 1), dtype=torch.float64, device=dev)
    # Per-kernel times from an untimed pass of the same steps with HIP events
    # around every launch: events between all the launches cost ~1 % of the
    # step (35-55 us per 128^4 step), so the timed pass records them around
    # the dominant kernel only.
    probe = {k: [] for k in names}
    run.advance(args.steps, events=probe, energy=energy)
    torch.cuda.synchronize()
    ms = {k: sum(e0.elapsed_time(e1) for e0, e1 in v) / len(v) for k, v in probe.items() if v}
    calls = {k: len(v) for k, v in probe.items() if v}
    sweep_keys = [k for k in ms if k not in ("rho", "poisson", "field_energy")]
    # dominant kernel = the sweep with the most total time over the steps
    dom = max(sweep_keys, key=lambda k: ms[k] * calls[k])

    events = {dom: []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.advance(args.steps, events=events, energy=energy)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    cells = n ** 4
    t_step = elapsed / args.steps
    dom_ms = sum(e0.elapsed_time(e1) for e0, e1 in events[dom]) / len(events[dom])  # live, timed region
    bytes_per_launch = 2 * 8 * cells  # one read + one write of f per pass
    achieved = bytes_per_launch / (dom_ms * 1e-3) / 1e9
    # the step's own algorithmic traffic: every sweep pass reads and writes f once
    passes = sum(calls[k] for k in sweep_keys) / args.steps
    step_bytes = passes * bytes_per_launch
    model_gbs = 13 * 8 * cells / t_step / 1e9  # the reference's model: 13 f-accesses per step
    traffic, traffic_note = traffic_from_profile(dom, n)

    out = {
        "metric": "phase-space cell-updates/s (2D2V 128^4 two-stream, FP64)",
        "value": cells / t_step,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        # fastest published 2D2V 128^4 FP64 rate: Kokkos s1 on H200, 17.16 ms/step
        "vs_baseline": (cells / t_step) / (cells / 17.16e-3) if n == 128 else None,
        "vs_baseline_ref": "2D2V 128^4 FP64 cell-updates/s of the reference, Kokkos s1 on H200 (17.16 ms/step, "
                           "artifact/vlasov2d_2v/kokkos/d875542/vlp2d_2v_H200_fp64_Nx128_Nvx128_solver1.txt:29)",
        "dtype": "f64",
        "data": "synthetic (closed-form two-stream initial condition)",
        "config": {"workload": f"2D2V two-stream instability {n}^4 FP64, k=0.2, v in [-8,8], dt=0.05",
                   "shape": [n, n, n, n], "parallelism": "single GPU",
                   "path": "fused" if run.fused else "generic"},
        "step_traffic": {"passes_per_step": passes, "algorithmic_bytes": step_bytes,
                         "GBps": step_bytes / t_step / 1e9, "frac_hbm_peak": step_bytes / t_step / 1e9 / HBM_PEAK_GBS,
                         "note": "every sweep pass reads and writes f once (16 B per cell); the Poisson solve "
                                 "and field energy move no f"},
        "reference_model": {"GBps_13access": model_gbs,
                            "note": "the reference's traffic model (13 f-accesses per step, "
                                    "python_scripts/plot/src/parse_cross_platform_vlp2d.py:15-20), for comparison "
                                    "with its published GB/s; the fused schedule makes fewer accesses, so this "
                                    "is not a fraction of the HBM peak"},
        "kernel_ms": ms,
        "kernel_calls": calls,
        "kernel_ms_source": "HIP events around every launch in an untimed pass of the same steps (the timed "
                            "pass records events around the dominant kernel only)",
        "field_energy_last": float(energy[args.steps - 1].item()),
        "roofline": {"bound": "hbm", "kernel": dom, "kernel_ms_timed": dom_ms,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "measured_copy_GBps": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs},
    }
    del run
    torch.cuda.empty_cache()
    if not args.no_fp32 and runner.fused_supported(f0.shape, torch.float32):
        _progress("fp32 storage")
        out["fp32_storage"] = fp32_storage(grid, f0, args.steps, args.warmup, args.fp32_error_steps)
    if not args.no_ensemble:
        _progress("ensemble")
        out["ensemble"] = ensemble_bench(args.ensemble_steps)
    if not args.no_reference_configs:
        _progress("reference configurations")
        fp32_ms = out.get("fp32_storage", {}).get("ms_per_step")
        out["reference_configs"] = reference_configs(t_step * 1e3, fp32_ms)
    if not args.no_surrogate:
        _progress("surrogate")
        out["surrogate"] = surrogate_bench(args.surrogate_iters)
    if not args.no_cpu_baseline:
        _progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
