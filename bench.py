#!/usr/bin/env python3
"""Benchmark of the CCTpfa 2p Newton path on one MI355X.

Workload (2p DNAPL infiltration, 64^3): 2p immiscible (pw-Sn) on a 64^3 grid of a 1 m box
(262 144 cells, 2x2 blocks), porosity 0.2, log10 K ~ N(-11, 0.5) from a seed, van Genuchten
alpha = 4e-4, n = 2.5, Swr = 0.05, hydrostatic initial pw with Sn = 0, DNAPL influx
0.04 kg/(m^2 s) on the central quarter of the top face, hydrostatic Dirichlet sides,
10 implicit-Euler steps of dt = 10 s, numeric (forward-difference) Jacobian, linear systems to
1e-10 relative residual.

Forms, all on the same GPU, inputs resident in HBM when a timed region starts:
  batched : --batch problems of that workload (seeds 2024, 2025, ...) in one batched executor,
            FP32 BiCGSTAB + block Jacobi inside FP64 iterative refinement -> `value`
  single  : one problem, the same mixed-precision solver with block Jacobi (fused kernels in a HIP graph)
  single multigrid : one problem, FP64 BiCGSTAB preconditioned by the aggregation multigrid
  plain   : one problem at a time, FP64 BiCGSTAB + ILU(0) (the reference's solver arithmetic)
Extra sections: an ensemble of many problems of mixed kinds and sizes in one batch, and a
served stream of requests with Poisson arrivals (p50/p99 latency against throughput).
Metric: Newton cell-steps/s = sum over problems of cells x Newton iterations / wall seconds.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

PEAK_HBM_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per axis")
    ap.add_argument("--tsteps", type=int, default=10, help="implicit Euler steps per solve")
    ap.add_argument("--dt", type=float, default=10.0)
    ap.add_argument("--batch", type=int, default=8, help="problems of the workload solved together")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU form")
    ap.add_argument("--no-single", action="store_true", help="skip the single-problem mixed form")
    ap.add_argument("--no-single-mg", action="store_true", help="skip the single-problem multigrid form")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the mixed-size ensemble")
    ap.add_argument("--no-serve", action="store_true", help="skip the served stream")
    ap.add_argument("--ensemble-count", type=int, default=256)
    ap.add_argument("--serve-rates", default="2,20,200")
    ap.add_argument("--serve-count", type=int, default=16)
    ap.add_argument("--hbm-budget-gb", type=float, default=96.0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape of the batched kernels")
    ap.add_argument("--batch-precond", choices=("amg", "jacobi"), default="amg",
                    help="preconditioner of the batched FP32 BiCGSTAB: aggregation multigrid or block Jacobi")
    ap.add_argument("--probe", action="store_true",
                    help="hardware-counter workload: one implicit-Euler step of the batched form, then 10 timed "
                         "launches of the batched SpMV + dot and a 512 MiB device copy (reference for the "
                         "byte counters); run under rocprofv3 --pmc")
    return ap.parse_args()


MIXED = dict(reduction=1e-10, inner_reduction=1e-4, max_iter=2000, max_refinements=30)


def workload_problem(dx, n, seed):
    return dx.Problem(kind=2, dim=3, n=(n, n, n), seed=seed, pid=0, mu_log10K=-11.0, sigma_log10K=0.5)


def batch_linear(dx, args):
    """Linear-solver parameters of the batched executor (FP32 BiCGSTAB in FP64 refinement)."""
    return dict(MIXED, precond=dx.PRECOND_AMG if args.batch_precond == "amg" else dx.PRECOND_JACOBI)


def run_batched(dx, args):
    """B problems of the workload in one batch; returns (seconds per solve, cell-steps, infos,
    solution of problem 0, kernel timings)."""
    n, B = args.n, args.batch
    probs = [workload_problem(dx, n, 2024 + i) for i in range(B)]
    asms = [dx.FVAssembler(p) for p in probs]

    tuned = [None]

    def solve_once():
        for p, a in zip(probs, asms):
            a.setSolution(p.u0)
        b = dx.Batch(linear=batch_linear(dx, args))
        if tuned[0]:
            b.set_tuning(*tuned[0][:3])
        for a in asms:
            b.add(a, nsteps=args.tsteps, dt=args.dt)
        for a in asms:
            a.synchronize()
        t0 = time.perf_counter()
        b.run()
        el = time.perf_counter() - t0
        infos = [b.status(a)[1] for a in asms]
        if not all(b.status(a)[0] == 1 for a in asms):
            raise RuntimeError("batched solve did not converge")
        return el, infos, b

    for w in range(args.warmup):
        el, infos, b = solve_once()
        if w == 0 and not args.no_tune:
            # launch shape of the BiCGSTAB kernels for these members (timing only; results unchanged)
            tuned[0] = b.autotune()
            log(f"[bench] autotune: coarse {tuned[0][0]}, vector grid {tuned[0][1]}, spmv grid {tuned[0][2]}, "
                f"{tuned[0][3]:.1f} us per iteration")
        b.close()
        log(f"[bench] batched warmup {w}: {el:.3f} s, newton {[i.newton_steps for i in infos]}")
    tot, cs, all_infos = 0.0, 0, []
    for s in range(args.steps):
        el, infos, b = solve_once()
        tot += el
        cs += sum(p.N * i.newton_steps for p, i in zip(probs, infos))
        all_infos.append(infos)
        log(f"[bench] batched step {s}: {el:.3f} s, newton {[i.newton_steps for i in infos]}, "
            f"linear its {[i.linear_iterations for i in infos]}, fp64 fallback {[i.fp64_fallback for i in infos]}")
        if s + 1 < args.steps:
            b.close()
    u0 = asms[0].solution()
    # live timing of the batched kernels (HIP events on the batch stream), state of the last solve
    kernels = {}
    for name, kid in [("batched_assemble", dx.KERNEL_ASSEMBLE), ("batched_spmv_dot_fp32", dx.KERNEL_BJ_SPMV_DOT_FP32),
                      ("batched_spmv_dots_fp32", dx.KERNEL_BJ_SPMV_DOTS_FP32),
                      ("batched_iteration_fp32", dx.KERNEL_BJ_ITERATION_FP32),
                      ("batched_mg_down0", dx.KERNEL_MG_DOWN0), ("batched_mg_up0", dx.KERNEL_MG_UP0),
                      ("batched_refine_residual_fp64", dx.KERNEL_REFINE_RESIDUAL)]:
        try:
            ms, by = b.time_kernel(kid, reps=20)
        except Exception as e:
            log(f"[bench] kernel {name}: not timed ({e})")
            continue
        kernels[name] = {"ms": ms, "bytes": by, "GBps": by / (ms * 1e-3) / 1e9}
        log(f"[bench] kernel {name}: {ms * 1e3:.1f} us, {kernels[name]['GBps']:.0f} GB/s")
    b.close()
    for a in asms:
        a.close()
    kernels["launch_shape"] = {"coarse": tuned[0][0], "vector_grid": tuned[0][1], "spmv_grid": tuned[0][2],
                               "us_per_iteration": tuned[0][3]} if tuned[0] else None
    return tot / args.steps, cs / tot, all_infos, u0, kernels


def run_single(dx, args, linear, label, kernel_ids):
    prob = workload_problem(dx, args.n, 2024)
    asm = dx.FVAssembler(prob)

    def solve_once():
        asm.setSolution(prob.u0)
        asm.setPreviousSolution(prob.u0)
        nw = dx.NewtonSolver(asm, linear=linear)
        info = nw.time_loop(args.tsteps, args.dt)
        if not info.converged:
            raise RuntimeError(f"{label}: Newton time loop did not converge")
        return info

    for w in range(args.warmup):
        info = solve_once()
        log(f"[bench] {label} warmup {w}: newton {info.newton_steps} linear its {info.linear_iterations}")
    asm.synchronize()
    t0 = time.perf_counter()
    infos = []
    for s in range(args.steps):
        infos.append(solve_once())
        log(f"[bench] {label} step {s}: newton {infos[-1].newton_steps} linear its {infos[-1].linear_iterations} "
            f"refinements {infos[-1].refinements}")
    asm.synchronize()
    el = time.perf_counter() - t0
    u = asm.solution()
    kernels = {}
    asm.assembleJacobianAndResidual(u)
    for name, kid, reps in kernel_ids:
        try:
            ms, by = asm.time_kernel(kid, reps=reps)
        except Exception as e:  # a timing that stopped early is reported as missing
            log(f"[bench] kernel {name}: not timed ({e})")
            continue
        kernels[name] = {"ms": ms, "bytes": by, "GBps": by / (ms * 1e-3) / 1e9}
        log(f"[bench] kernel {name}: {ms * 1e3:.1f} us, {kernels[name]['GBps']:.0f} GB/s")
    asm.close()
    newton = sum(i.newton_steps for i in infos)
    return {"value": prob.N * newton / el, "ms_per_step": 1e3 * el / args.steps,
            "newton_steps_per_solve": newton / args.steps,
            "linear_iterations_per_solve": sum(i.linear_iterations for i in infos) / args.steps,
            "refinements_per_solve": sum(i.refinements for i in infos) / args.steps}, u, kernels


def run_ensemble(dx, args):
    """Many independent problems of mixed kinds and sizes in one batch: half OneP 2D (side
    log-uniform in [32, 512], one stationary solve), half TwoP 3D (side log-uniform in [24, 96],
    one implicit-Euler step of dt), log10 K ~ N(-11.5, 0.75), each its own seed."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(11)
    count = args.ensemble_count
    specs = []
    for i in range(count):
        if i % 2 == 0:
            s = int(round(np.exp(rng.uniform(np.log(32), np.log(512)))))
            specs.append((1, 2, (s, s), True))
        else:
            s = int(round(np.exp(rng.uniform(np.log(24), np.log(96)))))
            specs.append((2, 3, (s, s, s), False))

    def make(i):
        kind, dim, n, lpde = specs[i]
        p = dx.Problem(kind=kind, dim=dim, n=n, seed=5000 + i, mu_log10K=-11.5, sigma_log10K=0.75)
        a = dx.FVAssembler(p)
        a.setSolution(p.u0)
        return p, a

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=8) as ex:
        members = list(ex.map(make, range(count)))
    t_setup = time.perf_counter() - t0
    b = dx.Batch(linear=batch_linear(dx, args))
    for (p, a), (kind, dim, n, lpde) in zip(members, specs):
        b.add(a, nsteps=1, dt=args.dt, linear_pde=lpde)
    for p, a in members:
        a.synchronize()
    t0 = time.perf_counter()
    phases = 0
    while b.step():
        phases += 1
    el = time.perf_counter() - t0
    cs, ok, cells, fb = 0, 0, 0, 0
    for (p, a) in members:
        f, info = b.status(a)
        ok += f == 1
        fb += info.fp64_fallback
        cs += p.N * info.newton_steps
        cells += p.N
    b.close()
    for p, a in members:
        a.close()
    res = {"problems": count, "converged": ok, "cells": cells, "seconds": el, "phases": phases,
           "cell_steps_per_s": cs / el, "host_setup_seconds": t_setup, "fp64_fallback_problems": fb,
           "mix": "50% 1p 2D side log-uniform [32,512] stationary; 50% 2p 3D side log-uniform [24,96], "
                  f"1 implicit-Euler step dt={args.dt:g} s; log10 K ~ N(-11.5, 0.75)"}
    log(f"[bench] ensemble: {count} problems, {cells} cells, {el:.2f} s, {cs / el:.3e} cell-steps/s, "
        f"{ok} converged, setup {t_setup:.1f} s")
    return res


def run_serve(dx, args):
    from dumux_amd import serve
    out = []
    budget = int(args.hbm_budget_gb * 2 ** 30)
    for rate in [float(r) for r in args.serve_rates.split(",") if r]:
        rng = np.random.default_rng(int(rate * 1000) + 3)
        reqs = serve.poisson_stream(rate, args.serve_count, rng, cells_lo=1e4, cells_hi=2e6, nsteps=5, dt=args.dt,
                                    seed0=int(rate * 1000), mu_log10K=-11.0, sigma_log10K=0.5)
        srv = serve.Server(hbm_budget=budget, linear=batch_linear(dx, args), prep_threads=4)
        peak = [0]
        done = srv.run(reqs, on_phase=lambda s: peak.__setitem__(0, max(peak[0], len(s.batch))))
        srv.close()
        st = serve.summarize(done)
        st.update({"rate_per_s": rate, "peak_concurrent": peak[0], "phases": srv.phases,
                   "mean_cells": float(np.mean([r.cells for r in reqs]))})
        log(f"[bench] serve rate {rate:g}/s: {st['requests']} requests, p50 {st['p50_s']:.2f} s, "
            f"p99 {st['p99_s']:.2f} s, {st['cell_steps_per_s']:.3e} cell-steps/s, peak {peak[0]} concurrent")
        out.append(st)
    return {"budget_bytes": budget, "requests_per_rate": args.serve_count,
            "stream": "2p 3D cubes, cells log-uniform [1e4, 2e6], 5 implicit-Euler steps, Poisson arrivals",
            "rates": out}


def run_probe(dx, args):
    """Workload for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: the batched kernels on --batch
    problems of the workload (one implicit-Euler step), then 10 launches of the batched SpMV + dot
    (all members active), with multigrid 10 launches each of level 0's down- and up-sweep, then
    three copies of 512 MiB (hipMemcpy device to device), whose counters calibrate bytes per
    counter unit."""
    import ctypes as C
    n, B = args.n, args.batch
    probs = [workload_problem(dx, n, 2024 + i) for i in range(B)]
    asms = [dx.FVAssembler(p) for p in probs]
    for p, a in zip(probs, asms):
        a.setSolution(p.u0)
    b = dx.Batch(linear=batch_linear(dx, args))
    for a in asms:
        b.add(a, nsteps=1, dt=args.dt)
    b.run()
    ms, by = b.time_kernel(dx.KERNEL_BJ_SPMV_DOT_FP32, reps=10)
    log(f"[probe] batched spmv+dot: {ms * 1e3:.1f} us, algorithmic {by / 2**10:.0f} KiB per launch")
    probes = {"batched_spmv_dot": {"avg_ms": ms, "algorithmic_bytes": by}}
    if args.batch_precond == "amg":
        # then 10 launches each of level 0's multigrid down- and up-sweep
        for name, kid in (("batched_mg_down0", dx.KERNEL_MG_DOWN0), ("batched_mg_up0", dx.KERNEL_MG_UP0)):
            ms_k, by_k = b.time_kernel(kid, reps=10)
            probes[name] = {"avg_ms": ms_k, "algorithmic_bytes": by_k}
            log(f"[probe] {name}: {ms_k * 1e3:.1f} us, algorithmic {by_k / 2**10:.0f} KiB per launch")
    b.close()
    for a in asms:
        a.close()
    hip = C.CDLL("libamdhip64.so")
    size = 512 << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(size)) == 0 and hip.hipMalloc(C.byref(dst), C.c_size_t(size)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(size)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, C.c_size_t(size), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"probe": probes, "problems": B, "n": n, "copy_bytes": size}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import dumux_amd as dx
    if args.probe:
        run_probe(dx, args)
        return

    n = args.n
    N = n ** 3
    t_sec, value, binfos, u_batched, kernels = run_batched(dx, args)
    newton_b = [i.newton_steps for i in binfos[-1]]
    log(f"[bench] batched: {1e3 * t_sec:.1f} ms per solve of {args.batch} problems, {value:.3e} cell-steps/s")

    single = None
    if not args.no_single:
        chosen = dict(precond=dx.PRECOND_JACOBI, precision=dx.PRECISION_MIXED, **MIXED)
        single, u_single, k = run_single(dx, args, chosen, "single", [
            ("assemble", dx.KERNEL_ASSEMBLE, 20), ("cell_state", dx.KERNEL_CELL_STATE, 20),
            ("spmv_fp32", dx.KERNEL_SPMV_FP32, 20), ("bj_spmv_dot_fp32", dx.KERNEL_BJ_SPMV_DOT_FP32, 20),
            ("bj_spmv_dots_fp32", dx.KERNEL_BJ_SPMV_DOTS_FP32, 20), ("bj_iteration_fp32", dx.KERNEL_BJ_ITERATION_FP32, 20),
            ("refine_residual_fp64", dx.KERNEL_REFINE_RESIDUAL, 20), ("newton_update", dx.KERNEL_NEWTON_UPDATE, 20)])
        kernels.update(k)
        single["batched_vs_single_err"] = {
            "pw_rel": float(np.max(np.abs(u_batched[0::2] - u_single[0::2]) / np.abs(u_single[0::2]))),
            "sn": float(np.max(np.abs(u_batched[1::2] - u_single[1::2]) / np.maximum(1.0, np.abs(u_single[1::2]))))}

    single_mg = None
    if not args.no_single_mg:
        lin = dict(precond=dx.PRECOND_AMG, precision=dx.PRECISION_FP64, reduction=1e-10, max_iter=2000)
        single_mg, u_mg, _ = run_single(dx, args, lin, "single multigrid", [])
        single_mg["form"] = "FP64 BiCGSTAB + aggregation multigrid V-cycle (FP32 levels), one problem"
        single_mg["batched_vs_single_mg_err"] = {
            "pw_rel": float(np.max(np.abs(u_batched[0::2] - u_mg[0::2]) / np.abs(u_mg[0::2]))),
            "sn": float(np.max(np.abs(u_batched[1::2] - u_mg[1::2]) / np.maximum(1.0, np.abs(u_mg[1::2]))))}
        log(f"[bench] single multigrid: {single_mg['ms_per_step']:.1f} ms per solve, {single_mg['value']:.3e} cell-steps/s")

    plain = None
    if not args.no_plain:
        lin = dict(precond=dx.PRECOND_ILU0, precision=dx.PRECISION_FP64, reduction=1e-10, max_iter=2000)
        plain, u_plain, k = run_single(dx, args, lin, "plain", [
            ("ilu0_apply_fp64", dx.KERNEL_ILU0_APPLY_FP64, 5), ("spmv_fp64", dx.KERNEL_SPMV_FP64, 20)])
        kernels.update(k)
        plain["form"] = ("FP64 BiCGSTAB + ILU(0) (sequential-order arithmetic; sweeps as one launch each over 4^3 boxes), "
                         "one problem at a time")
        # accuracy of the chosen form against the FP64 form (problem seed 2024)
        plain["batched_vs_plain_err"] = {
            "pw_rel": float(np.max(np.abs(u_batched[0::2] - u_plain[0::2]) / np.abs(u_plain[0::2]))),
            "sn": float(np.max(np.abs(u_batched[1::2] - u_plain[1::2]) / np.maximum(1.0, np.abs(u_plain[1::2]))))}
        log(f"[bench] plain: {plain['ms_per_step']:.1f} ms per solve, {plain['value']:.3e} cell-steps/s; "
            f"batched vs plain {plain['batched_vs_plain_err']}")

    ensemble = None if args.no_ensemble else run_ensemble(dx, args)
    served = None if args.no_serve else run_serve(dx, args)

    def roofline_of(name, traffic_file):
        """Roofline record of a timed kernel: achieved = algorithmic bytes / launch time; traffic
        from the committed rocprofv3 counter summary of the same kernel and workload."""
        if name not in kernels:
            return None
        k = kernels[name]
        traffic, traffic_src = None, None
        tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", traffic_file)
        if n == 64 and args.batch == 8 and os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f)["traffic_bytes_per_launch"]
            traffic_src = "rocprofv3 FETCH_SIZE/WRITE_SIZE of bench.py --probe, scaled by a 512 MiB copy (profiles/)"
        return {"bound": "hbm", "kernel": name, "achieved": k["GBps"], "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                "frac": k["GBps"] / PEAK_HBM_GBPS, "traffic": traffic, "traffic_source": traffic_src,
                "bytes_per_launch": k["bytes"], "avg_launch_ms": k["ms"]}

    # the dominant kernel (largest share of the solve's GPU time) and, with multigrid, level 0's sweeps
    roofline = roofline_of("batched_spmv_dot_fp32", "traffic_batched_spmv_2p64x8.json")
    roofline_mg = [r for r in (roofline_of("batched_mg_down0", "traffic_batched_mg_down0_2p64x8.json"),
                               roofline_of("batched_mg_up0", "traffic_batched_mg_up0_2p64x8.json")) if r]

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(n, args.dt)
        except Exception as e:  # the product result stands without the baseline
            log(f"[bench] cpu baseline failed: {e}")

    out = {
        "metric": "2p Newton cell-steps/s, batched mixed precision (plain FP64 GPU form beside it)",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_sec, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64 (residual, Jacobian, refinement, Newton) + fp32 Krylov",
        "data": "synthetic (seeded log-normal K, seeds 2024..)",
        "config": {"workload": f"{args.batch} x 2p DNAPL infiltration {n}^3 ({N} cells each), {args.tsteps} x dt={args.dt:g} s "
                               "implicit Euler, numeric Jacobian, linear systems to 1e-10, solved as one batch",
                   "form": "batched executor: FP32 BiCGSTAB in FP64 iterative refinement, preconditioned by "
                           + ("aggregation multigrid (2x2x2 aggregates, Galerkin, block-Jacobi smoothing)"
                              if args.batch_precond == "amg" else "block Jacobi"),
                   "newton_steps_per_problem": newton_b,
                   "linear_iterations_per_problem": [i.linear_iterations for i in binfos[-1]],
                   "fp64_fallback_per_problem": [i.fp64_fallback for i in binfos[-1]]},
        "single_mixed_gpu": single,
        "single_multigrid_gpu": single_mg,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "ensemble": ensemble,
        "serve": served,
        "kernels": kernels,
        "roofline": roofline,
        "roofline_multigrid_level0": roofline_mg or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(n, dt):
    """The oracle (CPU restatement of the reference path) on the first implicit-Euler step of
    the workload: colored OpenMP assembly, sequential ILU0-BiCGSTAB (as dune-istl's
    SeqILU/BiCGSTABSolver), FP64."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle_ffi as O
    threads = min(16, os.cpu_count() or 1)
    po = O.OracleProblem(kind=2, dim=3, n=(n, n, n), seed=2024, pid=0, mu_log10K=-11.0, sigma_log10K=0.5)
    log(f"[bench] cpu baseline: oracle, 1 implicit-Euler step of {n}^3 with {threads} assembly threads")
    t0 = time.perf_counter()
    u, info = po.time_loop(po.u0, 1, dt, lin_reduction=1e-10, lin_max_iter=2000, nthreads=threads)
    el = time.perf_counter() - t0
    if not info.converged:
        raise RuntimeError("oracle did not converge")
    return {"value": po.N * info.newton_steps / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"first implicit-Euler step (dt={dt:g} s, {info.newton_steps} Newton iterations, "
                      f"{info.linear_iterations} BiCGSTAB iterations) of one {n}^3 problem, {el:.1f} s; "
                      "colored OpenMP assembly, sequential ILU0-BiCGSTAB"}


if __name__ == "__main__":
    main()
