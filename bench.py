"""Benchmark: bundle-adjustment LM iterations/s at 1.2M observations, mixed precision vs FP64.

Workload (one "step" = one 30-iteration LM solve, fixed schedule, from the same initial guess, with
the problem resident on the GPU): synthetic configuration 3 -- 1,500 cameras on a noisy random
walk, ~100.8k points, truncated Zipf(1.3) tracks on [2, 60] drawn until exactly 1,200,000
observations, Snavely cameras, Huber(1.0), implicit-Schur block-Jacobi PCG (max 50 iterations,
tol 1e-6), all termination tolerances 0 so exactly 30 LM iterations run.

value = mixed-precision LM iterations/s; the plain FP64 GPU path is timed the same way and reported
beside it, with the final-cost difference between the two.  The roofline entry is the dominant
kernel's algorithmic bytes per launch over its average duration, timed with HIP events on the
solver's stream; cpu_baseline is the FP64 CPU oracle (one thread) on a bounded sample of the same
problem.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LM_ITERS = 30


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solver_options(S, precision):
    return S.default_options(precision=precision, linear_solver=2, max_num_iterations=LM_ITERS,
                             pcg_max_iterations=50, pcg_tolerance=1e-6, function_tolerance=0.0,
                             gradient_tolerance=0.0, parameter_tolerance=0.0)


def time_solver(S, p, precision, steps, warmup):
    s = S.Solver(p, solver_options(S, precision))
    for _ in range(warmup):
        s.reset()
        s.run()
    total, iters, last = 0.0, 0, None
    for _ in range(steps):
        s.reset()
        t0 = time.perf_counter()
        out = s.run()  # synchronous: returns after the device finished
        total += time.perf_counter() - t0
        iters += out["num_iterations"] - 1  # trace row 0 is the initial evaluation
        last = out
    return s, total, iters, last


def profile_kernels(s):
    s.reset()
    s.set_profiling(True)
    s.run()
    kt = s.kernel_times()
    s.set_profiling(False)
    return kt


def cpu_baseline(S, p):
    from oracle import oracle as O
    o = O.default_options(max_num_iterations=1, linear_solver=2, pcg_max_iterations=50, pcg_tolerance=1e-6,
                          function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    t0 = time.perf_counter()
    O.solve(p.num_cameras(), p.num_points(), p.camera_index(), p.point_index(), p.observations(),
            p.mutable_cameras(), p.mutable_points(), o)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "LM iters/s", "cores": 1, "kind": "port",
            "sample": "FP64 CPU oracle, 1 thread: initial linearization + 1 LM iteration (50 PCG iterations) "
                      "of the full 1.2M-observation problem, %.1f s" % dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; a multi-GPU figure is not measured "
              "(independent windows shard by window, a large BA by points -- see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import slamplay_amd as S
    log("bench: generating configuration 3 (seed %d)" % args.seed)
    p = S.synthetic(3, seed=args.seed)
    nc, npt, no = p.num_cameras(), p.num_points(), p.num_observations()
    log("bench: %d cameras, %d points, %d observations" % (nc, npt, no))
    log("bench: mixed precision, %d warmup + %d timed solves" % (args.warmup, args.steps))
    sm, t_mixed, it_mixed, out_mixed = time_solver(S, p, 1, args.steps, args.warmup)
    mixed_ips = it_mixed / t_mixed
    log("bench: mixed %.1f LM iters/s, final cost %.9e" % (mixed_ips, out_mixed["final_cost"]))
    kt = profile_kernels(sm)
    res = {}
    if not args.no_fp64:
        log("bench: FP64")
        s64, t64, it64, out64 = time_solver(S, p, 0, args.steps, args.warmup)
        fp64_ips = it64 / t64
        res.update(fp64_iters_per_s=fp64_ips, mixed_vs_fp64=mixed_ips / fp64_ips,
                   final_cost_fp64=out64["final_cost"],
                   final_cost_rel_diff=abs(out_mixed["final_cost"] - out64["final_cost"]) / out64["final_cost"])
        log("bench: fp64 %.1f LM iters/s, final cost %.9e" % (fp64_ips, out64["final_cost"]))
        s64.close()
    # dominant kernel by total time among those with stated algorithmic bytes
    cands = {k: v for k, v in kt.items() if v["bytes_per_launch"] > 0 and v["launches"] > 0}
    dom = max(cands, key=lambda k: cands[k]["total_ms"]) if cands else None
    roofline = None
    if dom:
        v = cands[dom]
        avg_s = v["total_ms"] / v["launches"] / 1e3
        achieved = v["bytes_per_launch"] / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                    "avg_launch_us": avg_s * 1e6, "launches_per_solve": v["launches"]}
    kernel_share = {k: round(v["total_ms"], 3) for k, v in sorted(kt.items(), key=lambda kv: -kv[1]["total_ms"])}
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU oracle baseline (bounded sample)")
        cpu = cpu_baseline(S, p)
    line = {
        "metric": "BA LM iters/s at 1.2M obs (mixed vs FP64) + p99 batched local-BA latency",
        "value": mixed_ips,
        "unit": "LM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Jacobian/Hessian/PCG matvec + f64 block inverses and accept-test cost",
        "data": "synthetic (seeded generator, configuration 3)",
        "config": {"workload": "BA config 3: %d cameras, %d points, %d observations, mixed precision, "
                               "implicit-Schur block-Jacobi PCG (max 50, tol 1e-6), %d LM iterations" %
                               (nc, npt, no, LM_ITERS),
                   "cameras": nc, "points": npt, "observations": no, "lm_iterations": LM_ITERS},
        "final_cost": out_mixed["final_cost"],
        "initial_cost": out_mixed["initial_cost"],
        "pcg_iterations": out_mixed["linear_solver_iterations"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernel_ms_per_solve": kernel_share,
    }
    line.update(res)
    sm.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
