"""Benchmark: batched SQP solves/s and fwd+bwd grads/s, fused vs per-launch GPU path.

Workload (default): quadrotor nonlinear MPC, nx=13, nu=4, horizon T=50
(N=50 dynamics steps), batch B=1024, 5 SQP iterations, PCG eps 1e-6 absolute,
synthetic seeded inputs.  One step = one forward solve of the whole batch plus
the implicit-differentiation backward of the tracking loss
0.5 * sum ||x* - x_ref||^2 with respect to the shared state/control weights and
the per-instance initial states.  `value` is the fused path's grads/s.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported
(instances are independent; one process drives one GPU).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SQP solves/s & fwd+bwd grads/s, quadrotor T=50 B=1024, fused vs plain GPU"
MI355X_FP64_PEAK_TFLOPS = 78.6
# HBM bytes per qp_iteration_kernel launch (one whole solve: five SQP
# iterations with their linearizations) at the default workload, from two
# rocprofv3 --pmc passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, scale factors
# from a 256 MiB device copy): 463.9 MB read + 686.6 MB written, against 731
# MB algorithmic; the excess is the linearization's register spills
QP_TRAFFIC_BYTES = 1150.5e6
# the register-resident forward of a batch of one horizon with more instances
# than CUs: one launch per solve (each workgroup linearizes one instance and
# runs one SQP iteration of it, as soon as the instance's previous iteration
# is done); otherwise one linearization and one QP launch per SQP iteration
QP_KERNEL_TEXT = ("qp_iteration_kernel (assembly + PCG + recovery of every SQP iteration; for a batch of one "
                  "horizon with more instances than CUs one launch per solve, linearizations included)")
QP_TRAFFIC_SOURCE = "profiles/qp_iteration_quadrotor_B1024_T50.md"
# HBM bytes per forward SQP iteration of the chain workload (stream_assemble_kernel
# + stream_qp_kernel, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0): 0.983 TB read + 0.757 TB
# written (the linear model's blocks are one shared copy per instance, staged
# in LDS but for D_0 and D_N; the traffic is the PCG vectors in the workspace)
CHAIN_TRAFFIC_BYTES = 1.740e12
CHAIN_TRAFFIC_SOURCE = "profiles/chain_nx48_B512_N256.md"
# HBM bytes per forward launch (stream_assemble_kernel + stream_qp_kernel, one SQP
# iteration, mean over the three) of the ragged mix's dominant group, the streamed
# quadrotor instances with N in 64..200: 275.3 GB read + 22.4 GB written
RAGGED_TRAFFIC_BYTES = 297.7e9
RAGGED_TRAFFIC_GROUP = "quadrotor/1"
RAGGED_TRAFFIC_SOURCE = "profiles/ragged_mixed_B8192_T10-200.md"


def pcg_flops(N, nx):
    """Algorithmic flops of one PCG iteration of one instance: two block-tridiagonal
    matvecs over (3N+1) nx^2 entries plus dots and axpys."""
    return 4 * (3 * N + 1) * nx * nx + 10 * (N + 1) * nx


def stream_block_bytes(nx, nu):
    """Algorithmic HBM bytes of one PCG iteration per time block of the streamed
    kernels: compact layout (nx <= 16) reads A three times and B once (float32)
    and the packed D_t twice (float64); the full layout (nx > 16) reads A_t and
    A_{t-1} twice each, B once (float32) and the full D_t twice (float64)."""
    if nx <= 16:
        return 12 * nx * nx + 4 * nx * nu + 8 * nx * (nx + 1)
    return 16 * nx * nx + 4 * nx * nu + 16 * nx * nx


def cpu_env():
    """Host facts the CPU baseline is quoted with."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                          "NUMEXPR_NUM_THREADS")}
    return {"cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model,
            "thread_env": env}


def _oracle_model(name, prob):
    from oracle import dynamics as odyn

    if name == "linear":
        return odyn.make_model("linear", prob["x0"].shape[1], prob["u_ref"].shape[2])
    return odyn.make_model(name)


def _cpu_warm(_):
    time.sleep(0.05)
    return os.getpid()


def _cpu_init():
    import numpy  # noqa: F401

    from diffmpc_amd import workloads  # noqa: F401
    from oracle import sqp  # noqa: F401


def cpu_chunk(job):
    """One worker's share of a CPU baseline: the float64 oracle (one thread)
    on instances [lo, hi) of a seeded workload, `reps` times; returns
    (instances solved, seconds)."""
    gen, kw, lo, hi, backward, reps = job
    import numpy as np
    from threadpoolctl import threadpool_limits

    from diffmpc_amd import workloads as W
    from oracle import sqp as osqp

    name, prob, sp = getattr(W, gen)(**kw)[:3]
    sub = W.slice_problem(prob, slice(lo, hi))
    sub["model"] = _oracle_model(name, prob)

    def once():
        ref = osqp.sqp_solve(sub, sp)
        if backward:
            osqp.sqp_backward(sub, sp, ref, ref["states"] - sub["x_ref"], np.zeros_like(ref["controls"]))

    times = []
    with threadpool_limits(limits=1):
        if reps > 1:
            once()  # warm-up
        for _ in range(reps):
            t0 = time.perf_counter()
            once()
            times.append(time.perf_counter() - t0)
    return hi - lo, times


def cpu_baseline(gen, kw, backward, sample_1core, sample_all, unit, what, reps_1core=1, chunk_reps=1):
    """The oracle timed on the host: one thread on `sample_1core` instances
    (the median over `reps_1core` repeats after a warm-up when repeated),
    then every core of this process's CPU share (at most 16 worker processes,
    one thread each) on `sample_all` instances split evenly, by wall clock;
    rates in instances per second."""
    import multiprocessing as mp
    import statistics

    n1, times = cpu_chunk((gen, kw, 0, sample_1core, backward, reps_1core))
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    per = max(1, sample_all // workers)
    jobs = [(gen, kw, i * per, (i + 1) * per, backward, chunk_reps) for i in range(workers)]
    with mp.get_context("spawn").Pool(workers, initializer=_cpu_init) as pool:
        pool.map(_cpu_warm, range(4 * workers))  # every worker started and imported
        t0 = time.perf_counter()
        res = pool.map(cpu_chunk, jobs)
        t_all = time.perf_counter() - t0
    n_all = sum(r[0] * len(r[1]) for r in res)
    rep = f" x {reps_1core} repeats (median)" if reps_1core > 1 else ""
    return {
        "value": n1 / statistics.median(times), "unit": unit, "cores": 1, "kind": "port",
        "sample": f"{what}: {sample_1core} instances{rep}, float64 numpy oracle, one thread",
        "all_cores": {"value": n_all / t_all, "cores": workers,
                      "sample": f"{n_all} instance solves on {workers} worker processes (one thread each), "
                                "wall clock" + (" including one warm-up solve per worker" if chunk_reps > 1 else "")},
        "host": cpu_env(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--horizon", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=192, help="instances in the one-thread CPU baseline sample")
    ap.add_argument("--fused-only", action="store_true",
                    help="time only the fused path (for profiler runs); prints the fused figures")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB mebibytes device-to-device once (reference kernel for HBM counters)")
    ap.add_argument("--blocks-per-group", type=int, default=0, metavar="K",
                    help="force K time blocks per lane group in the register-resident kernels "
                         "(1, 2 or 4; 0 = automatic) for every solve of the run")
    ap.add_argument("--workload", choices=["quadrotor", "config0", "spacecraft7", "ragged", "chain", "rollout",
                                           "dynparams", "imitation"],
                    default="quadrotor",
                    help="quadrotor: the headline line (default); config0: the single-instance LTI QP (nx=4 nu=2 "
                         "T=20) with the CPU solves/s headline; spacecraft7: attitude MPC B=4096 N=100 forward + "
                         "adjoint w.r.t. x0, log-weights and x_ref; ragged: mixed batch of 8192 with horizons "
                         "U{10..200} in one graph; chain: 48-state chain B=512 N=256; rollout: the reference's "
                         "RL benchmark (LTI 8/4, N=39, 50 closed-loop steps, B=--batch or 64); dynparams: the "
                         "headline workload with gradients w.r.t. mass and inertia as well; imitation: the "
                         "reference's cart-pole imitation-learning step (32 demonstrations, N=100)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "check": float(dst[-1])}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: instances are independent and one process drives one GPU; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from diffmpc_amd import _lib
    from diffmpc_amd import workloads as W

    if args.blocks_per_group:
        _lib.check(_lib.load().dmpc_set_blocks_per_group(args.blocks_per_group), "dmpc_set_blocks_per_group")
    if args.workload == "config0":
        return bench_config0(args)
    if args.workload == "spacecraft7":
        return bench_spacecraft7(args)
    if args.workload == "ragged":
        return bench_ragged(args)
    if args.workload == "chain":
        return bench_chain(args)
    if args.workload == "rollout":
        return bench_rollout(args)
    if args.workload == "dynparams":
        return bench_dynparams(args)
    if args.workload == "imitation":
        return bench_imitation(args)

    dev = torch.device("cuda:0")
    name, prob, sp = W.quadrotor(batch=args.batch, seed=1, horizon=args.horizon)
    N, nx = prob["horizon"], prob["x0"].shape[1]

    def make(mode):
        solver, params, guess = W.make_solver(name, prob, sp, device=dev, mode=mode)
        return solver, params, guess

    def fwd_bwd(solver, params, guess):
        # cold adjoint start every step: the backward writes its multipliers
        # into the guess (the reference's warm-start channel), which would
        # otherwise hand the next step a converged adjoint
        guess.dual_backward_guess.zero_()
        wx = params["weights_penalization_reference_state_trajectory"][0].clone().requires_grad_(True)
        wu = params["weights_penalization_control_squared"][0].clone().requires_grad_(True)
        x0 = params["initial_state"].clone().requires_grad_(True)
        weights = {"weights_penalization_reference_state_trajectory": wx,
                   "weights_penalization_control_squared": wu, "initial_state": x0}
        sol = solver.solve(guess, params, weights)
        loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
        grads = torch.autograd.grad(loss, [wx, wu, x0])
        return sol, grads

    def fwd(solver, params, guess):
        with torch.no_grad():
            return solver.solve(guess, params)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, out

    # ---- fused path -------------------------------------------------------
    solver, params, guess = make("fused")
    t_fwd, sol = timed(lambda: fwd(solver, params, guess), args.steps, args.warmup)
    lib = _lib.load()
    _lib.profile_read(_lib.PROF_QP_ITERATION)
    _lib.profile_read(_lib.PROF_BACKWARD)
    _lib.profile_read(_lib.PROF_LINEARIZE)
    lib.dmpc_profile_enable(1)
    t_fb, (sol_fb, _) = timed(lambda: fwd_bwd(solver, params, guess), args.steps, args.warmup)
    torch.cuda.synchronize()
    lib.dmpc_profile_enable(0)
    qp_ms, qp_n = _lib.profile_read(_lib.PROF_QP_ITERATION)
    bw_ms, bw_n = _lib.profile_read(_lib.PROF_BACKWARD)
    ln_ms, ln_n = _lib.profile_read(_lib.PROF_LINEARIZE)
    pcg_fwd = int(sol_fb.num_pcg_iters.sum())  # over instances and SQP iterations of one step
    pcg_bwd = int(solver.last_backward_pcg_iters.sum())
    steps_prof = args.steps + args.warmup
    # algorithmic PCG flops of the profiled solves over the QP kernels' time
    # (one launch per solve, or one per SQP iteration with ragged horizons)
    flops_qp_total = pcg_flops(N, nx) * pcg_fwd * steps_prof
    flops_qp_launch = flops_qp_total / max(qp_n, 1)
    qp_avg_ms = qp_ms / max(qp_n, 1)
    achieved = flops_qp_total / (qp_ms * 1e-3) / 1e12

    if args.fused_only:
        print(json.dumps({"fused_ms_fwd": t_fwd * 1e3, "fused_ms_fwd_bwd": t_fb * 1e3,
                          "qp_iteration_launch_avg_ms": qp_avg_ms, "qp_iteration_launches": qp_n,
                          "backward_launch_avg_ms": bw_ms / max(bw_n, 1), "pcg_forward": pcg_fwd,
                          "pcg_backward": pcg_bwd, "achieved_tflops": achieved}), flush=True)
        return

    # ---- per-launch path --------------------------------------------------
    psolver, pparams, pguess = make("plain")
    t_pfwd, psol = timed(lambda: fwd(psolver, pparams, pguess), args.plain_steps, 1)
    t_pfb, _ = timed(lambda: fwd_bwd(psolver, pparams, pguess), args.plain_steps, 1)
    same = float((sol.states - psol.states).abs().max() / sol.states.abs().max())

    B = args.batch
    out = {
        "metric": METRIC,
        "value": B / t_fb,
        "unit": "grads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_fb * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG",
        "data": "synthetic",
        "config": {"workload": f"quadrotor nx=13 nu=4 T={N} B={B}, 5 SQP iterations, PCG eps 1e-6",
                   "global_batch": B, "horizon": N, "parallelism": "single GPU"},
        "fused": {"solves_per_s": B / t_fwd, "grads_per_s": B / t_fb, "ms_fwd": t_fwd * 1e3, "ms_fwd_bwd": t_fb * 1e3},
        "plain": {"solves_per_s": B / t_pfwd, "grads_per_s": B / t_pfb, "ms_fwd": t_pfwd * 1e3,
                  "ms_fwd_bwd": t_pfb * 1e3},
        "fused_vs_plain": {"solves": t_pfwd / t_fwd, "grads": t_pfb / t_fb},
        "fused_plain_max_rel_diff": same,
        "pcg_iterations_per_step": {"forward": pcg_fwd, "backward": pcg_bwd},
        "kernel_ms_per_step": {"qp_iteration": qp_ms / steps_prof, "backward": bw_ms / steps_prof,
                               "linearize": ln_ms / steps_prof},
        "roofline": {
            "bound": "fp64_valu",
            "kernel": QP_KERNEL_TEXT,
            "achieved": achieved,
            "peak": MI355X_FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / MI355X_FP64_PEAK_TFLOPS,
            "traffic": QP_TRAFFIC_BYTES if (N, args.batch) == (50, 1024) else None,
            "traffic_source": QP_TRAFFIC_SOURCE,
            "launch_avg_ms": qp_avg_ms,
            "flops_per_launch": flops_qp_launch,
        },
    }

    # ---- CPU baseline: the float64 oracle on a bounded sample --------------
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(
            "quadrotor", dict(batch=args.batch, seed=1, horizon=args.horizon), True, args.cpu_sample, args.batch,
            "grads/s", f"forward + backward of the headline instances (of {args.batch})")
    print(json.dumps(out), flush=True)


def _timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        out = fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def _qp_roofline(lib_mod, fn, N, nx, pcg_count):
    """Run fn once with the library's launch events on; returns the QP-iteration
    kernel's roofline object (FP64 VALU: algorithmic PCG flops per launch over
    the mean launch time) and fn's result."""
    import torch

    lib_mod.profile_read(lib_mod.PROF_QP_ITERATION)
    lib_mod.load().dmpc_profile_enable(1)
    out = fn()
    torch.cuda.synchronize()
    lib_mod.load().dmpc_profile_enable(0)
    ms, n = lib_mod.profile_read(lib_mod.PROF_QP_ITERATION)
    avg = ms / max(n, 1)
    flops = pcg_flops(N, nx) * pcg_count(out)  # all SQP iterations of the solve
    achieved = flops / (ms * 1e-3) / 1e12
    return {"bound": "fp64_valu", "kernel": QP_KERNEL_TEXT,
            "achieved": achieved, "peak": MI355X_FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / MI355X_FP64_PEAK_TFLOPS, "traffic": None, "launch_avg_ms": avg,
            "flops_per_launch": flops / max(n, 1)}, out


def bench_config0(args):
    """Config 0: the plain LTI MPC QP (nx=4, nu=2, T=20, A = I + 0.1 G,
    Q = I, R = 0.1 I, Qf = 10 I, one SQP iteration, PCG 1e-10 / 200) -- the
    reference's CPU path.  One instance on the GPU (forward solve, and
    forward + backward, each captured as one hipGraph: launch-latency bound),
    the same configuration batched (B = 4096 seeded instances), and the
    float64 oracle on the host: one instance on one thread (median of 20
    solves after a warm-up), and one instance per worker on every core."""
    import torch

    from diffmpc_amd import _lib
    from diffmpc_amd import autotune
    from diffmpc_amd import workloads as W
    from diffmpc_amd.graph import GraphedStep

    dev = torch.device("cuda:0")
    res = {}
    for B in (1, 4096):
        name, prob, sp = W.lti(batch=B, seed=0)
        solver, params, guess = W.make_solver(name, prob, sp, device=dev)
        shapes = None

        def fwd():
            with torch.no_grad():
                return solver.solve(guess, params)

        def fwd_bwd():
            guess.dual_backward_guess.zero_()
            wx = params["weights_penalization_reference_state_trajectory"].clone().requires_grad_(True)
            x0 = params["initial_state"].clone().requires_grad_(True)
            sol = solver.solve(guess, params, {"weights_penalization_reference_state_trajectory": wx,
                                               "initial_state": x0})
            return torch.autograd.grad(0.5 * sol.states.square().sum(), [wx, x0])

        if B > 1:
            # the batch packs several instances per workgroup (an instance uses
            # 84 lanes): graph-replayed forward solve per packing, then the
            # autotuner's choice of launch shape for the timed lines
            shapes = {}
            for ipw in (1, 2, 3, 4):
                solver.params["instances_per_workgroup"] = 0 if ipw == 1 else ipw
                t, _ = _timed(GraphedStep(fwd).replay, args.steps, args.warmup)
                shapes[str(ipw)] = t * 1e3
            best, _ = autotune.tune(solver, params, guess, reps=3)
        roof, sol = _qp_roofline(_lib, fwd, prob["horizon"], 4, lambda o: int(o.num_pcg_iters.sum()))
        steps = max(args.steps, 100) if B == 1 else args.steps
        gf, gb = GraphedStep(fwd), GraphedStep(fwd_bwd)
        t_f, _ = _timed(gf.replay, steps, args.warmup)
        t_b, _ = _timed(gb.replay, steps, args.warmup)
        res[B] = {"solves_per_s": B / t_f, "grads_per_s": B / t_b, "ms_fwd": t_f * 1e3, "ms_fwd_bwd": t_b * 1e3,
                  "pcg_iterations": int(sol.num_pcg_iters.sum()), "roofline": roof}
        if shapes:
            res[B]["launch_shape"] = list(best)
            res[B]["ms_fwd_by_instances_per_workgroup"] = shapes
    out = {
        "metric": "config 0: LTI MPC QP solves/s, one instance (nx=4 nu=2 T=20, 1 SQP iteration, PCG 1e-10)",
        "value": res[1]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": max(args.steps, 100),
        "warmup": args.warmup, "ms_per_step": res[1]["ms_fwd"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG", "data": "synthetic",
        "config": {"workload": "LTI nx=4 nu=2 T=20, one instance, one hipGraph launch per solve", "global_batch": 1},
        "single_instance": res[1], "batched_4096": res[4096],
        "roofline": res[4096]["roofline"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline("lti", dict(batch=16, seed=0), False, 1, 16, "solves/s",
                                           "one config-0 instance solved", reps_1core=20, chunk_reps=200)
    print(json.dumps(out), flush=True)


def bench_spacecraft7(args):
    """Config 2: spacecraft attitude MPC, B = 4096, nx = 7, nu = 3, N = 100,
    dt = 0.1, Euler, inertia ~ U(0.5, 2), 3 SQP iterations, PCG 1e-8; the
    loss 0.5 sum ||x* - x_ref||^2 differentiated w.r.t. x0, the log-weights
    theta_Q, theta_R (w = exp(theta), shared by the batch) and x_ref: fused vs
    the per-launch path."""
    import torch

    from diffmpc_amd import _lib
    from diffmpc_amd import workloads as W

    dev = torch.device("cuda:0")
    B = args.batch if args.batch != 1024 else 4096
    name, prob, sp = W.spacecraft7(batch=B, seed=2)
    N = prob["horizon"]
    th_q0 = torch.tensor(prob["theta_q"], dtype=torch.float64, device=dev)
    th_r0 = torch.tensor(prob["theta_r"], dtype=torch.float64, device=dev)

    def make(mode):
        solver, params, guess = W.make_solver(name, prob, sp, device=dev, mode=mode)

        def fwd_bwd():
            guess.dual_backward_guess.zero_()
            th_q = th_q0.clone().requires_grad_(True)
            th_r = th_r0.clone().requires_grad_(True)
            x0 = params["initial_state"].clone().requires_grad_(True)
            xr = params["reference_state_trajectory"].clone().requires_grad_(True)
            sol = solver.solve(guess, params, {"weights_penalization_reference_state_trajectory": th_q.exp(),
                                               "weights_penalization_control_squared": th_r.exp(),
                                               "initial_state": x0, "reference_state_trajectory": xr})
            loss = 0.5 * (sol.states - xr).square().sum()
            return sol, torch.autograd.grad(loss, [x0, th_q, th_r, xr])

        def fwd():
            with torch.no_grad():
                return solver.solve(guess, params)
        return fwd, fwd_bwd, solver

    fwd, fwd_bwd, solver = make("fused")
    t_f, _ = _timed(fwd, args.steps, args.warmup)
    t_b, _ = _timed(fwd_bwd, args.steps, args.warmup)
    roof, (sol, _) = _qp_roofline(_lib, fwd_bwd, N, 7, lambda o: int(o[0].num_pcg_iters.sum()))
    pfwd, pfwd_bwd, _ = make("plain")
    t_pf, _ = _timed(pfwd, 1, 1)
    t_pb, _ = _timed(pfwd_bwd, 1, 1)
    out = {
        "metric": "config 2: spacecraft attitude fwd+adjoint grads/s (nx=7 nu=3 N=100 B=4096), fused vs plain GPU",
        "value": B / t_b, "unit": "grads/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_b * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG", "data": "synthetic",
        "config": {"workload": f"spacecraft attitude nx=7 nu=3 N={N} B={B}, 3 SQP iterations, PCG 1e-8, grads "
                               "w.r.t. x0, log Q/R weights, x_ref", "global_batch": B},
        "fused": {"solves_per_s": B / t_f, "grads_per_s": B / t_b, "ms_fwd": t_f * 1e3, "ms_fwd_bwd": t_b * 1e3},
        "plain": {"solves_per_s": B / t_pf, "grads_per_s": B / t_pb, "ms_fwd": t_pf * 1e3, "ms_fwd_bwd": t_pb * 1e3},
        "fused_vs_plain": {"solves": t_pf / t_f, "grads": t_pb / t_b},
        "pcg_iterations_per_step": {"forward": int(sol.num_pcg_iters.sum()),
                                    "backward": int(solver.last_backward_pcg_iters.sum())},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline("spacecraft7", dict(batch=B, seed=2), True, 64, 1024, "grads/s",
                                           f"forward + backward of config-2 instances (of {B})")
    print(json.dumps(out), flush=True)


def _fwd_bwd_step(groups, wkey="weights_penalization_reference_state_trajectory"):
    """One step over (solver, params, guess, weight) groups: solve, tracking
    loss 0.5 ||x* - x_ref||^2, gradient w.r.t. the state weights."""
    import torch

    def step():
        outs = []
        for solver, params, guess, w in groups:
            guess.dual_backward_guess.zero_()
            sol = solver.solve(guess, params, {wkey: w})
            loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
            (g,) = torch.autograd.grad(loss, [w])
            outs.append((sol.states, g))
        return outs

    return step


def bench_ragged(args):
    """Config 3: 8192 instances, horizons N ~ U{10..200}, systems drawn from
    {LTI 4/2, spacecraft attitude 7/3, quadrotor 13/4}, 3 SQP iterations;
    the autotuner picks each group's launch shape, then forward + backward of
    all groups is captured as one hipGraph and replayed."""
    import torch

    from diffmpc_amd import autotune
    from diffmpc_amd import workloads as W
    from diffmpc_amd.graph import GraphedStep

    from diffmpc_amd import _lib

    dev = torch.device("cuda:0")
    groups, info = [], {}
    model_ids = {"linear": (_lib.MODEL_LINEAR, 4, 2), "spacecraft7": (_lib.MODEL_SPACECRAFT7, 7, 3),
                 "quadrotor": (_lib.MODEL_QUADROTOR, 13, 4)}
    for name, (prob, sp, h) in W.ragged(batch=8192, seed=3).items():
        # instances short enough for the register-resident kernels form their own group
        limit = autotune.register_horizon_limit(*model_ids[name])
        for part, (sub, hs) in enumerate(W.split_by_horizon(prob, h, limit)):
            solver, params, guess = W.make_solver(name, sub, sp, device=dev, horizons=hs)
            best, timings = autotune.tune(solver, params, guess, reps=1)
            w = params["weights_penalization_reference_state_trajectory"].clone().requires_grad_(True)
            groups.append((solver, params, guess, w))
            info[f"{name}/{part}"] = {"instances": int(len(hs)), "max_horizon": int(hs.max()),
                                      "launch_shape": list(best),
                                      "tuning_ms": {"/".join(map(str, k)): v for k, v in timings.items()}}
    # per-group roofline of the forward QP kernels (one eager forward solve
    # with the library's launch events): register-resident groups against
    # the FP64 VALU peak (algorithmic PCG flops), streamed groups against HBM
    # (algorithmic block bytes per PCG iteration)
    lib = _lib.load()
    for (solver, params, guess, _), (key, gi) in zip(groups, info.items()):
        nx, nu = solver.program.num_state_variables, solver.program.num_control_variables
        hs = params["horizons"].cpu().numpy()
        _lib.profile_read(_lib.PROF_QP_ITERATION)
        lib.dmpc_profile_enable(1)
        with torch.no_grad():
            sol = solver.solve(guess, params)
        torch.cuda.synchronize()
        lib.dmpc_profile_enable(0)
        ms, n = _lib.profile_read(_lib.PROF_QP_ITERATION)
        iters = sol.num_pcg_iters.sum(dim=1).cpu().numpy()
        if gi["launch_shape"][0] == "streamed":
            amount = float(((hs + 1) * iters).sum()) * stream_block_bytes(nx, nu)
            gi["roofline"] = {"bound": "hbm", "achieved": amount / (ms * 1e-3) / 1e9, "peak": 8000.0,
                              "unit": "GB/s", "kernel_ms": ms, "launches": n}
        else:
            amount = float(sum(pcg_flops(int(h), nx) * int(k) for h, k in zip(hs, iters)))
            gi["roofline"] = {"bound": "fp64_valu", "achieved": amount / (ms * 1e-3) / 1e12,
                              "peak": MI355X_FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "kernel_ms": ms, "launches": n}
        gi["roofline"]["frac"] = gi["roofline"]["achieved"] / gi["roofline"]["peak"]
    dom = max(info, key=lambda k: info[k]["roofline"]["kernel_ms"])
    graphed = GraphedStep(_fwd_bwd_step(groups))
    for _ in range(args.warmup):
        graphed.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graphed.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    B = sum(v["instances"] for v in info.values())
    print(json.dumps({
        "metric": "ragged mixed batch fwd+bwd grads/s (8192 instances, N ~ U{10..200}, 3 SQP iterations)",
        "value": B / dt, "unit": "grads/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG", "data": "synthetic",
        "config": {"workload": "ragged LTI/spacecraft7/quadrotor mix, one hipGraph per step", "global_batch": B},
        "roofline": dict(info[dom]["roofline"], group=dom,
                         kernel=("stream assembly + stream_qp_kernel" if info[dom]["launch_shape"][0] == "streamed"
                                 else "qp_iteration_kernel") + " (forward, all SQP iterations of the group)",
                         traffic=RAGGED_TRAFFIC_BYTES if dom == RAGGED_TRAFFIC_GROUP else None,
                         traffic_unit="bytes per forward launch (one SQP iteration)",
                         traffic_source=RAGGED_TRAFFIC_SOURCE),
        "groups": info,
    }), flush=True)


def bench_chain(args):
    """Config 4: 12-mass chain, nx=48 nu=12, N=256, B=512, one SQP iteration
    (linear system), PCG eps 1e-8; fused (streamed variant) vs the per-launch
    path, forward + backward."""
    import torch

    from diffmpc_amd import _lib
    from diffmpc_amd import workloads as W

    dev = torch.device("cuda:0")
    name, prob, sp = W.chain(batch=512, seed=4)
    N, nx, nu = prob["horizon"], 48, 12
    res = {}
    for mode in ("fused",) if args.fused_only else ("fused", "plain"):
        solver, params, guess = W.make_solver(name, prob, sp, device=dev, mode=mode)
        w = params["weights_penalization_reference_state_trajectory"].clone().requires_grad_(True)
        step = _fwd_bwd_step([(solver, params, guess, w)])
        step()
        torch.cuda.synchronize()
        if mode == "fused":
            _lib.profile_read(_lib.PROF_QP_ITERATION)
            _lib.load().dmpc_profile_enable(1)
        n = args.steps if mode == "fused" else 1
        t0 = time.perf_counter()
        for _ in range(n):
            out = step()
        torch.cuda.synchronize()
        res[mode] = (time.perf_counter() - t0) / n
        if mode == "fused":
            _lib.load().dmpc_profile_enable(0)
            qp_ms, qp_n = _lib.profile_read(_lib.PROF_QP_ITERATION)
            with torch.no_grad():
                sol = solver.solve(guess, params)
            iters = int(sol.num_pcg_iters.sum())
    # The chain model is linear, so the streamed kernels keep one copy of its
    # blocks per instance and read them from cache: the launch is bound by
    # its FP64 instruction stream, not by HBM.  Roofline: algorithmic PCG
    # flops per launch over the launch time; `block_read_gbs` restates the
    # block bytes a PCG iteration reads per knot (A_t and A_{t-1} twice each,
    # B_{t-1} once, float; D_t twice, double) as a rate, served by the caches.
    per_block = 4 * nx * nx * 4 + nx * nu * 4 + 2 * nx * nx * 8
    launch_ms = qp_ms / max(qp_n, 1)
    block_gbs = per_block * (N + 1) * iters / (launch_ms * 1e-3) / 1e9
    achieved = pcg_flops(N, nx) * iters / (launch_ms * 1e-3) / 1e12
    B = prob["x0"].shape[0]
    if args.fused_only:
        print(json.dumps({"fused_ms_fwd_bwd": res["fused"] * 1e3, "qp_launch_avg_ms": launch_ms,
                          "pcg_iterations_forward": iters, "achieved_tflops": achieved,
                          "block_read_gbs": block_gbs}), flush=True)
        return
    print(json.dumps({
        "metric": "chain fwd+bwd grads/s (nx=48 nu=12 N=256 B=512), fused vs plain GPU",
        "value": B / res["fused"], "unit": "grads/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": res["fused"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG", "data": "synthetic",
        "config": {"workload": "12-mass chain nx=48 nu=12 N=256 B=512, 1 SQP iteration, PCG eps 1e-8",
                   "global_batch": B},
        "plain": {"grads_per_s": B / res["plain"], "ms_fwd_bwd": res["plain"] * 1e3},
        "fused_vs_plain": res["plain"] / res["fused"],
        "pcg_iterations_forward": iters,
        "roofline": {"bound": "fp64_valu", "kernel": "stream assembly + stream_qp_kernel (forward SQP iteration)",
                     "achieved": achieved, "peak": MI355X_FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MI355X_FP64_PEAK_TFLOPS,
                     "traffic": CHAIN_TRAFFIC_BYTES if (N, B) == (256, 512) else None,
                     "traffic_source": CHAIN_TRAFFIC_SOURCE, "launch_avg_ms": launch_ms,
                     "block_read_gbs": block_gbs},
    }), flush=True)


# reference JAX GPU run of the same rollout (B=64, PCG 1e-12, warm start),
# benchmarking/reinforcement-learning/timing_results/timing_64_40_12_pcg=1e-12_warmstart=True
RL_REFERENCE_MS = {"fwd": 132.8, "fwd_bwd": 214.9}


def bench_rollout(args):
    """The reference's RL benchmark (benchmark_diffmpc.py): LTI nx=8 nu=4,
    horizon 40 (N=39), 50 closed-loop steps, one SQP iteration, PCG 1e-12,
    warm-started solves; forward rollout and gradient of the summed reward
    w.r.t. the state weights, each captured as one hipGraph."""
    import torch

    from diffmpc_amd import workloads as W
    from diffmpc_amd.graph import GraphedStep
    from diffmpc_amd.rollout import Rollout, linear_simulator

    dev = torch.device("cuda:0")
    B = args.batch if args.batch != 1024 else 64
    steps_sim = 50
    name, prob, sp, (A, Bm, b) = W.rl_lti(batch=B, seed=0)
    solver, params, _ = W.make_solver(name, prob, sp, device=dev)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    ro = Rollout(solver, params, linear_simulator(t(A), t(Bm), t(b)),
                 lambda x, u: -(x.square().sum(-1) + u.square().sum(-1)), steps_sim, adjoint_warm_start=False)
    wx = params["weights_penalization_reference_state_trajectory"][0].clone().requires_grad_(True)
    weights = {"weights_penalization_reference_state_trajectory": wx}
    x_init = params["initial_state"]

    def fwd():
        with torch.no_grad():
            return ro(x_init, weights)[0]

    def fwd_bwd():
        total = ro(x_init, weights)[0]
        return torch.autograd.grad(total.sum(), [wx])[0]

    res = {}
    for key, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
        g = GraphedStep(fn)
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        res[key] = (time.perf_counter() - t0) / args.steps
    solves = B * steps_sim
    print(json.dumps({
        "metric": "RL closed-loop rollout fwd+bwd, LTI nx=8 nu=4 N=39, 50 steps (solves/s in the gradient pass)",
        "value": solves / res["fwd_bwd"], "unit": "grads/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["fwd_bwd"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": (RL_REFERENCE_MS["fwd_bwd"] / (res["fwd_bwd"] * 1e3)) if B == 64 else None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG", "data": "synthetic (the reference's seeded data)",
        "config": {"workload": f"RL rollout LTI 8/4 N=39, 50 steps, PCG 1e-12, B={B}", "global_batch": B},
        "fwd": {"ms": res["fwd"] * 1e3, "solves_per_s": solves / res["fwd"],
                "vs_reference_jax_gpu": (RL_REFERENCE_MS["fwd"] / (res["fwd"] * 1e3)) if B == 64 else None},
        "reference_jax_gpu_ms": RL_REFERENCE_MS,
    }), flush=True)


def bench_dynparams(args):
    """Headline workload (quadrotor B=1024 N=50, 5 SQP iterations) with the
    gradient w.r.t. the per-instance mass and inertia next to the cost
    weights and x0: forward + backward + the dynamics-parameter contraction."""
    import torch

    from diffmpc_amd import _lib
    from diffmpc_amd import workloads as W

    dev = torch.device("cuda:0")
    name, prob, sp = W.quadrotor(batch=args.batch, seed=1, horizon=args.horizon)
    solver, params, guess = W.make_solver(name, prob, sp, device=dev)
    keys = ["weights_penalization_reference_state_trajectory", "weights_penalization_control_squared",
            "initial_state", "mass", "inertia"]

    def fwd_bwd(with_dyn):
        guess.dual_backward_guess.zero_()
        ks = keys if with_dyn else keys[:3]
        w = {k: params[k].clone().requires_grad_(True) for k in ks}
        sol = solver.solve(guess, params, w)
        loss = (sol.states - params["reference_state_trajectory"]).square().sum()
        return torch.autograd.grad(loss, [w[k] for k in ks])

    lib = _lib.load()
    res = {}
    for with_dyn in (False, True):
        for _ in range(args.warmup):
            fwd_bwd(with_dyn)
        torch.cuda.synchronize()
        _lib.profile_read(_lib.PROF_DYN_GRADS)
        _lib.profile_read(_lib.PROF_BACKWARD)
        lib.dmpc_profile_enable(1)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fwd_bwd(with_dyn)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        lib.dmpc_profile_enable(0)
        dg_ms, dg_n = _lib.profile_read(_lib.PROF_DYN_GRADS)
        bw_ms, bw_n = _lib.profile_read(_lib.PROF_BACKWARD)
        res[with_dyn] = (dt, dg_ms / max(dg_n, 1), bw_ms / max(bw_n, 1))
    B = args.batch
    t_dyn, dg_avg, bw_avg = res[True]
    print(json.dumps({
        "metric": "quadrotor fwd+bwd grads/s incl. d/d(mass, inertia) (T=50 B=1024)",
        "value": B / t_dyn, "unit": "grads/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_dyn * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG, f64 hyper-dual parameter contraction",
        "data": "synthetic",
        "config": {"workload": f"quadrotor nx=13 nu=4 T={args.horizon} B={B}, 5 SQP iterations, grads w.r.t. "
                               "w_x, w_u, x0, mass, inertia", "global_batch": B},
        "without_dynamics_params_ms": res[False][0] * 1e3,
        "dyn_param_grad_kernel_ms": dg_avg, "backward_kernel_ms": bw_avg,
    }), flush=True)


# train_imitation_diffmpc.py's recorded time: 7.04 s per 200 training steps
# (mean of the 20 runs in train_il_diffmpc.pkl; device not recorded)
IL_REFERENCE_MS = 35.2


def bench_imitation(args):
    """The reference's imitation-learning benchmark
    (benchmarking/imitation-learning/train_imitation_diffmpc.py): cart-pole
    N=100, 32 expert demonstrations (collected first, untimed: 501 SQP
    iterations, PCG 1e-24, line search), one training step = Euler rollout of
    the demonstrations as the guess, batched solve (<= 5 SQP iterations,
    tolerance 0.1, PCG 1e-14, line search), imitation loss, backward, weight
    update, plus the model loss -- captured as one hipGraph."""
    import numpy as np
    import torch

    from diffmpc_amd import imitation as il

    dev = torch.device("cuda:0")
    B = args.batch if args.batch != 1024 else 32
    t0 = time.perf_counter()
    x0, us, expert = il.collect_expert_data(num_trajs=B, seed=123, device=dev)
    torch.cuda.synchronize()
    t_expert = time.perf_counter() - t0
    problem = il.cartpole_problem()
    solver = il.SQPSolver(program=problem, params=il.training_solver_params())
    tr = il.ImitationLearning(solver, x0, us, device=dev)
    g = tr.captured_step()  # one replay = one training step, chained through the graph's weight buffer
    w0 = torch.tensor(il.INITIAL_WEIGHTS[0], dtype=torch.float64, device=dev)
    g["w"].copy_(w0)
    for _ in range(args.warmup):
        g["graph"].replay()
    g["w"].copy_(w0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses, model_losses = [], []
    for _ in range(args.steps):
        g["graph"].replay()
        losses.append(g["loss"].clone())
        model_losses.append(g["model_loss"].clone())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    out = {
        "metric": "imitation-learning training step, cart-pole N=100, 32 demonstrations (ms/step)",
        "value": dt * 1e3, "unit": "ms/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": (IL_REFERENCE_MS / (dt * 1e3)) if B == 32 else None,
        "dtype": "f32 Jacobians, f64 dynamics values/Schur/PCG", "data": "synthetic (expert demonstrations solved here "
                                                                     "from the reference's seeded initial states)",
        "config": {"workload": f"cart-pole imitation learning N=100 B={B}, <=5 SQP iterations, tol 0.1, PCG 1e-14, "
                               "line search; one hipGraph per step", "global_batch": B},
        "imitation_loss": float(losses[-1]), "model_loss": float(model_losses[-1]), "weights": g["w"].tolist(),
        "expert_collection_s": t_expert, "expert_max_convergence_error": float(expert.convergence_error.max()),
        "reference_ms_per_step": IL_REFERENCE_MS,
    }
    if not args.no_cpu_baseline:
        from threadpoolctl import threadpool_limits

        from diffmpc_amd import workloads as W
        from oracle import imitation as oil

        m = min(4, B)
        x0n, usn = x0[:m].cpu().numpy(), us[:m].cpu().numpy()
        prob = oil.cartpole_problem(x0n, problem.params)
        sp = W.solver_params(5, 1e-1, 1e-14, linesearch=True)
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            oil.il_loss_and_grad(prob, sp, usn, il.INITIAL_WEIGHTS[0])
            t1 = time.perf_counter()
        out["cpu_baseline"] = {
            "value": (t1 - t0) * 1e3 * B / m, "unit": "ms/step", "cores": 1, "kind": "port",
            "sample": f"one training step's loss and gradient for {m} of the {B} demonstrations, float64 numpy "
                      "oracle, one thread, scaled to the batch",
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
