"""IMPES throughput on one MI355X: the three-phase gas-injection black-oil case at
1M cells (100x100x100, dx=dy=10 m, dz=1 m, sigma_ln 1.5, 9 wells, dt = 1 day),
pressure by FGMRES (FP64) preconditioned with the RS-AMG hierarchy in FP32, to a
1e-8 relative residual, beside the same run with the FP64 hierarchy.

A step is one IMPES time step: TPFA assembly, AMG numeric update + FGMRES solve,
explicit upwind transport with its substeps and ledgers. Inputs are resident in
HBM; the first warmup step builds the AMG hierarchy (pattern frozen afterwards,
as in the reference's cached solver).

Prints ONE JSON line (rank 0). ``--gpus N`` with N > 1 exits: this build runs on
one GPU (see the sharding section of the design notes)."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oil-simulator-py_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "IMPES cell-steps/s @1M cells, FP32-AMG+FP64-GMRES vs FP64 AMG at 1e-8 rel res"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=18)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dims", default="100x100x100")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-hierarchy comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-dims", default="50x50x40")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="concurrent single-threaded oracle copies of the CPU baseline (0: one per available CPU, "
                         "at most 16)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cold-steps", type=int, default=20, help="steps of the cold run (setup included)")
    ap.add_argument("--kernels", default="csr_jacobi_f32,st7_jacobi_f32",
                    help="kernels timed for the roofline figures: the first (the dominant one, the L1-Jacobi "
                         "sweep of the CSR levels) is `roofline`, the others `roofline_other`")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing of the history-match case")
    ap.add_argument("--no-rl", action="store_true", help="skip the batched reinforcement-learning environment leg")
    ap.add_argument("--rl-envs", type=int, default=64, help="copies stepped together by the batched environment")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (FP64) and exit: the known-size kernel that "
                         "calibrates FETCH_SIZE/WRITE_SIZE counters")
    return ap.parse_args()


def make_sim(dims, precision):
    """The benchmark case with the shared solver settings (simulator.cases.BENCH_AMG,
    the ones the GPU parity tests run): FGMRES to a 1e-8 relative residual and an
    estimated relative pressure error below 1e-8."""
    from simulator.cases import BENCH_AMG, baseline_case, build

    cfg, fields = baseline_case(2, dims=dims)
    cfg["simulation"]["amg"] = dict(BENCH_AMG, precision=precision)
    return cfg, build(cfg, fields, "cuda")


def run(dims, precision, steps, warmup, kernel_steps=0, kernels=()):
    """warmup + `steps` timed steps (no instrumentation), then for each name in
    `kernels`, `kernel_steps` more steps with every launch of that kernel
    bracketed by HIP events on its stream (V-cycle graph replay is bypassed while
    a kernel is timed). Returns the simulator, the timed seconds, the FGMRES
    iterations per step, the kernel totals by name and the setup figures: wall
    time of the first warmup step (it builds the AMG hierarchy) and of the
    hierarchy build inside it."""
    import torch
    import impes_hip as H

    cfg, sim = make_sim(dims, precision)
    dt = cfg["simulation"]["time_step_days"] * 86400.0
    setup = {}
    for w in range(warmup):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if not sim.run_step(dt):
            raise RuntimeError("warmup step failed")
        if w == 0:
            torch.cuda.synchronize()
            setup = {"first_step_ms": 1e3 * (time.perf_counter() - t0),
                     "hierarchy_setup_ms": 1e3 * sim._model.amg_setup_seconds()}
    torch.cuda.synchronize()
    its, est = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        if not sim.run_step(dt):
            raise RuntimeError("timed step failed (pressure solve did not converge)")
        its.append((int(sim.last_solve.get("iterations", 0)), int(sim.last_solve.get("checks", 0))))
        est.append(float(sim.last_solve.get("err_estimate", 0.0)))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kts = {}
    for kernel in (kernels if kernel_steps else ()):
        H.ktimer_reset(enable=True, kernel=kernel)
        for _ in range(kernel_steps):
            if not sim.run_step(dt):
                raise RuntimeError("kernel-timing step failed")
        torch.cuda.synchronize()
        kts[kernel] = H.ktimer_read()
        H.ktimer_reset(enable=False)
    setup["max_error_estimate"] = max(est) if est else None
    return sim, el, its, kts, setup


def _cpu_sample(dims, steps):
    """One oracle run of the sample (see cpu_baseline); returns (timed seconds,
    seconds of the hierarchy-building first step, the last pressure system)."""
    sys.path.insert(0, ROOT)
    from oracle import amg as oa, impes as oi
    from simulator.cases import baseline_case

    cfg, fields = baseline_case(2, dims=dims)
    case = oi.Case(cfg, fields, pvt=oi.PVT(cfg["fluid"]["pvt_table"]))
    sim = oi.Simulation(case)
    state = {}

    def solver(A, b):
        if "H" not in state:
            state["H"] = oa.Hierarchy(A, max_levels=25)
        else:
            state["H"].update(A)
        state["A"], state["b"] = A, b
        x, st = oa.fgmres(A, b, x0=None, precond=state["H"].precondition, rtol=1e-8, maxit=400)
        return x

    dt = cfg["simulation"]["time_step_days"] * 86400.0
    t0 = time.perf_counter()
    sim.step(dt, solver=solver)
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(dt, solver=solver)
    return time.perf_counter() - t0, t_first, state["A"], state["b"]


def _cpu_worker(dims, steps):
    """Child process of cpu_baseline: one single-threaded oracle run, its timed
    seconds on stdout."""
    el, _, _, _ = _cpu_sample(dims, steps)
    print(json.dumps({"timed_s": el}), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(dims, steps, procs, splu_dims=(30, 30, 20)):
    """The oracle (CPU restatement, same algorithm as the device path: FP64 AMG +
    FGMRES to 1e-8) on a bounded sample of the workload: a dims-sized grid of the
    same case (same field generator, wells placed proportionally), one untimed
    step that builds the hierarchy, then `steps` timed steps. The oracle is
    single-threaded (numpy/scipy sparse kernels), so the host is measured as
    `procs` independent copies of the sample run at once in child processes, one
    thread each (BLAS threads pinned to 1): value = procs x cells x steps / the
    slowest copy's timed seconds. Beside it: one copy alone, and scipy's sparse
    direct solve (SuperLU) of the sample's last pressure system against the
    oracle's FGMRES-AMG solve of the same system, on a splu_dims grid of the same
    case (SuperLU's fill on the 3D sample itself takes minutes)."""
    import subprocess

    import numpy as np
    import scipy.sparse.linalg as spl

    sys.path.insert(0, ROOT)
    from oracle import amg as oa

    n = int(np.prod(dims))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    args = [sys.executable, os.path.abspath(__file__), "--cpu-worker", "x".join(map(str, dims)),
            "--cpu-steps", str(steps)]
    ps = [subprocess.Popen(args, env=env, stdout=subprocess.PIPE, text=True) for _ in range(procs)]
    times = []
    for p in ps:
        out, _ = p.communicate(timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"CPU baseline worker failed with exit code {p.returncode}")
        times.append(json.loads(out.strip().splitlines()[-1])["timed_s"])
    # one copy alone (this process, BLAS threads as set for it), then the direct solve
    el1, t_first, _, _ = _cpu_sample(dims, steps)
    _, _, A, b = _cpu_sample(splu_dims, 1)
    t0 = time.perf_counter()
    lu = spl.splu(A.tocsc())
    x = lu.solve(b)
    t_lu = time.perf_counter() - t0
    H = oa.Hierarchy(A, max_levels=25)
    t0 = time.perf_counter()
    H.update(A)
    xa, st = oa.fgmres(A, b, x0=None, precond=H.precondition, rtol=1e-8, maxit=400)
    t_amg = time.perf_counter() - t0
    return {"value": procs * n * steps / max(times), "unit": "cell-steps/s", "cores": procs, "kind": "port",
            "sample": f"oracle (numpy/scipy, 1 thread per process) on a {dims[0]}x{dims[1]}x{dims[2]} grid of the "
                      f"same case, {steps} steps after one hierarchy-building step, {procs} copies at once "
                      f"(slowest {max(times):.1f} s)",
            "cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "single_process_value": n * steps / el1, "single_process_first_step_s": t_first,
            "splu_dims": list(splu_dims), "splu_solve_s": t_lu, "fgmres_amg_solve_s": t_amg,
            "splu_vs_fgmres_rel_diff": float(np.linalg.norm(x - xa) / np.linalg.norm(x))}


def cold_run(dims, steps):
    """A fresh simulator of the benchmark case run for `steps` steps from its
    initial state, timed end to end: the first step builds the AMG hierarchy and
    captures the V-cycle graphs, so the rate includes the setup."""
    import torch

    cfg, sim = make_sim(dims, "fp32")
    dt = cfg["simulation"]["time_step_days"] * 86400.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        if not sim.run_step(dt):
            raise RuntimeError("cold-run step failed")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    n = 1
    for d in dims:
        n *= d
    return {"steps": steps, "total_s": el, "cell_steps_per_s_incl_setup": n * steps / el,
            "hierarchy_setup_ms": 1e3 * sim._model.amg_setup_seconds()}


def vcycle_level_table(sim, cycles=20):
    """Per-level figures of the FP32 V-cycle of the benchmark's hierarchy:
    each level's own launches timed with HIP events over `cycles` eager cycles,
    and their algorithmic bytes per visit (row pointers, indices and values of
    every operator pass once, each vector pass once)."""
    ms, by, graph_ms = sim._model.amg_level_times(32, cycles)
    rows = []
    for l, ((n, nnz), t, b) in enumerate(zip(sim.amg_levels(), ms, by)):
        rows.append({"level": l, "rows": n, "nnz": nnz, "us_per_visit": 1e3 * t, "bytes_per_visit": b,
                     "GBs": b / (t * 1e-3) / 1e9 if t > 0 else None})
    return {"levels": rows, "eager_cycle_us": 1e3 * sum(ms), "graph_cycle_us": 1e3 * graph_ms,
            "bytes_per_cycle": sum(by)}


def adjoint_timing(steps=None):
    """The two-phase history-match case (64x64x8, 30 steps of 5 days): wall time of
    one forward run and of the misfit gradient by the hand-written adjoint
    (checkpoints every 5 steps, each segment recomputed), both with the HIP AMG in
    FP64 FGMRES on the GPU, plus the gradient's agreement with a central finite
    difference along one random unit direction of ln k."""
    import numpy as np
    import torch
    from adjoint import twophase as tp
    from adjoint.history import problem

    solve = tp.device_solver(rtol=1e-12)
    P = problem(device="cuda", solve=solve, steps=steps)
    case, lnk, rates, data, sigma, n = P["case"], P["lnk_prior"], P["rates"], P["data"], P["sigma"], P["steps"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    obs, _ = tp.simulate(case, lnk, rates, solve, n)
    J0 = float(tp.misfit(obs, data, sigma))
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, g_lnk, g_rates = tp.gradient(case, lnk, rates, data, sigma, solve, n, checkpoint_every=5)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    v = torch.tensor(np.random.default_rng(3).standard_normal(tuple(lnk.shape)), device=lnk.device)
    v /= v.norm()
    h = 1e-4
    jp = float(tp.misfit(tp.simulate(case, lnk + h * v, rates, solve, n)[0], data, sigma))
    jm = float(tp.misfit(tp.simulate(case, lnk - h * v, rates, solve, n)[0], data, sigma))
    fd = (jp - jm) / (2 * h)
    an = float((g_lnk * v).sum())
    return {"case": f"two-phase history match {'x'.join(map(str, case.dims))}, {n} steps, 5 wells",
            "forward_s": t_fwd, "gradient_s": t_adj, "gradient_over_forward": t_adj / t_fwd,
            "misfit": J0, "fd_rel_err_one_direction": abs(fd - an) / abs(fd), "grad_lnk_norm": float(g_lnk.norm()),
            "grad_rates": [float(x) for x in g_rates]}


def adjoint_config4(steps=20, checkpoint_every=4, fd_dirs=2, fd_h=1e-4):
    """Adjoint at scale: the three-phase gas-injection case at 1M cells, 20 steps
    of 1 day. Truth: the case's ln k field; data: its producers' surface oil,
    water and gas rates with 2% Gaussian noise; the gradient of the rate misfit
    is taken at a prior ln k (another seed) with respect to the 1M per-cell ln k
    values and the 9 well controls, by the hand-written HIP reverse sweep
    (checkpoints every 4 steps, segments recomputed). Forward and transposed
    solves: FP32 RS-AMG inside FP64 FGMRES to 1e-10. Reported: one forward run's
    wall time, the gradient's (forward with checkpoints + reverse sweep), their
    ratio, and the gradient's relative difference from the same adjoint with the
    FP64 hierarchy. Check: central differences at h = 1e-4 (fixed beforehand) of
    the torch restatement (solves to 1e-11) with the base run's decisions
    replayed, along `fd_dirs` random unit directions of ln k."""
    import numpy as np
    import torch
    from adjoint import device as ad
    from adjoint import threephase as tp
    from adjoint.twophase import device_solver
    from simulator.cases import SEEDS

    cfg, fields, lnk_truth, lnk_prior = ad.history_match_case(steps=steps)
    solver = {"precision": 32, "rtol": 1e-10, "adj_rtol": 1e-10, "etol": 0.0, "maxit": 600}
    dev = ad.DeviceAdjoint(cfg, fields, solver=solver)
    truth = torch.tensor(lnk_truth.reshape(-1), dtype=torch.float64, device="cuda")
    data, sigma = ad.rate_data(dev.observe(truth, dev.values0, steps), SEEDS["noise"])
    prior = torch.tensor(lnk_prior.reshape(-1), dtype=torch.float64, device="cuda")
    dev.observe(prior, dev.values0, 1)  # warm: hierarchy build and graphs of the prior's first system
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.observe(prior, dev.values0, steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    J, g32, gv32, obs = dev.gradient(prior, dev.values0, steps, data, sigma, checkpoint_every)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    peak = torch.cuda.max_memory_allocated()
    timing = dict(dev.timing)
    del dev
    dev64 = ad.DeviceAdjoint(cfg, fields, solver=dict(solver, precision=64))
    t0 = time.perf_counter()
    J64, g64, gv64, _ = dev64.gradient(prior, dev64.values0, steps, data, sigma, checkpoint_every)
    torch.cuda.synchronize()
    t_grad64 = time.perf_counter() - t0
    del dev64
    # replayed central differences of the restatement (FP64 hierarchy, solves to 1e-11)
    case = tp.ThreePhaseCase.from_config(cfg, fields, device="cuda")
    solve = device_solver(rtol=1e-11, precision=64)
    dt_, st_ = torch.tensor(data, device="cuda"), torch.tensor(sigma, device="cuda")
    pr = prior.reshape(case.dims)

    def J_of(lnk, branches):
        with torch.no_grad():
            return float(tp.rate_misfit(tp.observe(case, lnk, case.rates0, solve, steps, branches=branches), dt_,
                                        st_))

    base = tp.BranchLog()
    J_t = J_of(pr, base)
    fd = []
    rng = np.random.default_rng(SEEDS["fd"])
    for _ in range(fd_dirs):
        v = torch.tensor(rng.standard_normal(case.dims), device="cuda")
        v /= v.norm()
        jp = J_of(pr + fd_h * v, tp.BranchLog(replay=base))
        jm = J_of(pr - fd_h * v, tp.BranchLog(replay=base))
        f = (jp - jm) / (2 * fd_h)
        an = float((g32.reshape(case.dims) * v).sum())
        fd.append({"fd": f, "adjoint": an, "rel_err": abs(f - an) / max(abs(an), 1e-300)})
    return {"case": f"three-phase gas injection {'x'.join(map(str, _dims(cfg)))}, {steps} steps, 9 wells; "
                    "objective: producer surface-rate misfit (2% noise), gradient at a prior ln k",
            "checkpoint_every": checkpoint_every, "forward_s": t_fwd, "gradient_s": t_grad,
            "gradient_over_forward": t_grad / t_fwd, "gradient_fp64_s": t_grad64, "torch_allocator_peak_gb": peak / 1e9,
            "misfit": J, "misfit_restatement_rel_diff": abs(J_t - J) / abs(J),
            "fp32_vs_fp64_gradient_rel": float((g32 - g64).norm() / g64.norm()),
            "fp32_vs_fp64_controls_rel": float(np.linalg.norm(gv32 - gv64) / np.linalg.norm(gv64)),
            "fd_h": fd_h, "fd_checks": fd, "fd_max_rel_err": max(c["rel_err"] for c in fd),
            "grad_lnk_norm": float(g32.norm()), "grad_controls": [float(x) for x in gv32],
            "substeps": timing.get("substeps"), "forward_in_gradient_s": timing.get("forward_s"),
            "reverse_s": timing.get("reverse_s")}


def _dims(cfg):
    return tuple(int(v) for v in cfg["reservoir"]["dimensions"])


def rl_timing(num_envs, steps=8):
    """Reinforcement-learning environments (rl_env.py) on the 20x20x3 three-phase
    environment: env-steps/s of BatchedReservoirEnv with num_envs copies in one
    device step against ReservoirEnv stepped alone, both with random actions of
    +-5 % and the observation copied to the host every step (as the reference's
    step does). Two warmup steps each after a reset; the reset and the first
    step (which builds the AMG hierarchy, as after every reset) are timed too."""
    import tempfile

    import numpy as np
    import torch
    from rl_env import BatchedReservoirEnv, ReservoirEnv
    from simulator.cases import rl_env_case

    rng = np.random.default_rng(0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "env.json")
        with open(path, "w") as f:
            json.dump(rl_env_case(), f)

        first = {}

        def timed(env, shape):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            env.reset(seed=0)
            env.step(rng.uniform(-0.05, 0.05, shape))  # builds the AMG hierarchy
            torch.cuda.synchronize()
            first[len(first)] = 1e3 * (time.perf_counter() - t0)
            env.step(rng.uniform(-0.05, 0.05, shape))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                out = env.step(rng.uniform(-0.05, 0.05, shape))
                if not out[4]["converged"]:
                    raise RuntimeError("environment step did not converge")
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        benv = BatchedReservoirEnv(path, num_envs)
        tb = timed(benv, (num_envs, benv.wells_per_env))
        senv = ReservoirEnv(path)
        ts = timed(senv, (len(senv.well_manager.wells),))
    batched = num_envs * steps / tb
    single = steps / ts
    return {"case": "three-phase 20x20x3, 4 wells, dt = 1 day, FP32-AMG FGMRES to 1e-8", "num_envs": num_envs,
            "steps": steps, "batched_ms_per_step": tb / steps * 1e3, "single_ms_per_step": ts / steps * 1e3,
            "batched_env_steps_per_s": batched, "single_env_steps_per_s": single, "speedup": batched / single,
            "reset_and_first_step_ms": {"batched": first[0], "single": first[1]}}


def main():
    a = parse()
    if a.cpu_worker:
        _cpu_worker(tuple(int(x) for x in a.cpu_worker.split("x")), a.cpu_steps)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is described in the design notes, not implemented)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import impes_hip as H
    from simulator.cases import BENCH_AMG

    H.load()  # fails loudly when the HIP library is missing
    if a.calibrate_copy:
        n = a.calibrate_copy * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            H.copy_f64(y, x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 8, "calibrate_copy_bytes_written": n * 8}), flush=True)
        return
    dims = tuple(int(x) for x in a.dims.split("x"))
    ncell = int(np.prod(dims))

    sim32, el32, its32, _, setup32 = run(dims, "fp32", a.steps, a.warmup)
    value = ncell * a.steps / el32
    p32 = sim32.fluid.pressure.clone()
    del sim32

    extra = {}
    if not a.no_fp64:
        sim64, el64, its64, _, setup64 = run(dims, "fp64", a.steps, a.warmup)
        p64 = sim64.fluid.pressure
        extra = {"fp64_value": ncell * a.steps / el64, "fp64_ms_per_step": 1e3 * el64 / a.steps,
                 "speedup_fp32_vs_fp64": el64 / el32,
                 "fp32_vs_fp64_pressure_max_rel": float((p32 - p64).abs().max() / p64.abs().max()),
                 "fgmres_iterations_fp64": [i for i, _ in its64], "error_checks_fp64": [c for _, c in its64],
                 "setup_fp64": setup64}
        del sim64

    # rooflines of the dominant kernel and of the others named, on a fresh run of
    # the same workload
    names = [k for k in a.kernels.split(",") if k]
    sim_k, _, _, kts, _ = run(dims, "fp32", 0, a.warmup, kernel_steps=3, kernels=names)
    levels = sim_k.amg_levels()
    extra["vcycle_levels"] = vcycle_level_table(sim_k)
    del sim_k
    if a.cold_steps > 0:
        extra["cold_run"] = cold_run(dims, a.cold_steps)
    tf = os.path.join(ROOT, "profiles", "hbm_traffic_1M_gas_injection.json")
    traffic_table = json.load(open(tf)) if os.path.exists(tf) else {}

    def roofline_of(kt):
        if kt is None or kt["launches"] <= 0:
            return None
        # a kernel that runs on several levels: total algorithmic bytes over total
        # launch time (bytes and time per launch are both averages over the levels)
        per_launch_s = kt["ms"] * 1e-3 / kt["launches"]
        bytes_per_launch = kt["bytes"] / kt["launches"]
        achieved = bytes_per_launch / per_launch_s / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic_table.get(kt["name"], {}).get("hbm_bytes_per_launch"), "kernel": kt["name"],
                "launches": kt["launches"], "avg_us": per_launch_s * 1e6,
                "algorithmic_bytes_per_launch": bytes_per_launch}

    roof = roofline_of(kts.get(names[0])) if names else None
    if len(names) > 1:
        extra["roofline_other"] = {k: roofline_of(kts.get(k)) for k in names[1:]}

    if not a.no_adjoint:
        extra["adjoint"] = adjoint_timing()
        extra["adjoint_three_phase"] = adjoint_config4()
    if not a.no_rl:
        extra["rl_batch"] = rl_timing(a.rl_envs)

    cpu = None
    if not a.no_cpu:
        cdims = tuple(int(x) for x in a.cpu_dims.split("x"))
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cpu = cpu_baseline(cdims, a.cpu_steps, a.cpu_procs or max(1, min(16, avail)))

    out = {"metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": 1e3 * el32 / a.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (FP32 AMG preconditioner)", "data": "synthetic",
           "config": {"workload": "three-phase gas-injection black-oil IMPES, 100x100x100 (1M cells)",
                      "dims": list(dims), "cells": ncell, "dt_days": 1.0, "wells": 9, "rel_residual": 1e-8,
                      "solver": dict(BENCH_AMG), "amg_levels": [list(x) for x in levels],
                      "fgmres_iterations": [i for i, _ in its32], "error_checks": [c for _, c in its32],
                      "parallelism": "single GPU"},
           "roofline": roof, "cpu_baseline": cpu, "setup": setup32}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
