"""Benchmark of the batched temporalCone control step on one MI355X.

Workload (default): the floating 29-DoF humanoid on a 256x256 Perlin
heightfield, 4096 worlds, fp32, 4 microsteps of 5 ms per 20 ms control step,
implicit PD targets U[-0.3, 0.3] rad around a standing pose, elliptic cone
(mu_u 1.0, mu_v 0.7), 4+2 cone sweeps. One "step" is one control step of all
worlds (one launch of the fused kernel); the metric is env-steps/s.

Prints one JSON line (see the keys below). Inputs are generated up front and
resident in HBM before the timed region; the timed region is bracketed by
device synchronisation on both sides.
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "numi-lab_amd"))

HBM_PEAK_GBPS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "control_step_humanoid_4096_traffic.md")


def bytes_per_env_step(nq, nv, s_dyn, s_static, m):
    """Compulsory HBM bytes per env-step: the persistent per-environment
    state (q, v, dynamic bodies, manifolds) read and written every microstep,
    static bodies read, plus control input, reset mask, world status and
    contact status (4 microsteps per control step)."""
    return 4 * (2 * (4 * nq + 4 * nv + 128 * s_dyn + 320 * m) + 128 * s_static) + 4 * nv + 4 + 48 + 288


def read_traffic_profile(path):
    if not os.path.isfile(path):
        return None
    text = open(path).read()
    m = re.search(r"hbm_bytes_per_launch\s*[=:]\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def cpu_baseline(kind, seed, envs, steps, threads, solver=0):
    """Times the oracle (FP32 CPU restatement of the same microstep) on a
    bounded sample of the same workload: `envs` worlds over `steps` control
    steps, one contiguous block of worlds per std::thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as ct

    import metalrobo as mr
    import oracle_lib as ol

    model = mr.WorldModel.synthetic(kind, envs, seed)
    cfg = model.default_config()
    cfg.solver = solver
    st = ol.OracleState(model, cfg)
    cmds = model.commands(0, steps)
    lib = ol.load()
    E = envs
    status = np.zeros(E, mr.WORLD_STATUS_DTYPE)
    cstatus = np.zeros(E, mr.CONTACT_STATUS_DTYPE)
    acc = np.zeros((E, model.max_nv), np.float32)
    p = lambda a: a.ctypes.data  # noqa: E731
    t0 = time.perf_counter()
    for k in range(steps):
        rc = lib.mr_oracle_control_step_parallel(ct.byref(model.desc), ct.byref(cfg), E, k, threads, p(st.q),
                                                 p(st.v), p(st.scene), p(st.mh), p(st.mp), p(st.mcount),
                                                 p(np.ascontiguousarray(cmds[k])), p(status), p(cstatus), p(acc))
        if rc != 0:
            raise RuntimeError("oracle failed")
    dt = time.perf_counter() - t0
    return {"value": round(E * steps / dt, 1), "unit": "env-steps/s", "cores": threads, "kind": "port",
            "sample": f"{E} worlds x {steps} control steps of the same workload (oracle, FP32, "
                      f"{threads} std::threads)", "seconds": round(dt, 2)}


def residual_leg(mr, model, commands, args, reorder=False):
    """The same workload with the adaptive cone solve: sweeps continue until
    the final-sweep max |dlambda| of a microstep is at or below the tolerance
    (capped at --residual-max-iterations sweeps). Reports throughput, control
    step latency and how many microsteps reached the tolerance. With `reorder`,
    the worlds run as a work queue over the resident workgroups, re-sorted on
    the device after every control step by the sweeps each world took, most
    first (the results are the same as the plain launch's)."""
    import torch

    cfg = model.default_config()
    cfg.solver = SOLVERS[args.solver]
    cfg.residual_tolerance = args.residual_tolerance
    cfg.max_iterations = args.residual_max_iterations
    world = mr.World(model, cfg)
    if reorder:
        world.set_schedule(np.arange(model.environment_count), world.resident_workgroups())
        world.set_schedule_reorder(True)
    world.set_step_timing(True)
    warm = min(args.warmup, commands.shape[0])
    if warm > 0:
        world.enqueue_device(commands[:warm])
        world.wait()
    steps = min(args.residual_steps, commands.shape[0] - warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.enqueue_device(commands[warm:warm + steps])
    world.wait()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_ms = world.step_milliseconds()
    cs = world.contact_statuses()
    with_contacts = cs["constraints"] > 0
    res = cs["residuals"][..., 0][with_contacts]
    iters = cs["solver_iterations"][with_contacts]
    return {
        "tolerance": args.residual_tolerance, "max_iterations": args.residual_max_iterations, "steps": int(steps),
        "env_steps_per_s": round(args.envs * steps / (t1 - t0), 1),
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4)},
        "final_microstep_iterations_mean": round(float(iters.mean()), 2) if iters.size else 0.0,
        "final_microstep_at_tolerance": round(float(np.mean(res <= args.residual_tolerance)), 4) if res.size else 1.0,
        "final_sweep_max_dlambda_p99": float(np.percentile(res, 99)) if res.size else 0.0,
    }


def warm_start_leg(mr, args):
    """Learned warm start on the chain-and-boxes workload: capture (features, converged impulse)
    samples from the plain adaptive solve, train the MLP with the library's own kernels, then run
    the same settled worlds at the residual tolerance with the plain and the learned warm start
    (plain fallback on) and report sweeps per control step."""
    import torch
    from metalrobo import warmstart as ws

    tol = args.residual_tolerance if args.residual_tolerance > 0.0 else 1.0e-5
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, args.warm_start_collect_worlds, args.seed)
    cfg = model.default_config()
    cfg.residual_tolerance = tol
    cfg.max_iterations = args.residual_max_iterations
    t0 = time.perf_counter()
    f, t = ws.collect_samples(model, cfg, control_steps=args.warm_start_collect_steps, settle_steps=40,
                              max_samples=args.warm_start_samples)
    collect_s = time.perf_counter() - t0
    mlp = ws.WarmStartMLP(seed=args.seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = mlp.fit(f, t, steps=args.warm_start_train_steps, batch=8192, lr=1.0e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del f, t
    ev_model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, args.warm_start_eval_worlds, args.seed + 1)
    ev_cfg = ev_model.default_config()
    ev_cfg.residual_tolerance = tol
    ev_cfg.max_iterations = args.residual_max_iterations
    ev = ws.evaluate(ev_model, ev_cfg, mlp, control_steps=args.warm_start_eval_steps, settle_steps=40)
    plain, learned = ev["plain"]["sweeps_per_step"], ev["mlp"]["sweeps_per_step"]
    return {
        "workload": "chain7_boxes20_plane (config 1)", "residual_tolerance": tol,
        "samples": rep.kept + rep.dropped, "samples_kept": rep.kept, "collect_s": round(collect_s, 2),
        "train_steps": rep.steps, "train_batch": 8192, "train_s": round(train_s, 2),
        "loss_first": round(rep.first_loss, 5), "loss_final": round(rep.final_loss, 5),
        "eval": ev, "sweeps_saved_fraction": round(1.0 - learned / plain, 4) if plain > 0 else None,
    }


def heterogeneous_leg(mr, args):
    """Heterogeneous batch: per-world random trees (6-40 DoF) with 0-63 free bodies, settled
    so the free bodies have landed; one horizon timed on the plain per-world path (eager
    launches, one workgroup per world) and on the autotuned schedule replayed from a captured
    graph, from the same state. The two end states must agree bit for bit."""
    import torch
    from metalrobo import autotune as at

    E = args.hetero_envs
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_RANDOM_TREES, E, args.seed)
    cfg = model.default_config()
    world = mr.World(model, cfg)
    world.advance(model.commands(0, args.hetero_settle))
    horizon = args.hetero_horizon
    cmds = torch.from_numpy(model.commands(args.hetero_settle, horizon)).to("cuda:0")
    start = at.snapshot(world)
    # plain per-world path
    world.set_schedule(None, 0)
    times = []
    for _ in range(2):
        at.restore(world, start)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        world.enqueue_device(cmds)
        world.wait()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    plain_s = min(times)
    plain_end = at.snapshot(world)
    cstat = world.contact_statuses()
    # autotuned schedule on graph-captured horizons
    at.restore(world, start)
    t0 = time.perf_counter()
    tuned = at.autotune(world, cmds, trial_horizons=2)
    tune_s = time.perf_counter() - t0
    times = []
    for _ in range(2):
        at.restore(world, start)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        world.enqueue_graph(cmds)
        world.wait()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    tuned_s = min(times)
    tuned_end = at.snapshot(world)
    same = all(np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
               for a, b in zip(tuned_end[:3], plain_end[:3]))
    h = model.models()
    return {
        "workload": f"random_trees_{E}", "worlds": E, "dof_range": [int(h["nv"].min()), int(h["nv"].max())],
        "bodies_range": [int(h["body_count"].min()), int(h["body_count"].max())],
        "constraints_per_world_range": [int(cstat["constraints"][-1].min()), int(cstat["constraints"][-1].max())],
        "horizon_steps": horizon,
        "plain_env_steps_per_s": round(E * horizon / plain_s, 1),
        "tuned_env_steps_per_s": round(E * horizon / tuned_s, 1),
        "speedup": round(plain_s / tuned_s, 4),
        "schedule": {"name": tuned.best.name, "grid": tuned.best.grid,
                     "resident_workgroups": tuned.resident_workgroups, "trials": tuned.trials,
                     "autotune_s": round(tune_s, 2)},
        "bit_exact_vs_plain": bool(same),
        "failed_env_steps": int((world.statuses()["code"] != 0).sum()),
        "failed_by_status": {STATUS_NAMES.get(int(c), str(int(c))): int(n) for c, n in
                             zip(*np.unique(world.statuses()["code"], return_counts=True)) if int(c) != 0},
    }


SOLVERS = {"sweeps": 0, "jacobi": 1}  # MRSolverMode
STATUS_NAMES = {1: "fixed_budget_complete", 2: "nonfinite_input", 3: "nonfinite_result", 4: "pair_capacity_overflow",
                5: "contact_capacity_overflow", 6: "manifold_capacity_overflow", 7: "constraint_capacity_overflow",
                8: "island_capacity_overflow", 9: "ccd_event_budget_exhausted", 10: "factorization_failed",
                11: "did_not_converge", 12: "unsupported", 13: "ccd_capacity_overflow"}


def server_leg(mr, args):
    """Step server on the 29-DoF humanoid: Poisson arrivals at 200 requests/s, N log-uniform in
    [1, 1024] worlds, horizons 10-500 control steps, a resident pool admitted by a 200 GB HBM
    budget; p50/p99 per-control-step latency and aggregate env-steps/s, next to the same first
    requests served one after another with plain per-request launches."""
    from metalrobo import server as sv

    pool = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, args.serve_pool, args.seed)
    srv = sv.StepServer(pool, hbm_budget_bytes=200e9)
    reqs = sv.poisson_requests(200.0, args.serve_requests, seed=args.seed)
    rep = srv.serve(reqs)
    base_reqs = sv.poisson_requests(200.0, args.serve_requests, seed=args.seed)[:args.serve_serial_requests]
    base = sv.serve_serially(lambda n: mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, n, args.seed),
                             base_reqs)
    return {
        "workload": "humanoid29 requests, Poisson 200/s, N log-uniform 1-1024, horizon 10-500",
        "requests": rep.requests, "pool_slots": int(pool.environment_count), "capacity_slots": rep.capacity_slots,
        "bytes_per_slot": round(rep.bytes_per_slot, 1), "hbm_budget_bytes": rep.hbm_budget_bytes,
        "env_steps": rep.env_steps, "wall_s": round(rep.wall_s, 3),
        "env_steps_per_s": round(rep.env_steps_per_s, 1),
        "latency_ms": {"p50": round(rep.latency_ms_p50, 4), "p99": round(rep.latency_ms_p99, 4)},
        "ticks": rep.ticks, "mean_active_worlds": round(rep.mean_active_worlds, 1),
        "max_active_worlds": rep.max_active_worlds,
        "queue_wait_s": {"p50": round(rep.queue_wait_s_p50, 4), "p99": round(rep.queue_wait_s_p99, 4)},
        "serial_plain": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in base.items()},
        "speedup_vs_serial": round(rep.env_steps_per_s / base["env_steps_per_s"], 3)
        if base["env_steps_per_s"] > 0 else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--envs", type=int, default=4096)
    ap.add_argument("--config", choices=["humanoid", "chain"], default="humanoid")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-envs", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=500)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-cpu-baseline", action="store_true")
    ap.add_argument("--residual-tolerance", type=float, default=1.0e-5,
                    help="final-sweep max |dlambda| target of the adaptive leg (0 disables the leg)")
    ap.add_argument("--residual-max-iterations", type=int, default=128)
    ap.add_argument("--residual-steps", type=int, default=50)
    ap.add_argument("--stage-profile", type=int, default=0, metavar="STEPS",
                    help="after the timed region, run STEPS more control steps with the per-stage shader-clock "
                         "counters on and report each stage's share of the cycles")
    ap.add_argument("--solver", choices=list(SOLVERS), default="sweeps",
                    help="cone-solve iteration: ordered block sweeps or the relaxed block Jacobi")
    ap.add_argument("--skip-server", action="store_true", help="skip the step-server leg")
    ap.add_argument("--serve-pool", type=int, default=16384)
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-serial-requests", type=int, default=16)
    ap.add_argument("--skip-heterogeneous", action="store_true",
                    help="skip the heterogeneous random-tree leg (work queue, graphs, autotuner)")
    ap.add_argument("--hetero-envs", type=int, default=8192)
    ap.add_argument("--hetero-settle", type=int, default=30)
    ap.add_argument("--hetero-horizon", type=int, default=50)
    ap.add_argument("--skip-warm-start", action="store_true",
                    help="skip the learned-warm-start leg (chain-and-boxes workload)")
    ap.add_argument("--warm-start-samples", type=int, default=10_000_000)
    ap.add_argument("--warm-start-collect-worlds", type=int, default=4096)
    ap.add_argument("--warm-start-collect-steps", type=int, default=40)
    ap.add_argument("--warm-start-train-steps", type=int, default=1500)
    ap.add_argument("--warm-start-eval-worlds", type=int, default=16384)
    ap.add_argument("--warm-start-eval-steps", type=int, default=10)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the timed region, run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic that scales FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()

    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py measures one GPU only (--gpus 1): this build runs on a single GPU. Worlds are "
              "independent, so a multi-GPU run would shard them with no exchange step.", file=sys.stderr)
        sys.exit(2)

    import torch

    import metalrobo as mr

    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    kind = mr.MR_SYNTHETIC_HUMANOID_TERRAIN if args.config == "humanoid" else mr.MR_SYNTHETIC_CHAIN_BOXES
    model = mr.WorldModel.synthetic(kind, args.envs, args.seed)
    cfg = model.default_config()
    cfg.solver = SOLVERS[args.solver]
    world = mr.World(model, cfg)
    world.set_step_timing(True)
    total = args.warmup + args.steps
    commands = torch.from_numpy(model.commands(0, total)).to("cuda:0")
    torch.cuda.synchronize()

    if args.warmup > 0:
        world.enqueue_device(commands[:args.warmup])
        world.wait()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.enqueue_device(commands[args.warmup:])
    adv = world.wait()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = args.envs * args.steps / (t1 - t0)

    step_ms = world.step_milliseconds()
    cstat = world.contact_statuses()
    stat = world.statuses()
    mean_manifolds = float(cstat["manifolds"].mean())
    h = model.models()[0]
    nq, nv = int(model.max_nq), int(model.max_nv)
    bodies = model.bodies()
    scene = bodies[h["link_count"]:h["body_count"]]
    s_dyn = int(np.sum(scene["motion_type"] == 2))
    s_static = int(len(scene) - s_dyn)
    b_env = bytes_per_env_step(nq, nv, s_dyn, s_static, mean_manifolds)
    avg_launch_s = float(step_ms.mean()) * 1e-3
    achieved = args.envs * b_env / avg_launch_s / 1e9
    traffic = read_traffic_profile(TRAFFIC_PROFILE)

    out = {
        "metric": "env-steps/s (29-DoF, 4096 worlds) vs plain; p99 step latency at residual 1e-5"
        if args.config == "humanoid" else "env-steps/s (7-DoF chain + 20 boxes)",
        "value": round(value, 1),
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded worlds and commands; no datasets)",
        "config": {
            "workload": ("humanoid29_heightfield256_pd_elliptic" if args.config == "humanoid"
                         else "chain7_boxes20_plane_effort"),
            "worlds": args.envs, "control_steps": args.steps, "microsteps_per_step": int(cfg.physics_substeps),
            "control_dt_s": float(cfg.control_timestep),
            "cone_sweeps": f"{cfg.velocity_iterations}+{cfg.final_velocity_iterations}",
            "solver": args.solver,
            "eligible_pairs_per_world": int(h["pair_count"]), "plain_path": True,
        },
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4),
                       "max": round(float(step_ms.max()), 4)},
        "residual": {"final_sweep_max_dlambda_p99": float(np.percentile(cstat["residuals"][..., 0], 99)),
                     "normal_p99": float(np.percentile(cstat["residuals"][..., 1], 99)),
                     "cone_violation_max": float(cstat["residuals"][..., 2].max())},
        "contacts": {"mean_manifolds": round(mean_manifolds, 3),
                     "mean_constraints": round(float(cstat["constraints"].mean()), 3),
                     "max_constraints": int(cstat["max_constraints"].max()),
                     "failed_env_steps": int(np.sum(stat["code"] != 0))},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 6), "traffic": traffic,
                     "algorithmic_bytes_per_env_step": round(b_env, 1),
                     "kernel_ms_per_launch": round(avg_launch_s * 1e3, 4)},
        "cpu_baseline": None,
    }
    if args.stage_profile > 0:
        prof_world = mr.World(model, cfg)
        prof_world.enqueue_device(commands[:args.warmup])
        prof_world.wait()
        prof_world.set_stage_profiling(True)
        n = min(args.stage_profile, commands.shape[0] - args.warmup)
        prof_world.enqueue_device(commands[args.warmup:args.warmup + n])
        prof_world.wait()
        cyc = prof_world.stage_cycles().astype(np.float64)[:, :len(mr.World.STAGE_NAMES)]
        out["stage_cycles"] = {
            "steps": int(n), "cycles_per_env_step": round(float(cyc.sum() / args.envs / n), 0),
            "fraction": {name: round(float(f), 4) for name, f in zip(mr.World.STAGE_NAMES, cyc.sum(0) / cyc.sum())},
        }
        del prof_world
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.residual_tolerance > 0.0 and args.residual_steps > 0:
        out["at_residual"] = residual_leg(mr, model, commands, args)
        out["at_residual_reordered"] = residual_leg(mr, model, commands, args, reorder=True)
    if not args.skip_heterogeneous:
        out["heterogeneous"] = heterogeneous_leg(mr, args)
    if not args.skip_server:
        out["server"] = server_leg(mr, args)
    if not args.skip_warm_start:
        del world
        out["warm_start"] = warm_start_leg(mr, args)
    if not args.skip_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(kind, args.seed, args.cpu_envs, args.cpu_steps, threads, SOLVERS[args.solver])
    print(json.dumps(out), flush=True)
    del adv


if __name__ == "__main__":
    main()
