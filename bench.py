"""Benchmark of the batched temporalCone control step on one MI355X.

Workload (default): the floating 29-DoF humanoid on a 256x256 Perlin
heightfield, 4096 worlds, fp32, 4 microsteps of 5 ms per 20 ms control step,
implicit PD targets U[-0.3, 0.3] rad around a standing pose, elliptic cone
(mu_u 1.0, mu_v 0.7). The cone solve runs at the metric's residual: sweeps
continue until a microstep's final-sweep max |dlambda| is at or below 1e-5
(capped at 128 sweeps). One "step" is one control step of all worlds; the
metric is env-steps/s.

The worlds are first rolled into steady contact for a fixed number of control
steps (--preroll, not controlled by --warmup), so that any --steps measures
the same phase of the horizon. From that saved state three paths run the same
commands:
  * batched (the headline `value`): the persistent work-queue kernel over all
    worlds, re-sorted on the device after every control step by the sweeps
    each world took, heaviest first;
  * plain at the same residual: one workgroup per world, hardware order;
  * plain at the fixed 4+2 sweep budget.
`vs_plain` is batched / plain at residual 1e-5. Extended legs (heterogeneous
batch, step server at the same residual, learned warm start) follow.

Prints one JSON line. Inputs are generated up front and resident in HBM
before the timed region; the timed region is bracketed by device
synchronisation on both sides.
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
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "control_horizon_queue_humanoid_4096_traffic.md")


def bytes_per_env_step(nq, nv, s_dyn, s_static, m):
    """Compulsory HBM bytes per env-step: the persistent per-environment
    state (q, v, dynamic bodies, manifolds) read and written every microstep,
    static bodies read, plus control input, reset mask, world status and
    contact status (4 microsteps per control step)."""
    return 4 * (2 * (4 * nq + 4 * nv + 128 * s_dyn + 320 * m) + 128 * s_static) + 4 * nv + 4 + 48 + 288


def read_traffic_profile(path):
    """HBM bytes per env-step of the headline kernel from the committed counter profile (FETCH_SIZE
    + WRITE_SIZE of the driver's command, scaled by the copy calibration), and the mean manifolds
    per world of the run it was measured on."""
    if not os.path.isfile(path):
        return None, None
    text = open(path).read()
    m = re.search(r"hbm_bytes_per_env_step\s*[=:]\s*([0-9.eE+]+)", text)
    k = re.search(r"mean_manifolds\s*[=:]\s*([0-9.eE+]+)", text)
    return (float(m.group(1)) if m else None), (float(k.group(1)) if k else None)


def host_cores():
    """Threads the CPU baseline uses and how they were chosen: the process affinity mask,
    bounded by the host's share for one GPU (the OMP_NUM_THREADS the GPU host sets, 16 when
    unset)."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(affinity, share)), {"hardware_concurrency": os.cpu_count(), "affinity_mask_cpus": affinity,
                                          "cpu_share_per_gpu": share}


def cpu_baseline(kind, seed, envs, preroll, steps, threads, cfg_of):
    """Times the oracle (FP32 CPU restatement of the same microstep) on a bounded sample of the
    same workload: `envs` worlds rolled `preroll` control steps into steady contact (untimed),
    then `steps` control steps timed, one contiguous block of worlds per std::thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as ct

    import metalrobo as mr
    import oracle_lib as ol

    model = mr.WorldModel.synthetic(kind, envs, seed)
    cfg = cfg_of(model)
    st = ol.OracleState(model, cfg)
    cmds = model.commands(0, preroll + steps)
    lib = ol.load()
    E = envs
    status = np.zeros(E, mr.WORLD_STATUS_DTYPE)
    cstatus = np.zeros(E, mr.CONTACT_STATUS_DTYPE)
    acc = np.zeros((E, model.max_nv), np.float32)
    p = lambda a: a.ctypes.data  # noqa: E731

    def run(k):
        rc = lib.mr_oracle_control_step_parallel(ct.byref(model.desc), ct.byref(cfg), E, k, threads, p(st.q),
                                                 p(st.v), p(st.scene), p(st.mh), p(st.mp), p(st.mcount),
                                                 p(np.ascontiguousarray(cmds[k])), p(status), p(cstatus), p(acc))
        if rc != 0:
            raise RuntimeError("oracle failed")

    for k in range(preroll):
        run(k)
    t0 = time.perf_counter()
    for k in range(preroll, preroll + steps):
        run(k)
    dt = time.perf_counter() - t0
    return {"value": round(E * steps / dt, 1), "unit": "env-steps/s", "cores": threads, "kind": "port",
            "sample": f"{E} worlds x {steps} control steps of the same workload after a {preroll}-step pre-roll "
                      f"(oracle, FP32, same cone-solve configuration, {threads} std::threads)",
            "seconds": round(dt, 2)}


def timed_run(world, commands, first, steps, envs):
    """Runs commands[first:first+steps] eagerly (per-launch HIP events on the world's stream)
    between two device synchronisations; returns throughput, per-step latencies and statuses."""
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.enqueue_device(commands[first:first + steps])
    world.wait()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_ms = world.step_milliseconds()
    return {"seconds": t1 - t0, "env_steps_per_s": envs * steps / (t1 - t0), "step_ms": step_ms,
            "world_step_us": world.world_step_microseconds(), "cstat": world.contact_statuses(),
            "stat": world.statuses()}


def timed_horizon(world, commands, first, steps, horizon, envs):
    """Runs commands[first:first+steps] as persistent horizon launches of `horizon` control steps
    (each world runs its steps back to back inside a launch); returns throughput, the per-world
    per-step wall-clock times the kernel stamps, and the statuses of every step."""
    import torch

    step_us, cstat, stat, kernel_ms = [], [], [], 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c0 in range(first, first + steps, horizon):
        n = min(horizon, first + steps - c0)
        world.enqueue_device_horizon(commands[c0:c0 + n])
        world.wait()
        kernel_ms += world.last_kernel_milliseconds() * n  # the library reports the per-step average
        step_us.append(world.world_step_microseconds())
        cstat.append(world.contact_statuses())
        stat.append(world.statuses())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return {"seconds": t1 - t0, "env_steps_per_s": envs * steps / (t1 - t0), "kernel_ms": kernel_ms,
            "world_step_us": np.concatenate(step_us), "cstat": np.concatenate(cstat), "stat": np.concatenate(stat)}


def summarize(run, steps, tol):
    cs = run["cstat"]
    with_contacts = cs["constraints"] > 0
    res = cs["residuals"][..., 0][with_contacts]
    iters = cs["solver_iterations"][with_contacts]
    out = {
        "env_steps_per_s": round(run["env_steps_per_s"], 1), "steps": int(steps),
        "ms_per_step": round(run["seconds"] * 1e3 / steps, 4),
        "final_microstep_iterations_mean": round(float(iters.mean()), 2) if iters.size else 0.0,
        "final_sweep_max_dlambda_p99": float(np.percentile(res, 99)) if res.size else 0.0,
        "mean_manifolds": round(float(cs["manifolds"].mean()), 3),
        "failed_env_steps": int(np.sum(run["stat"]["code"] != 0)),
    }
    if tol > 0.0:
        out["final_microstep_at_tolerance"] = round(float(np.mean(res <= tol)), 4) if res.size else 1.0
    if "step_ms" in run:
        # what a caller waiting on each control step of every world sees: the launch
        out["launch_latency_ms"] = {"p50": round(float(np.percentile(run["step_ms"], 50)), 4),
                                    "p99": round(float(np.percentile(run["step_ms"], 99)), 4),
                                    "max": round(float(run["step_ms"].max()), 4)}
    if "world_step_us" in run:
        # each world's own control step, start to published statuses, on the GPU's constant clock
        w = run["world_step_us"] * 1e-3
        out["world_step_latency_ms"] = {"p50": round(float(np.percentile(w, 50)), 4),
                                        "p99": round(float(np.percentile(w, 99)), 4),
                                        "max": round(float(w.max()), 4)}
    return out


_STARTED = time.perf_counter()


def progress(leg):
    """One stderr line per finished leg, so a long run shows where it is."""
    print(f"bench: {leg} done at {time.perf_counter() - _STARTED:.1f} s", file=sys.stderr, flush=True)


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
    saved = 1.0 - learned / plain if plain > 0 else 0.0
    return {
        # the learned warm start is switched on only where it saves sweeps at equal residual; its
        # kernel instance runs one wave per SIMD, so a network that saves none only costs time
        "enabled": bool(saved > 0.0),
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


def task_leg(mr, args):
    """The task layer either side of the step, driven by a device policy: 4096 humanoid worlds run
    observe -> dense-layer policy (actor 256-128 ELU, stochastic head; the pack's critic of the same
    shape runs only when rollouts are recorded, which this leg does not) ->
    action sampling and scoring -> apply_actions -> control step -> complete, all on the world's
    stream. The synthetic policy has random weights of that shape. The same worlds stepped with
    plain PD targets and no task give the layer's cost."""
    import torch
    from metalrobo import packs as pk
    from metalrobo import task as tk

    E, K, W = args.task_envs, args.task_steps, 5
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, E, args.seed)
    cfg = model.default_config()
    world = mr.World(model, cfg)
    prog = tk.humanoid_locomotion_program(model, seed=args.seed)
    task = mr.LocomotionTask(world, prog)
    O, A = task.history_length * task.observation_count, task.action_count
    rng = np.random.default_rng(args.seed)

    def layer(i, o, act):
        return pk.DenseLayer((rng.standard_normal((o, i)) / np.sqrt(i)).astype(np.float32),
                             (0.05 * rng.standard_normal(o)).astype(np.float32), act)

    elu, ident = pk.ACTIVATION_ELU, pk.ACTIVATION_IDENTITY
    pack = pk.PolicyPack(
        id="bench_policy", revision=1, contract=(1, 1, 2, 3, 4),
        layers=[layer(O, 256, elu), layer(256, 128, elu), layer(128, A, ident)],
        observation_mean=np.zeros(O, np.float32), observation_inverse_std=np.ones(O, np.float32),
        critic_layers=[layer(O, 256, elu), layer(256, 128, elu), layer(128, 1, ident)],
        action_log_std=np.full(A, -1.0, np.float32), action_bias=np.zeros(A, np.float32),
        action_scale=np.ones(A, np.float32), observation_clip=5.0, action_clip=1.0)
    task.set_policy(pack)
    for _ in range(W):
        task.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        task.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _, tr, _ = task.read()
    # the task enqueues its control steps itself: read the last step's statuses through the C ABI
    from metalrobo import native as nat
    stat = np.zeros(E, nat.WORLD_STATUS_DTYPE)
    mr.load().check(mr.load().lib.mr_world_read_statuses(world._h, stat.ctypes.data, stat.size))
    task.close()
    world.close()
    plain = mr.World(model, cfg)
    plain.set_step_timing(True)
    commands = torch.from_numpy(model.commands(0, W + K)).cuda()
    plain.enqueue_device(commands[:W])
    plain.wait()
    run = timed_run(plain, commands, W, K, E)
    plain.close()
    return {"workload": "humanoid29_heightfield locomotion task with a device policy", "worlds": E, "steps": K,
            "policy": {"actor": [O, 256, 128, A], "critic": [O, 256, 128, 1], "critic_evaluated": False,
                       "activation": "elu",
                       "stochastic": True, "weights": "random (synthetic)"},
            "env_steps_per_s": round(E * K / dt, 1), "ms_per_step": round(1000.0 * dt / K, 4),
            "world_only_env_steps_per_s": round(run["env_steps_per_s"], 1),
            # a throughput ratio, not a kernel share: the world-only run steps PD-target commands
            # (different physics) with per-step timing events
            "throughput_ratio_vs_world_only": round((E * K / dt) / run["env_steps_per_s"], 4),
            "failed_env_steps_last_step": int((stat["code"] != 0).sum()),
            "terminated_last_step": int((tr["termination"] != 0).sum())}


def ccd_leg(mr, args):
    """Hybrid continuous collision on the chain-and-boxes worlds with every free box CCD-enabled and
    thrown at the ground (horizontal U(-8, 8) m/s, downward U(20, 30) m/s): the same control steps
    from the same state in the speculative and the hybrid mode. Reports throughput, the impact
    segments the hybrid mode splits microsteps at, status codes, and the deepest penetration seen.
    The hybrid mode transfers the thrown boxes' momentum into the light, undamped chain links that
    the speculative mode lets them pass through; a few chains are then spun past the microstep's
    stability limit for explicit velocity terms (the workload's own failure mechanism), and those
    worlds fail non-finite and roll back."""
    import ctypes as ct

    import torch
    from metalrobo import native as nat

    E, K = args.ccd_envs, args.ccd_steps
    base = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, E, args.seed)
    bcfg = base.default_config()
    shapes = base.shapes().copy()
    boxes = shapes["shape_type"] == 2
    shapes["flags"][boxes] |= 2  # MR_SHAPE_FLAG_ENABLE_CCD
    desc = nat.WorldModelDescC.from_buffer_copy(base.desc)
    desc.shapes = shapes.ctypes.data
    b = mr.load()
    h = b.lib.mr_world_model_create(ct.byref(desc), bcfg.control_timestep / bcfg.physics_substeps)
    if not h:
        raise mr.MetalRoboError(b.last_error())
    model = mr.WorldModel(h, b)
    q, v, scene = base.initial_state()
    rng = np.random.default_rng(args.seed)
    free = scene["flags_indices"][..., 0] == 2  # dynamic scene bodies
    n = int(free.sum())
    lin = scene["linear_velocity_inv_mass"]
    lin[free, 0] = rng.uniform(-8.0, 8.0, n)
    lin[free, 1] = rng.uniform(-8.0, 8.0, n)
    lin[free, 2] = -rng.uniform(20.0, 30.0, n)
    scene["linear_velocity_inv_mass"] = lin
    commands = torch.from_numpy(base.commands(0, K)).cuda()
    out = {"workload": "chain7_boxes20_plane, boxes CCD-enabled and thrown at 20-30 m/s", "worlds": E, "steps": K}
    for mode, name in ((0, "speculative"), (1, "hybrid")):
        cfg = bcfg.copy()
        cfg.ccd_mode = mode
        cfg.max_ccd_events = 8
        cfg.max_ccd_iterations = 16
        cfg.ccd_distance_tolerance = 1.0e-3
        cfg.ccd_time_tolerance = 1.0e-5
        cfg.ccd_minimum_advance = 1.0e-6
        world = mr.World(model, cfg)
        world.set_state(q, v, scene)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        impacts, codes, pens = 0.0, np.zeros(16, np.int64), []
        for k in range(K):
            world.enqueue_device(commands[k:k + 1])
            world.wait()
            cs, st = world.contact_statuses()[0], world.statuses()[0]
            impacts += float(cs["diagnostics"][:, 3].sum())
            ok = st["code"] == 0  # a failed step's contact status describes a rolled-back state
            pens.append(cs["diagnostics"][ok, 0])
            codes += np.bincount(st["code"], minlength=16)[:16]
        pens = np.concatenate(pens)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        world.close()
        out[name] = {"env_steps_per_s": round(E * K / dt, 1), "ms_per_step": round(1000.0 * dt / K, 3),
                     "impact_segments_per_env_step": round(impacts / (E * K), 4),
                     "penetration_m_successful_steps": {"p50": round(float(np.percentile(pens, 50)), 5),
                                                        "p99": round(float(np.percentile(pens, 99)), 5),
                                                        "max": round(float(pens.max()), 5)},
                     "status_codes": {str(c): int(x) for c, x in enumerate(codes) if x}, "thrown_boxes": n}
    return out


def server_leg(mr, args):
    """Step server on the 29-DoF humanoid: Poisson arrivals at 200 requests/s, N log-uniform in
    [1, 1024] worlds, horizons 10-500 control steps, a resident pool admitted by a 200 GB HBM
    budget; p50/p99 per-control-step latency and aggregate env-steps/s, next to the same first
    requests served one after another with plain per-request launches."""
    from metalrobo import server as sv

    pool = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, args.serve_pool, args.seed)
    srv = sv.StepServer(pool, headline_config(pool, args), hbm_budget_bytes=200e9)
    reqs = sv.poisson_requests(200.0, args.serve_requests, seed=args.seed)
    rep = srv.serve(reqs)
    base_reqs = sv.poisson_requests(200.0, args.serve_requests, seed=args.seed)[:args.serve_serial_requests]
    base = sv.serve_serially(lambda n: mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, n, args.seed),
                             base_reqs, config_of=lambda m: headline_config(m, args))
    return {
        "workload": "humanoid29 requests, Poisson 200/s, N log-uniform 1-1024, horizon 10-500",
        "residual_tolerance": args.residual_tolerance, "budget_blocked_ticks": rep.budget_blocked_ticks,
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


def headline_config(model, args, fixed_budget=False):
    """The metric's cone-solve configuration: sweeps until the final-sweep max |dlambda| is at or
    below --residual-tolerance (capped at --residual-max-iterations), or the fixed 4+2 budget."""
    cfg = model.default_config()
    cfg.solver = SOLVERS[args.solver]
    if not fixed_budget and args.residual_tolerance > 0.0:
        cfg.residual_tolerance = args.residual_tolerance
        cfg.max_iterations = args.residual_max_iterations
    return cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--preroll", type=int, default=200,
                    help="control steps that bring the worlds into steady contact before --warmup and the "
                         "timed steps (fixed, so that any --steps measures the same phase)")
    ap.add_argument("--horizon", type=int, default=50,
                    help="control steps per persistent launch of the headline path (0: one launch per step)")
    ap.add_argument("--envs", type=int, default=4096)
    ap.add_argument("--config", choices=["humanoid", "chain"], default="humanoid")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-envs", type=int, default=256)
    ap.add_argument("--cpu-preroll", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-cpu-baseline", action="store_true")
    ap.add_argument("--residual-tolerance", type=float, default=1.0e-5,
                    help="final-sweep max |dlambda| the cone solve runs to (0: the fixed 4+2 budget)")
    ap.add_argument("--residual-max-iterations", type=int, default=128)
    ap.add_argument("--stage-profile", type=int, default=0, metavar="STEPS",
                    help="after the timed region, run STEPS more control steps with the per-stage shader-clock "
                         "counters on and report each stage's share of the cycles")
    ap.add_argument("--solver", choices=list(SOLVERS), default="sweeps",
                    help="cone-solve iteration: ordered block sweeps or the relaxed block Jacobi")
    ap.add_argument("--skip-comparisons", action="store_true",
                    help="time only the headline path (no plain-at-residual or fixed-budget legs)")
    ap.add_argument("--skip-server", action="store_true", help="skip the step-server leg")
    ap.add_argument("--skip-task", action="store_true", help="skip the policy-driven task-layer leg")
    ap.add_argument("--task-envs", type=int, default=4096)
    ap.add_argument("--task-steps", type=int, default=50)
    ap.add_argument("--skip-ccd", action="store_true", help="skip the hybrid continuous-collision leg")
    ap.add_argument("--ccd-envs", type=int, default=4096)
    ap.add_argument("--ccd-steps", type=int, default=20)
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
        print("bench.py measures one GPU only (--gpus 1): this library runs on a single GPU. Worlds are "
              "independent, so a multi-GPU run would shard them with no exchange step.", file=sys.stderr)
        sys.exit(2)

    import torch

    import metalrobo as mr
    from metalrobo import autotune as at

    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    kind = mr.MR_SYNTHETIC_HUMANOID_TERRAIN if args.config == "humanoid" else mr.MR_SYNTHETIC_CHAIN_BOXES
    E, K, W, P = args.envs, args.steps, args.warmup, args.preroll
    model = mr.WorldModel.synthetic(kind, E, args.seed)
    cfg = headline_config(model, args)
    tol = float(cfg.residual_tolerance)
    commands = torch.from_numpy(model.commands(0, P + W + K)).to("cuda:0")
    torch.cuda.synchronize()

    # pre-roll into steady contact (same configuration; results do not depend on the schedule)
    world = mr.World(model, cfg)
    if P > 0:
        world.enqueue_device(commands[:P])
        world.wait()
    start = at.snapshot(world)

    # headline: the batched path, a persistent work queue over every world (re-sorted by
    # sweeps after each launch), each launch running every world through --horizon steps
    H = args.horizon if args.horizon > 0 else 1
    world.set_schedule(np.arange(E), world.resident_workgroups())
    world.set_schedule_reorder(True)
    world.set_step_timing(True)
    if W > 0:
        # warmup on the per-step queue kernel: the horizon kernel then launches only in the timed region
        world.enqueue_device(commands[P:P + W])
        world.wait()
    if args.horizon > 0:
        head = timed_horizon(world, commands, P + W, K, H, E)
        launches = (K + H - 1) // H
        avg_launch_s = head["kernel_ms"] * 1e-3 / launches
        steps_per_launch = K / launches
        kernel_name = "mr_control_horizon_queue_kernel"
    else:
        head = timed_run(world, commands, P + W, K, E)
        avg_launch_s = float(head["step_ms"].mean()) * 1e-3
        steps_per_launch = 1.0
        kernel_name = "mr_control_step_queue_kernel"
    head_sum = summarize(head, K, tol)
    ms_per_step = head["seconds"] * 1e3 / K
    value = head["env_steps_per_s"]
    cstat, stat = head["cstat"], head["stat"]
    mean_manifolds = float(cstat["manifolds"].mean())

    legs = {}
    if not args.skip_comparisons:
        variants = [("plain_at_residual", cfg, False), ("queue_per_step_at_residual", cfg, True),
                    ("plain_fixed_budget", headline_config(model, args, True), False)]
        for name, c, queue in variants:
            if tol <= 0.0 and name != "plain_fixed_budget":
                continue
            w = mr.World(model, c)
            at.restore(w, start)
            if queue:
                w.set_schedule(np.arange(E), w.resident_workgroups())
                w.set_schedule_reorder(True)
            w.set_step_timing(True)
            if W > 0:
                w.enqueue_device(commands[P:P + W])
                w.wait()
            legs[name] = summarize(timed_run(w, commands, P + W, K, E), K, float(c.residual_tolerance))
            w.close()
    plain = legs.get("plain_at_residual")

    h = model.models()[0]
    nq, nv = int(model.max_nq), int(model.max_nv)
    bodies = model.bodies()
    scene = bodies[h["link_count"]:h["body_count"]]
    s_dyn = int(np.sum(scene["motion_type"] == 2))
    s_static = int(len(scene) - s_dyn)
    b_env = bytes_per_env_step(nq, nv, s_dyn, s_static, mean_manifolds)
    achieved = E * steps_per_launch * b_env / avg_launch_s / 1e9
    traffic_per_env_step, traffic_m = read_traffic_profile(TRAFFIC_PROFILE)
    traffic = (round(traffic_per_env_step * E * steps_per_launch, 1)
               if traffic_per_env_step is not None and args.horizon > 0 else None)
    p99 = head_sum["world_step_latency_ms"]["p99"]
    vs_plain = round(value / plain["env_steps_per_s"], 4) if plain else None

    out = {
        "metric": "env-steps/s (29-DoF, 4096 worlds) vs plain; p99 step latency at residual 1e-5"
        if args.config == "humanoid" else "env-steps/s (7-DoF chain + 20 boxes)",
        "value": round(value, 1),
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "vs_plain": vs_plain,
        "dtype": "fp32",
        "data": "synthetic (seeded worlds and commands; no datasets)",
        "config": {
            "workload": ("humanoid29_heightfield256_pd_elliptic" if args.config == "humanoid"
                         else "chain7_boxes20_plane_effort"),
            "worlds": E, "control_steps": K, "preroll_steps": P, "microsteps_per_step": int(cfg.physics_substeps),
            "control_dt_s": float(cfg.control_timestep), "solver": args.solver,
            "residual_tolerance": tol, "max_iterations": int(cfg.max_iterations) if tol > 0.0 else None,
            "cone_sweeps": None if tol > 0.0 else f"{cfg.velocity_iterations}+{cfg.final_velocity_iterations}",
            "path": (f"batched: persistent work queue over every world, each launch running every world "
                     f"through {H} control steps back to back, re-sorted by sweeps after each launch"
                     if args.horizon > 0 else
                     "batched: persistent work queue over every world, re-sorted by sweeps after each step"),
            "horizon_steps_per_launch": H if args.horizon > 0 else 1,
            "eligible_pairs_per_world": int(h["pair_count"]),
        },
        # per-control-step latency of each world (start of its step to its published statuses)
        "latency_ms": head_sum["world_step_latency_ms"],
        "target": {"vs_plain": 3.0, "p99_ms": 2.0,
                   "met": bool(vs_plain is not None and vs_plain >= 3.0 and p99 < 2.0)},
        "batched": {k: v for k, v in head_sum.items() if k not in ("world_step_latency_ms",)},
        "residual": {"final_sweep_max_dlambda_p99": head_sum["final_sweep_max_dlambda_p99"],
                     "final_microstep_at_tolerance": head_sum.get("final_microstep_at_tolerance"),
                     "final_microstep_iterations_mean": head_sum["final_microstep_iterations_mean"],
                     "normal_p99": float(np.percentile(cstat["residuals"][..., 1], 99)),
                     "cone_violation_max": float(cstat["residuals"][..., 2].max())},
        "contacts": {"mean_manifolds": round(mean_manifolds, 3),
                     "mean_constraints": round(float(cstat["constraints"].mean()), 3),
                     "max_constraints": int(cstat["max_constraints"].max()),
                     "failed_env_steps": int(np.sum(stat["code"] != 0))},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 6), "traffic": traffic,
                     "kernel": kernel_name,
                     "traffic_mean_manifolds": traffic_m, "traffic_bytes_per_env_step": traffic_per_env_step,
                     "algorithmic_bytes_per_env_step": round(b_env, 1),
                     "algorithmic_bytes_per_launch": round(E * steps_per_launch * b_env, 1),
                     "kernel_ms_per_launch": round(avg_launch_s * 1e3, 4)},
        "cpu_baseline": None,
    }
    out.update(legs)
    if args.stage_profile > 0:
        prof_world = mr.World(model, cfg)
        at.restore(prof_world, start)
        prof_world.set_stage_profiling(True)
        n = min(args.stage_profile, K + W)
        prof_world.enqueue_device(commands[P:P + n])
        prof_world.wait()
        cyc = prof_world.stage_cycles().astype(np.float64)[:, :len(mr.World.STAGE_NAMES)]
        out["stage_cycles"] = {
            "steps": int(n), "cycles_per_env_step": round(float(cyc.sum() / E / n), 0),
            "fraction": {name: round(float(f), 4) for name, f in zip(mr.World.STAGE_NAMES, cyc.sum(0) / cyc.sum())},
        }
        prof_world.close()
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    world.close()
    del commands
    progress("headline")
    for name, skip, leg in (("heterogeneous", args.skip_heterogeneous, heterogeneous_leg),
                            ("server", args.skip_server, server_leg),
                            ("warm_start", args.skip_warm_start, warm_start_leg),
                            ("task", args.skip_task, task_leg),
                            ("ccd", args.skip_ccd, ccd_leg)):
        if not skip:
            out[name] = leg(mr, args)
            progress(name)
    if not args.skip_cpu_baseline:
        threads, how = host_cores()
        if args.cpu_threads:
            threads = args.cpu_threads
        out["cpu_baseline"] = cpu_baseline(kind, args.seed, args.cpu_envs, args.cpu_preroll, args.cpu_steps,
                                           threads, lambda m: headline_config(m, args))
        out["cpu_baseline"].update(how)
        progress("cpu_baseline")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
