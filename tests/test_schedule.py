"""Launch schedules on the heterogeneous batch (per-world random trees with
uneven contact counts): the persistent work-queue kernel in any order, a
subset queue, and graph-captured horizons all give the per-world path's
results bit for bit (the same per-environment kernel body runs in each)."""

import numpy as np
import pytest

import metalrobo as mr

pytestmark = pytest.mark.gpu


def snapshot(world):
    q, v, s = world.state()
    h, p, c = world.manifolds()
    for e in range(len(c)):
        h[e, c[e]:] = np.zeros(1, h.dtype)
        p[e, c[e]:] = np.zeros(1, p.dtype)
    return q, v, s, h, p, c


def same(a, b):
    return all(np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))
               for x, y in zip(a, b))


def trees(envs=96, seed=4):
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_RANDOM_TREES, envs, seed)
    return model, model.default_config()


def test_work_queue_matches_the_per_world_grid(gpu):
    model, cfg = trees()
    cmds = model.commands(0, 12)
    ref = mr.World(model, cfg)
    ref.advance(cmds)
    ref_status = ref.statuses()
    queue = mr.World(model, cfg)
    order = np.random.default_rng(0).permutation(model.environment_count)
    queue.set_schedule(order, grid=17)  # fewer workgroups than worlds: each takes several
    queue.advance(cmds)
    assert same(snapshot(ref), snapshot(queue))
    np.testing.assert_array_equal(ref_status["code"], queue.statuses()["code"])
    assert ref.contact_statuses()["constraints"].sum() > 0


def test_subset_queue_steps_only_the_listed_worlds(gpu):
    model, cfg = trees(envs=48)
    cmds = model.commands(0, 6)
    full = mr.World(model, cfg)
    full.advance(cmds)
    part = mr.World(model, cfg)
    listed = np.arange(0, model.environment_count, 3)
    part.set_schedule(listed[::-1], grid=8)
    before = snapshot(part)
    out = part.advance(cmds)
    after, stepped = snapshot(part), snapshot(full)
    others = np.setdiff1d(np.arange(model.environment_count), listed)
    for k in range(3):  # q, v, scene
        np.testing.assert_array_equal(after[k][listed], stepped[k][listed])
        np.testing.assert_array_equal(after[k][others], before[k][others])
    # the advance summary counts the listed worlds only
    codes = part.statuses()["code"][:, listed]
    cs = part.contact_statuses()[:, listed]
    assert out.successful_environment_steps + out.failed_environment_steps == len(cmds) * len(listed)
    assert out.failed_environment_steps == int(np.sum(codes != 0))
    assert out.maximum_active_pairs == int(cs["max_active_pairs"].max())
    assert out.maximum_constraints == int(cs["max_constraints"].max())
    if out.failed_environment_steps:
        assert out.first_failing_environment in set(listed.tolist())


def test_graph_horizon_replays_match_eager(gpu):
    import torch

    model, cfg = trees(envs=64)
    steps = 5
    eager = mr.World(model, cfg)
    graph = mr.World(model, cfg)
    graph.set_schedule(np.arange(model.environment_count)[::-1], grid=32)
    cmd_dev = torch.zeros((steps, model.environment_count, model.max_nv), dtype=torch.float32, device="cuda")
    for rep in range(4):
        host = model.commands(rep * steps, steps)
        eager.advance(host)
        cmd_dev.copy_(torch.from_numpy(host))
        graph.enqueue_graph(cmd_dev)
        graph.wait()
        # eager work of the process between replays
        _ = torch.ones(1 << 16, device="cuda").sum().item()
        assert same(snapshot(eager), snapshot(graph)), f"replay {rep}"
        np.testing.assert_array_equal(eager.statuses()["code"], graph.statuses()["code"])


def test_autotuner_picks_a_schedule_that_reproduces_the_plain_path(gpu):
    import torch

    from metalrobo import autotune as at

    model, cfg = trees(envs=256, seed=8)
    world = mr.World(model, cfg)
    world.advance(model.commands(0, 20))
    cmds = torch.from_numpy(model.commands(20, 4)).cuda()
    start = at.snapshot(world)
    world.enqueue_device(cmds)
    world.wait()
    plain_end = snapshot(world)
    at.restore(world, start)
    res = at.autotune(world, cmds, trial_horizons=1)
    assert len(res.trials) == 5 and res.resident_workgroups > 0
    assert res.best.name == min(res.trials, key=lambda t: t["ms_per_step"])["name"]
    world.enqueue_graph(cmds)
    world.wait()
    assert same(snapshot(world), plain_end)


def test_sweep_reordered_queue_matches_the_per_world_grid(gpu):
    """Adaptive residual solves with the queue re-sorted after every control step by the
    sweeps each world took (mr_world_set_schedule_reorder), eager and graph-captured: the
    per-world path's results bit for bit; the schedule must not leak into the physics."""
    import torch

    model, cfg = trees(envs=64, seed=6)
    cfg.residual_tolerance = 1.0e-5
    cfg.max_iterations = 48
    cmds = model.commands(0, 10)
    ref = mr.World(model, cfg)
    ref.advance(cmds)
    ref_snap = snapshot(ref)
    eager = mr.World(model, cfg)
    eager.set_schedule(np.arange(model.environment_count), grid=13)
    eager.set_schedule_reorder(True)
    eager.advance(cmds)
    assert same(ref_snap, snapshot(eager))
    np.testing.assert_array_equal(ref.contact_statuses()["residuals"], eager.contact_statuses()["residuals"])
    sweeps = ref.contact_statuses()["residuals"][..., 3]
    assert sweeps.max() > sweeps.min(), "every world took the same number of sweeps"
    graphed = mr.World(model, cfg)
    graphed.set_schedule(np.arange(model.environment_count), grid=13)
    graphed.set_schedule_reorder(True)
    dev = torch.from_numpy(cmds[:5]).cuda()
    graphed.enqueue_graph(dev)
    graphed.wait()
    dev.copy_(torch.from_numpy(cmds[5:]))
    graphed.enqueue_graph(dev)  # replay of the same capture
    graphed.wait()
    assert same(ref_snap, snapshot(graphed))


def test_reorder_needs_a_full_work_queue(gpu):
    model, cfg = trees(envs=16)
    w = mr.World(model, cfg)
    with pytest.raises(mr.MetalRoboError, match="every environment"):
        w.set_schedule_reorder(True)
    w.set_schedule(np.arange(8), grid=4)
    with pytest.raises(mr.MetalRoboError, match="every environment"):
        w.set_schedule_reorder(True)


def test_learned_warm_start_capture_refuses_a_work_queue(gpu):
    """The MLP instance runs one workgroup per environment; a world with a work-queue
    schedule must refuse it rather than step the environments its order left out."""
    model, cfg = trees(envs=16)
    w = mr.World(model, cfg)
    w.set_schedule(np.arange(8), grid=4)
    with pytest.raises(mr.MetalRoboError, match="work-queue"):
        w.set_mlp_capture(True)


@pytest.mark.parametrize("kind,envs,grid", [(mr.MR_SYNTHETIC_RANDOM_TREES, 64, 24),
                                            (mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 48, 16)])
def test_persistent_horizon_matches_one_launch_per_step(gpu, kind, envs, grid):
    """mr_world_enqueue_device_horizon: every world runs all control steps in one persistent
    launch (fewer workgroups than worlds, reversed order, adaptive residual solve with the
    reordered queue, a reset mask mid-horizon). States, manifolds, statuses and contact
    statuses equal one plain launch per control step bit for bit, and the per-world step
    times are positive and within the launch."""
    import torch

    print(f"horizon: kind {kind}, {envs} worlds, grid {grid}", flush=True)
    model = mr.WorldModel.synthetic(kind, envs, 5)
    cfg = model.default_config()
    cfg.residual_tolerance = 1.0e-5
    cfg.max_iterations = 64
    steps = 6
    host = model.commands(0, steps)
    masks = np.zeros((steps, envs), np.uint32)
    masks[3, 1] = 1
    masks[4, envs - 2] = 1
    ref = mr.World(model, cfg)
    ref.set_step_timing(True)
    ref.advance(host, masks)
    # one launch per step: every world's step ends within its launch
    assert np.all(ref.world_step_microseconds() <= ref.step_milliseconds()[:, None] * 1e3 * 1.01)
    hz = mr.World(model, cfg)
    hz.set_schedule(np.arange(envs)[::-1], grid=grid)
    hz.set_schedule_reorder(True)
    hz.set_step_timing(True)
    cmds = torch.from_numpy(host).cuda()
    m_dev = torch.from_numpy(masks.view(np.int32)).cuda()
    hz.enqueue_device_horizon(cmds, m_dev)
    adv = hz.wait()
    assert same(snapshot(ref), snapshot(hz))
    for a, b in ((ref.statuses(), hz.statuses()), (ref.contact_statuses(), hz.contact_statuses())):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert ref.contact_statuses()["constraints"].sum() > 0
    assert adv.control_step_count == steps and adv.failed_environment_steps == int(
        np.sum(ref.statuses()["code"] != 0))
    us = hz.world_step_microseconds()
    assert us.shape == (steps, envs) and np.all(us > 0.0)
    launch_us = hz.last_kernel_milliseconds() * steps * 1e3  # the per-step average times the steps
    print(f"world step times: max {us.max():.1f} us, largest per-world sum {us.sum(axis=0).max():.1f} us, "
          f"launch {launch_us:.1f} us", flush=True)
    assert us.sum(axis=0).max() <= launch_us * 1.01
    with pytest.raises(mr.MetalRoboError):
        hz.step_milliseconds()
    plain = mr.World(model, cfg)
    with pytest.raises(mr.MetalRoboError, match="work-queue schedule"):
        plain.enqueue_device_horizon(cmds)


def test_one_submission_in_flight_per_world(gpu):
    """A resident world takes one submission at a time (the reference's contextBusy rule for a
    resident continuation): a second enqueue before mr_world_wait is refused and changes
    nothing; after the wait the next one is accepted, for every enqueue form."""
    import torch

    model, cfg = trees(envs=16)
    w = mr.World(model, cfg)
    cmds = torch.from_numpy(model.commands(0, 2)).cuda()
    w.enqueue_device(cmds)
    with pytest.raises(mr.MetalRoboError, match="context busy"):
        w.enqueue_device(cmds)
    with pytest.raises(mr.MetalRoboError, match="context busy"):
        w.advance(model.commands(0, 1))
    w.wait()
    ref = mr.World(model, cfg)
    ref.advance(model.commands(0, 2))
    assert same(snapshot(ref), snapshot(w))  # the refused submissions never ran
    w.set_schedule(np.arange(16), grid=8)
    w.enqueue_device_horizon(cmds)
    with pytest.raises(mr.MetalRoboError, match="context busy"):
        w.enqueue_graph(cmds)
    w.wait()
    w.enqueue_graph(cmds)
    w.wait()
    w.advance(model.commands(0, 1))
