"""Task layer (include/mr_task.h) and policy layers.

CPU tests pin the oracle restatement (oracle/oracle_task.cpp) with checks a
plausible mistake fails:
  * SplitMix64: the published outputs for seeds 0 and 1234567;
  * observations of roots at closed-form rotations: projected gravity
    (0, -sin t, -cos t) for a roll t, tilt t, local angular and linear velocity
    R^T w, root height, joint errors, scales and offsets, gravity normalization
    after noise;
  * rewards: yaw-frame tracking of a yawed root moving along its heading is
    exact (exp(0)), with a known lateral error exp(-d^2/width); integration by
    the control step; failure penalty; termination priority; timeout; physics
    error;
  * actions: the first-order response filter a (1 - (1 - f)^n) with
    f = 1 - exp(-dt/tau), the delay slot, clamping to the binding bounds;
  * command segments: resampled when the countdown reaches one, else counted down;
  * dense layer against float64 numpy; sample log probability against the
    Gaussian density of the returned latent.
GPU tests run the HIP kernels through the C ABI against the oracle: the task
loop (apply_actions -> physics -> complete -> observe) on the humanoid with
episodes ending by timeout and termination, observations, states and world
state bit for bit, rewards within 1e-5 relative (expf, atan2f of the device
library against the host's); the dense layer bit for bit for identity and
ReLU, within 4e-7 relative for tanh, ELU and SiLU; the sample within 1e-5."""

import math

import numpy as np
import pytest

import metalrobo as mr
import oracle_lib as ol
from metalrobo import task as tk


def humanoid(envs=4, seed=1):
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, envs, seed)
    return model, model.default_config()


def quat_axis(axis, angle):
    s = math.sin(angle / 2.0)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, math.cos(angle / 2.0)], np.float32)


# ------------------------------------------------------------------ CPU pins

def test_splitmix64_published_outputs():
    lib = ol._task_lib()
    assert lib.mr_oracle_mix64(0) == 0xE220A8397B1DCDAF
    assert lib.mr_oracle_mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4  # second output of seed 0
    assert lib.mr_oracle_mix64(1234567) == 6457827717110365317
    assert lib.mr_oracle_mix64(1234567 + 0x9E3779B97F4A7C15) == 3203168211198807973
    u = np.array([lib.mr_oracle_task_random_unit(7, e, 1, 0, 32) for e in range(4096)])
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.02


def _observe_once(model, cfg, program, q, v):
    ot = ol.OracleTask(model, cfg, program)
    ot.physics.q[:] = q
    ot.physics.v[:] = v
    ot.states["status"][:, 0] = 1  # running episode: observe the world state, not the reset state
    ot.states["status"][:, 1] = 0
    ot.observe()
    return ot


def test_observations_of_closed_form_root_states():
    model, cfg = humanoid(envs=2)
    prog = tk.humanoid_locomotion_program(model, noise=False)
    q = np.tile(prog["default_q"], (2, 1))
    v = np.zeros((2, model.max_nv), np.float32)
    t = 0.3
    q[0, 3:7] = quat_axis((1, 0, 0), t)  # roll by t
    v[0, 3:6] = (0.0, 0.0, 1.0)          # yaw rate in world
    v[0, 0:3] = (0.5, 0.0, 0.0)
    q[0, 7] += 0.125                      # first joint off its nominal angle
    v[0, 6] = 2.0
    ot = _observe_once(model, cfg, prog, q, v)
    frame = ot.observations[0, -1]
    s, c = math.sin(t), math.cos(t)
    np.testing.assert_allclose(frame[0:3], 0.25 * np.array([0.0, s, c]), atol=2e-7)  # R^T w, scaled
    np.testing.assert_allclose(frame[3:6], [0.0, -s, -c], atol=2e-7)                 # projected gravity
    A = len(prog["actions"])
    assert abs(frame[9] - 0.125) < 1e-6                                              # joint position error
    assert abs(frame[9 + A] - 0.05 * 2.0) < 1e-7                                     # joint velocity x 0.05
    np.testing.assert_allclose(frame[-2], 2.0 * 0.5, atol=1e-6)                      # R^T v (x unchanged by roll)
    np.testing.assert_allclose(frame[-1], q[0, 2] - 0.75, atol=1e-6)                 # root height with offset
    # the episode's command, scaled (2, 2, 0.25)
    cmd = ot.states["command_and_phase"][0, :3]
    np.testing.assert_allclose(frame[6:9], cmd * np.array([2.0, 2.0, 0.25], np.float32), rtol=1e-7)
    # env 1 is upright and at rest
    np.testing.assert_allclose(ot.observations[1, -1, 3:6], [0.0, 0.0, -1.0], atol=1e-7)


def test_projected_gravity_is_normalized_after_noise():
    model, cfg = humanoid(envs=3)
    prog = tk.humanoid_locomotion_program(model, noise=True)
    ot = ol.OracleTask(model, cfg, prog)
    g = ot.observations[:, -1, 3:6].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-6)
    assert np.any(np.abs(g[:, :2]) > 1e-4)  # the noise moved it off the axis before normalization


def _complete_once(model, cfg, prog, q, v, codes=None, steps_done=0, schedule=5):
    ot = ol.OracleTask(model, cfg, prog)
    E = model.environment_count
    ot.physics.q[:] = q
    ot.physics.v[:] = v
    ot.states["status"][:, 0] = 1
    ot.states["status"][:, 1] = 0
    ot.states["episode"][:, 0] = steps_done
    ot.states["schedule"][:, 0] = schedule
    ot.previous[:, 0, :] = v[:, 6:6 + ot.A]
    ot.complete(np.zeros(E, np.uint32) if codes is None else codes)
    return ot


def test_tracking_reward_in_the_yaw_frame_and_integration():
    model, cfg = humanoid(envs=2)
    prog = tk.humanoid_locomotion_program(model, noise=False)
    only = lambda names: np.array([r for r in prog["rewards"] if r["source"][0] in names], prog["rewards"].dtype)
    prog["rewards"] = only((0,))
    prog["terminations"] = prog["terminations"][:0]
    psi, u, d = 0.7, 0.6, 0.2
    q = np.tile(prog["default_q"], (2, 1))
    v = np.zeros((2, model.max_nv), np.float32)
    q[:, 3:7] = quat_axis((0, 0, 1), psi)
    v[0, 0:2] = (u * math.cos(psi), u * math.sin(psi))                       # along the heading
    v[1, 0:2] = (u * math.cos(psi) - d * math.sin(psi), u * math.sin(psi) + d * math.cos(psi))  # plus lateral d
    ot = ol.OracleTask(model, cfg, prog)
    ot.physics.q[:] = q
    ot.physics.v[:] = v
    ot.states["status"][:, 0] = 1
    ot.states["status"][:, 1] = 0
    ot.states["schedule"][:, 0] = 5
    ot.states["command_and_phase"][:, :3] = (u, 0.0, 0.0)
    ot.previous[:, 0, :] = 0.0
    ot.complete(np.zeros(2, np.uint32))
    dt = float(cfg.control_timestep)
    r = ot.transitions["reward_and_state"][:, 0]
    np.testing.assert_allclose(r[0], 1.0 * dt, rtol=1e-6)                     # exp(0), weight 1, times dt
    np.testing.assert_allclose(r[1], math.exp(-d * d / 0.25) * dt, rtol=2e-5)
    np.testing.assert_allclose(ot.transitions["reward_breakdown0"][:, 0], r, rtol=1e-7)


def test_terminations_priorities_penalty_timeout_and_physics_error():
    model, cfg = humanoid(envs=4)
    prog = tk.humanoid_locomotion_program(model, noise=False)
    prog["rewards"] = prog["rewards"][:0]
    prog["max_episode_steps"] = 10
    q = np.tile(prog["default_q"], (4, 1))
    v = np.zeros((4, model.max_nv), np.float32)
    q[0, 2] = 0.2                                  # below 0.35: height
    q[1, 2] = 0.2
    q[1, 3:7] = quat_axis((0, 1, 0), 1.3)          # and tilted beyond 1.0 rad: tilt has the higher priority
    ot = _complete_once(model, cfg, prog, q, v, codes=np.array([0, 0, 0, 3], np.uint32), steps_done=0)
    term = ot.transitions["termination"]
    np.testing.assert_array_equal(term[:, 0], [1, 1, 0, 1])
    np.testing.assert_array_equal(term[:, 3], [1, 2, 0, 5])
    np.testing.assert_allclose(ot.transitions["reward_and_state"][:3, 0], [-1.0, -1.0, 0.0], atol=1e-7)
    assert ot.transitions["reward_and_state"][3, 0] == 0.0          # no penalty for a physics error
    np.testing.assert_allclose(ot.transitions["reward_and_state"][1, 3], 1.3, atol=1e-6)  # tilt of a pitched root
    np.testing.assert_array_equal(ot.states["status"][:, 1], [1, 1, 0, 1])
    # the step that reaches max_episode_steps is a timeout, without penalty
    q2 = np.tile(prog["default_q"], (4, 1))
    ot2 = _complete_once(model, cfg, prog, q2, np.zeros_like(v), steps_done=9)
    np.testing.assert_array_equal(ot2.transitions["termination"][:, 1], [1, 1, 1, 1])
    np.testing.assert_array_equal(ot2.transitions["termination"][:, 3], [4, 4, 4, 4])


def test_command_segments_count_down_then_resample():
    model, cfg = humanoid(envs=3)
    prog = tk.humanoid_locomotion_program(model, noise=False)
    prog["rewards"] = prog["rewards"][:0]
    prog["terminations"] = prog["terminations"][:0]
    q = np.tile(prog["default_q"], (3, 1))
    v = np.zeros((3, model.max_nv), np.float32)
    ot = ol.OracleTask(model, cfg, prog)
    before = ot.states.copy()
    ot.physics.q[:] = q
    ot.physics.v[:] = v
    ot.states["schedule"][:, 0] = (1, 2, 7)
    ot.complete(np.zeros(3, np.uint32))
    s = ot.states
    np.testing.assert_array_equal(s["schedule"][1:, 0], [1, 6])
    np.testing.assert_array_equal(s["command_and_phase"][1:, :3], before["command_and_phase"][1:, :3])
    dt = float(cfg.control_timestep)
    assert 2.0 / dt - 1 <= s["schedule"][0, 0] <= 6.0 / dt
    assert not np.array_equal(s["command_and_phase"][0, :3], before["command_and_phase"][0, :3])
    lo, hi = np.array(prog["command_lower"][:3]), np.array(prog["command_upper"][:3])
    c = s["command_and_phase"][0, :3]
    assert np.all((c == 0.0)) or (np.all(c >= lo) and np.all(c <= hi))
    np.testing.assert_allclose(s["command_and_phase"][:, 3], 2.0 * np.float32(np.pi) * dt / 0.8, rtol=1e-6)


def test_action_filter_delay_and_clamp():
    model, cfg = humanoid(envs=1)
    tau, delay = 0.05, 1
    prog = tk.humanoid_locomotion_program(model, noise=False, response_time=tau, actuator_delay=delay, delay_slots=4)
    ot = ol.OracleTask(model, cfg, prog)
    dt = float(cfg.control_timestep)
    f = 1.0 - math.exp(-dt / tau)
    A = ot.A
    a = np.zeros((1, A), np.float32)
    a[0, 0] = 0.8
    a[0, 1] = 100.0  # far beyond the joint limit: clamped
    targets = []
    for n in range(4):
        ot.apply_actions(a)
        targets.append(ot.targets[0].copy())
    scale = 0.25
    j0 = prog["actions"][0]
    q0 = prog["default_q"][j0["indices"][2]]
    # one step of delay: the first step sees the empty slot, then the filter approaches 0.8
    for n, t in enumerate(targets):
        filtered = 0.8 * (1.0 - (1.0 - f) ** n) if n > 0 else 0.0
        assert abs(t[j0["indices"][3]] - (q0 + scale * filtered)) < 2e-6, (n, t[j0["indices"][3]])
    j1 = prog["actions"][1]
    assert targets[-1][j1["indices"][3]] == np.float32(j1["parameters"][2])
    assert np.all(targets[-1][:6] == 0.0)  # the root has no drive


def test_dense_layer_and_sample_against_closed_forms():
    rng = np.random.default_rng(3)
    E, I, O = 7, 13, 9
    w = rng.standard_normal((O, I)).astype(np.float32)
    b = rng.standard_normal(O).astype(np.float32)
    x = rng.standard_normal((E, I)).astype(np.float32)
    mean = rng.standard_normal(I).astype(np.float32) * 0.1
    inv = (1.0 + rng.random(I)).astype(np.float32)
    ref = np.clip((x.astype(np.float64) - mean) * inv, -2.0, 2.0) @ w.astype(np.float64).T + b
    for act, fn in ((0, lambda z: z), (1, lambda z: np.maximum(z, 0)), (2, np.tanh),
                    (3, lambda z: np.where(z >= 0, z, np.exp(z) - 1)), (4, lambda z: z / (1 + np.exp(-z)))):
        out = ol.policy_dense(w, b, x, O, act, mean, inv, 2.0)
        np.testing.assert_allclose(out, fn(ref), rtol=2e-5, atol=2e-5)
    states = np.zeros(E, mr.TASK_STATE_DTYPE)
    states["episode"][:, 0] = np.arange(E)
    states["episode"][:, 1] = 3
    A = 4
    mu = rng.standard_normal((E, A)).astype(np.float32)
    ls = np.array([-0.5, 0.0, 0.3, -1.0], np.float32)
    acts, lat, lp = ol.policy_sample(mu, np.zeros(A), np.ones(A), ls, 1.0, 11, 2, states, 5)
    z = (lat - mu) / np.exp(ls)
    expect = np.sum(-0.5 * z * z - ls - 0.5 * math.log(2 * math.pi), axis=1)
    np.testing.assert_allclose(lp, expect, rtol=1e-4, atol=1e-5)
    np.testing.assert_array_equal(acts, np.clip(lat, -1.0, 1.0))
    acts2, lat2, lp2 = ol.policy_sample(mu, np.zeros(A), np.ones(A), None, 1.0, 11, 2, states, 5)
    np.testing.assert_array_equal(lat2, mu)
    assert np.all(lp2 == 0.0)
    assert np.std(z) > 0.5  # the sample is not the mean


def test_task_create_refuses_unsupported_programs_on_the_cpu_abi():
    """Creation validates before touching the GPU only where the world exists; here the
    struct layout the binding builds is checked against the C header's."""
    import ctypes as ct

    from metalrobo.native import LocomotionTaskProgramC
    assert ct.sizeof(LocomotionTaskProgramC) == 160
    assert LocomotionTaskProgramC.seed.offset == 152
    assert LocomotionTaskProgramC.command_seconds.offset == 144
    assert mr.TASK_STATE_DTYPE.itemsize == 176 and mr.TASK_TRANSITION_DTYPE.itemsize == 128


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
def test_gpu_task_loop_matches_oracle(gpu):
    import torch

    E, steps = 12, 9
    model, cfg = humanoid(envs=E, seed=5)
    prog = tk.humanoid_locomotion_program(model, seed=9, response_time=0.03, actuator_delay=1)
    prog["max_episode_steps"] = 4                               # timeouts and restarts inside the horizon
    prog["terminations"][0]["parameters"][0] = np.float32(prog["base_height_target"] - 0.01)  # early falls
    print(f"task loop: {E} humanoid worlds, {steps} steps", flush=True)
    world = mr.World(model, cfg)
    task = mr.LocomotionTask(world, prog)
    otask = ol.OracleTask(model, cfg, prog)
    obs, tr, st = task.read()
    assert np.array_equal(obs.view(np.uint32), otask.observations.view(np.uint32))
    assert np.array_equal(st.view(np.uint8), otask.states.view(np.uint8))
    rng = np.random.default_rng(2)
    ended = 0
    for k in range(steps):
        a = rng.uniform(-1.2, 1.2, (E, task.action_count)).astype(np.float32)
        task.step(torch.from_numpy(a).cuda())
        otask.step(a, control_step=0)
        obs, tr, st = task.read()
        q, v, _ = world.state()
        assert np.array_equal(q.view(np.uint32), otask.physics.q.view(np.uint32)), f"step {k}: q"
        assert np.array_equal(v.view(np.uint32), otask.physics.v.view(np.uint32)), f"step {k}: v"
        # every task-state lane is exact except the episode return and tracking sums, which
        # accumulate expf results of the device and host libraries
        ost = otask.states
        for field in st.dtype.names:
            if field == "air_return_tracking":
                assert np.array_equal(st[field][:, :2].view(np.uint32), ost[field][:, :2].view(np.uint32)), k
                np.testing.assert_allclose(st[field][:, 2:], ost[field][:, 2:], rtol=1e-5, atol=1e-7)
            else:
                assert np.array_equal(st[field].view(np.uint32), ost[field].view(np.uint32)), (k, field)
        np.testing.assert_array_equal(tr["termination"], otask.transitions["termination"])
        for field in ("reward_and_state", "reward_breakdown0", "reward_breakdown1"):
            np.testing.assert_allclose(tr[field], otask.transitions[field], rtol=1e-5, atol=1e-7, err_msg=field)
        assert np.array_equal(obs.view(np.uint32), otask.observations.view(np.uint32)), f"step {k}: observations"
        ended += int(tr["termination"][:, 0].sum())
    assert ended > 0 and np.all(st["episode"][:, 1] >= 2)  # episodes ended and restarted


@pytest.mark.gpu
def test_gpu_policy_layers_match_oracle(gpu):
    import torch

    rng = np.random.default_rng(4)
    E, I, O = 300, 294, 256
    w = (rng.standard_normal((O, I)) / math.sqrt(I)).astype(np.float32)
    b = (0.1 * rng.standard_normal(O)).astype(np.float32)
    x = rng.standard_normal((E, I)).astype(np.float32)
    mean = (0.1 * rng.standard_normal(I)).astype(np.float32)
    inv = (1.0 + rng.random(I)).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    for act in range(5):
        out = torch.zeros((E, O), dtype=torch.float32, device="cuda")
        mr.native.policy_dense_layer(dev(w), dev(b), dev(x), out, act, dev(mean), dev(inv), 5.0)
        got = out.cpu().numpy()
        ref = ol.policy_dense(w, b, x, O, act, mean, inv, 5.0)
        if act in (0, 1):
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), act
        else:
            np.testing.assert_allclose(got, ref, rtol=4e-7, atol=1e-7, err_msg=str(act))
    # no normalization, a strided output
    out = torch.zeros((E, O + 3), dtype=torch.float32, device="cuda")
    mr.native.policy_dense_layer(dev(w), dev(b), dev(x), out, 0)
    ref = ol.policy_dense(w, b, x, O, 0)
    assert np.array_equal(out.cpu().numpy()[:, :O].view(np.uint32), ref.view(np.uint32))
    # ragged shapes: every neuron-tile width (64 / 32 / 16 / 8), partial
    # environment tiles, feature counts off the 16-feature chunk, one feature
    for e, i, o in ((1, 7, 1), (65, 294, 29), (1100, 128, 128), (513, 17, 9), (300, 33, 64), (70, 1, 3),
                    (4099, 40, 16), (129, 16, 257)):
        wr = (rng.standard_normal((o, i)) / math.sqrt(i)).astype(np.float32)
        br = (0.1 * rng.standard_normal(o)).astype(np.float32)
        xr = rng.standard_normal((e, i)).astype(np.float32)
        mr_ = (0.1 * rng.standard_normal(i)).astype(np.float32)
        ir = (1.0 + rng.random(i)).astype(np.float32)
        for act, norm in ((0, True), (1, False)):
            out = torch.full((e, o), float("nan"), dtype=torch.float32, device="cuda")
            if norm:
                mr.native.policy_dense_layer(dev(wr), dev(br), dev(xr), out, act, dev(mr_), dev(ir), 1.5)
                ref = ol.policy_dense(wr, br, xr, o, act, mr_, ir, 1.5)
            else:
                mr.native.policy_dense_layer(dev(wr), dev(br), dev(xr), out, act)
                ref = ol.policy_dense(wr, br, xr, o, act)
            assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32)), (e, i, o, act)
    # sample and score
    A = 29
    states = np.zeros(E, mr.TASK_STATE_DTYPE)
    states["episode"][:, 0] = rng.integers(0, 100, E)
    states["episode"][:, 1] = rng.integers(1, 9, E)
    mu = rng.standard_normal((E, A)).astype(np.float32)
    ls = (-1.0 + 0.2 * rng.standard_normal(A)).astype(np.float32)
    bias = (0.1 * rng.standard_normal(A)).astype(np.float32)
    scale = (0.5 + rng.random(A)).astype(np.float32)
    acts = torch.zeros((E, A), dtype=torch.float32, device="cuda")
    lat = torch.zeros_like(acts)
    lp = torch.zeros(E, dtype=torch.float32, device="cuda")
    st_dev = torch.from_numpy(states.view(np.uint8).copy()).cuda()
    mr.native.policy_sample_and_score(dev(mu), dev(bias), dev(scale), dev(ls), 1.0, 77, 3, st_dev, 12, acts, lat, lp)
    ra, rl, rp = ol.policy_sample(mu, bias, scale, ls, 1.0, 77, 3, states, 12)
    np.testing.assert_allclose(lat.cpu().numpy(), rl, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(acts.cpu().numpy(), ra, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(lp.cpu().numpy(), rp, rtol=1e-5, atol=1e-5)


@pytest.mark.gpu
def test_gpu_task_refuses_unsupported_programs(gpu):
    model, cfg = humanoid(envs=2)
    world = mr.World(model, cfg)
    prog = tk.humanoid_locomotion_program(model)
    bad = dict(prog)
    bad["observations"] = prog["observations"].copy()
    bad["observations"][0]["source"][0] = 17  # masked depth: not in the subset
    with pytest.raises(mr.MetalRoboError, match="unsupported opcode"):
        mr.LocomotionTask(world, bad)
    bad = dict(prog)
    bad["rewards"] = prog["rewards"].copy()
    bad["rewards"][0]["source"][0] = 11  # mechanical power
    with pytest.raises(mr.MetalRoboError, match="unsupported opcode"):
        mr.LocomotionTask(world, bad)
    bad = dict(prog)
    bad["actions"] = prog["actions"].copy()
    bad["actions"][0]["indices"][2] = 8  # not the DoF's coordinate
    with pytest.raises(mr.MetalRoboError, match="coordinate"):
        mr.LocomotionTask(world, bad)
    bad = dict(prog)
    bad["delay_slots"] = 2
    with pytest.raises(mr.MetalRoboError, match="delay_slots"):
        mr.LocomotionTask(world, bad)
