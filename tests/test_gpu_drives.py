"""Actuator commands on the GPU: effort commands and implicit PD drives with
active motor profiles (torque-speed envelope, MetalWorld.metal:252-269,
345-365), dry friction and effort limits, stepped by the HIP kernel and
compared with the oracle bit for bit."""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol

ACTUATED, POSITION_LIMIT, EFFORT_LIMIT, DRIVE = 2, 4, 16, 32


def motor_chain(envs=5):
    """A fixed-base 4-link chain hanging under gravity; joints 0 and 2 have
    active motor profiles, joint 1 only an effort limit, joint 3 a dry
    friction drive."""
    links = [mu.body(1.0, [0.01] * 3)]
    dofs, acts = [], []
    for i in range(1, 5):
        links.append(mu.body(1.0 + 0.3 * i, [0.01, 0.004, 0.01], parent=i - 1, joint_type=0, q_index=i - 1,
                             v_index=i - 1, axis=(1, 0, 0) if i % 2 else (0, 1, 0),
                             parent_anchor=(0, 0, -0.12) if i > 1 else (0, 0, 1.0), child_anchor=(0, 0, 0.12)))
        dofs.append(mu.dof(i, i - 1, flags=ACTUATED | EFFORT_LIMIT | DRIVE, limits=(0, 0, 0, 8.0 + i),
                           drive=(60.0, 2.0, 0.01, 0.6 if i == 4 else 0.0)))
        if i in (1, 3):
            acts.append(mu.actuator(i - 1, torque_constant=1.5, current_limit=4.0, no_load_speed=3.0 + i,
                                    efficiency=0.9))
        else:
            acts.append(mu.actuator(i - 1, active=False))
    return mu.CustomWorld(links, dofs, [], link_count=5, root_type=0, nq=4, nv=4, envs=envs, actuators=acts)


@pytest.mark.gpu
@pytest.mark.parametrize("implicit", [False, True])
def test_gpu_actuator_envelope_matches_oracle(gpu, implicit):
    w = motor_chain()
    cfg = w.config(actuation_mode=1 if implicit else 0)
    world = mr.World(w.model, cfg)
    st = ol.OracleState(w.model, cfg)
    rng = np.random.default_rng(2 if implicit else 1)
    q0 = rng.uniform(-0.5, 0.5, (w.envs, 4)).astype(np.float32)
    v0 = rng.uniform(-4.0, 4.0, (w.envs, 4)).astype(np.float32)
    st.q[:] = q0
    st.v[:] = v0
    world.set_state(q0, v0, st.scene)
    steps = 6
    scale = 0.8 if implicit else 30.0  # targets in rad, or efforts well past the envelope
    cmds = (scale * rng.uniform(-1, 1, (steps, w.envs, 4))).astype(np.float32)
    for k in range(steps):
        print(f"drives: implicit={implicit} control step {k}", flush=True)
        world.advance(cmds[k:k + 1])
        status, *_ = st.step(cmds[k], control_step=k)
        assert np.all(status["code"] == 0)
        q, v, _ = world.state()
        assert np.array_equal(q.view(np.uint32), st.q.view(np.uint32)), f"step {k}: q differs"
        assert np.array_equal(v.view(np.uint32), st.v.view(np.uint32)), f"step {k}: v differs"
    # the envelope was active: some joint ran past half its no-load speed
    assert np.max(np.abs(st.v[:, [0, 2]])) > 2.0
