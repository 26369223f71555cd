"""Pins the oracle's kinematics and articulated-body dynamics (CPU only).

Checks:
  * deterministic sine/cosine against float64 libm;
  * the free-body known answer of the reference ABA probe
    (apps/articulated_aba_gpu_probe.mm:927-981): m = 1, I = 0.1, gravity
    (0, -9.81, 0), root effort + body wrench -> a = (1.5, -10.31, 0.3, 0.2, 0.2, 0.7);
  * FP32 ABA against an independent FP64 Jacobian/Newton-Euler formulation on
    the chain and humanoid models at random states, within the reference's
    ABA tolerance of 5e-5 scaled acceleration (articulated_aba_gpu_probe.mm:811-813);
  * inverse-ABA point responses against FP64 M^{-1} J^T f;
  * forward kinematics against the FP64 kinematics.
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol
from fp64_dynamics import TreeModel


def test_det_sincos_accuracy():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-4, 4, 20000), rng.uniform(-200, 200, 20000),
                        np.array([0.0, 1e-8, -1e-8, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 4])]).astype(np.float32)
    s, c = ol.sincos(x)
    xs = x.astype(np.float64)
    err_s = np.abs(s.astype(np.float64) - np.sin(xs))
    err_c = np.abs(c.astype(np.float64) - np.cos(xs))
    assert err_s.max() < 2.5e-7 and err_c.max() < 2.5e-7, (err_s.max(), err_c.max())
    assert np.all(np.sign(s[np.abs(np.sin(xs)) > 1e-6]) == np.sign(np.sin(xs)[np.abs(np.sin(xs)) > 1e-6]))


def free_body_world():
    b = mu.body(1.0, [0.1, 0.1, 0.1])
    dofs = [mu.dof(0, i if i < 3 else mu.INVALID, flags=1) for i in range(6)]
    return mu.CustomWorld([b], dofs, [], link_count=1, root_type=1, nq=7, nv=6, gravity=(0.0, -9.81, 0.0),
                          microstep=1.0e-3)


def test_free_body_known_answer():
    w = free_body_world()
    cfg = w.config(control_timestep=1.0e-3, physics_substeps=1, apply_body_damping=0)
    q = np.array([0.2, -0.1, 0.8, 0, 0, 0, 1], np.float32)
    v = np.array([0.1, -0.2, 0.3, 0, 0, 0], np.float32)
    # root effort plus the probe's body wrench: force (0.3, 0.2, -0.1), torque (-0.01, 0.04, 0.02)
    tau = np.array([1.2 + 0.3, -0.7 + 0.2, 0.4 - 0.1, 0.03 - 0.01, -0.02 + 0.04, 0.05 + 0.02], np.float32)
    code, acc = ol.forward_dynamics(w.model, cfg, 0, q, v, tau)
    assert code == 0
    expect = np.array([1.5, -10.31, 0.3, 0.2, 0.2, 0.7])
    np.testing.assert_allclose(acc, expect, atol=2.0e-5)


def random_state(tm, rng, model_q0=None, scale=1.0):
    q = np.zeros(tm.nq, np.float64) if model_q0 is None else np.array(model_q0, np.float64)
    v = rng.normal(0, scale, tm.nv)
    if tm.floating:
        quat = rng.normal(size=4)
        q[3:7] = quat / np.linalg.norm(quat)
        q[7:] += rng.uniform(-0.8, 0.8, tm.nq - 7)
    else:
        q += rng.uniform(-2.0, 2.0, tm.nq)
    return q.astype(np.float32), v.astype(np.float32)


@pytest.mark.parametrize("kind,envs", [(mr.MR_SYNTHETIC_CHAIN_BOXES, 6), (mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 2)])
def test_aba_matches_fp64_dynamics(kind, envs):
    model = mr.WorldModel.synthetic(kind, envs, 17)
    cfg = model.default_config()
    cfg.actuation_mode = mr.native.MR_ACTUATION_EFFORT
    headers = model.models()
    env_model = model.environment_model()
    bodies, dofs = model.bodies(), model.dofs()
    q0, _, _ = model.initial_state()
    rng = np.random.default_rng(1)
    worst = 0.0
    for env in range(envs):
        tm = TreeModel(headers[env_model[env]], bodies, dofs, model.desc.gravity.x and [0, 0, -9.81] or [0, 0, -9.81])
        arm = np.array([d["drive"][2] for d in tm.dofs], np.float64)
        for trial in range(4):
            q, v = random_state(tm, rng, q0[env], scale=1.0)
            tau = rng.normal(0, 5.0, tm.nv).astype(np.float32)
            code, acc = ol.forward_dynamics(model, cfg, env, q, v, tau)
            assert code == 0
            ref = tm.forward_dynamics(q, v, tau, arm, damping=True)
            scaled = np.max(np.abs(acc - ref)) / max(1.0, np.max(np.abs(ref)))
            worst = max(worst, scaled)
            assert scaled < 5.0e-5, (env, trial, scaled)
    print(f"worst scaled acceleration error {worst:.2e}")


def test_aba_reduces_to_inverse_mass_for_point_forces():
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 1, 3)
    cfg = model.default_config()
    cfg.actuation_mode = mr.native.MR_ACTUATION_EFFORT
    tm = TreeModel(model.models()[0], model.bodies(), model.dofs(), [0, 0, -9.81])
    arm = np.array([d["drive"][2] for d in tm.dofs], np.float64)
    q0, v0, _ = model.initial_state()
    rng = np.random.default_rng(5)
    for link in [6, 12, 15, 22, 29, 0]:
        q, v = random_state(tm, rng, q0[0])
        pos, rot, joint, axis = tm.kinematics(q, v)
        x = (pos[link] + rng.normal(0, 0.05, 3)).astype(np.float32)
        f = rng.normal(0, 1.0, 3).astype(np.float32)
        code, out = ol.point_response(model, cfg, 0, q, v, link, x, f)
        assert code == 0
        ref = tm.point_response(q, arm, link, x, f)
        rel = np.max(np.abs(out - ref)) / max(1e-6, np.max(np.abs(ref)))
        assert rel < 1.0e-4, (link, rel)


@pytest.mark.parametrize("kind", [mr.MR_SYNTHETIC_CHAIN_BOXES, mr.MR_SYNTHETIC_HUMANOID_TERRAIN])
def test_kinematics_matches_fp64(kind):
    model = mr.WorldModel.synthetic(kind, 2, 9)
    cfg = model.default_config()
    tm = TreeModel(model.models()[model.environment_model()[1]], model.bodies(), model.dofs(), [0, 0, -9.81])
    q0, v0, _ = model.initial_state()
    rng = np.random.default_rng(2)
    for _ in range(5):
        q, v = random_state(tm, rng, q0[1])
        code, poses = ol.link_poses(model, cfg, 1, q, v)
        assert code == 0
        pos, rot, _, _ = tm.kinematics(q, v)
        np.testing.assert_allclose(poses[:, :3], pos, atol=2e-6)
        same = np.sign(np.sum(poses[:, 3:] * rot, axis=1))[:, None]
        np.testing.assert_allclose(poses[:, 3:] * same, rot, atol=2e-6)


def test_pendulum_energy_is_conserved_without_damping():
    """A frictionless chain in gravity conserves kinetic + potential energy
    (ArticulatedDynamics.hpp:213-221 invariants); with symplectic microsteps
    of 0.5 ms the drift over 0.5 s stays small."""
    links = [mu.body(1.0, [0.01, 0.01, 0.01])]
    dofs = []
    for i in range(1, 4):
        links.append(mu.body(1.0, [0.02, 0.002, 0.02], parent=i - 1, joint_type=0, q_index=i - 1, v_index=i - 1,
                             axis=(1, 0, 0), parent_anchor=(0, 0.15, 0) if i > 1 else (0, 0, 0),
                             child_anchor=(0, -0.15, 0)))
        dofs.append(mu.dof(i, i - 1))
    w = mu.CustomWorld(links, dofs, [], link_count=4, root_type=0, nq=3, nv=3, gravity=(0, 0, -9.81),
                       microstep=0.0005)
    cfg = w.config(control_timestep=0.0005, physics_substeps=1, apply_body_damping=0)
    st = ol.OracleState(w.model, cfg)
    st.q[0] = [0.6, -0.3, 0.2]
    tm = TreeModel(w.model.models()[0], w.model.bodies(), w.model.dofs(), [0, 0, -9.81])
    arm = np.zeros(3)
    e0 = tm.kinetic_energy(st.q[0], st.v[0], arm) + tm.potential_energy(st.q[0])
    cmd = np.zeros((1, 3), np.float32)
    for k in range(1000):
        status, *_ = st.step(cmd, control_step=k)
        assert status["code"][0] == 0
    e1 = tm.kinetic_energy(st.q[0], st.v[0], arm) + tm.potential_energy(st.q[0])
    assert abs(e1 - e0) < 2e-2 * max(1.0, abs(e0)), (e0, e1)


def test_chain_blow_up_is_the_step_size_limit_of_explicit_velocity_terms():
    """The chain + boxes workload's rare non-finite worlds come from the
    microstep, not from the contacts or the kernel: world 158 (seed 7) spins
    its effort-driven links until h |w| per microstep passes ~2 rad and
    blows up at 4 substeps (h = 5 ms); at 8 substeps (h = 2.5 ms) the same
    world, the same torques, stays finite. Semi-implicit Euler integrates the
    velocity-product (Coriolis, centrifugal) terms explicitly, which is
    stable only while h |w| stays of order 1."""
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 160, 7)
    cmds = model.commands(0, 60)
    outcome = {}
    for subs in (4, 8):
        cfg = model.default_config()
        cfg.physics_substeps = subs
        st = ol.OracleState(model, cfg)
        h = cfg.control_timestep / subs
        peak_hw, code = 0.0, 0
        for k in range(60):
            status, *_ = st.step(cmds[k], control_step=k, env_begin=158, env_end=159)
            code = int(status["code"][158])
            if code != 0:
                break
            peak_hw = max(peak_hw, h * float(np.abs(st.v[158]).max()))
        outcome[subs] = (code, k, peak_hw)
    code4, step4, hw4 = outcome[4]
    assert code4 == 3 and step4 == 37, outcome  # MR_STEP_NONFINITE_RESULT at control step 37
    assert hw4 > 2.0, outcome                    # the last accepted step was already past the limit
    code8, _, hw8 = outcome[8]
    assert code8 == 0 and hw8 < 1.0, outcome
