"""The task-rollout stepping surface end to end (c_api.h:583-746 semantics):
mr_create_task_rollout from a manifest, mr_task_rollout_advance with
normalized actions and reset masks, the advance record, the published
status spans, the layout counters and the final-state readback, checked
against the oracle stepping the same commands."""

import numpy as np
import pytest

import metalrobo as mr
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def oracle_commands(kind, model, actions):
    """The action mapping of the synthetic sources: PD targets of 0.3 rad per
    unit around the standing pose (humanoid) or efforts of 5 N m per unit
    (chain), actions clamped to [-1, 1], in float32 as the library computes."""
    E, nv = model.environment_count, model.max_nv
    a = np.clip(actions, np.float32(-1.0), np.float32(1.0)).astype(np.float32)
    cmd = np.zeros((actions.shape[0], E, nv), np.float32)
    if kind == mr.MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN:
        standing = model.initial_state()[0][0, 7:]  # joint coordinates of the standing pose
        cmd[:, :, 6:] = standing[None, None, :] + np.float32(0.3) * a
    else:
        cmd[:, :, :] = np.float32(5.0) * a
    return cmd


@pytest.mark.parametrize("source,kind,envs", [
    (mr.MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN, mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 6),
    (mr.MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES, mr.MR_SYNTHETIC_CHAIN_BOXES, 4),
])
def test_gpu_task_rollout_advance_matches_oracle(gpu, source, kind, envs):
    seed, steps = 11, 4
    print(f"task rollout: source {source}, {envs} environments", flush=True)
    ro = mr.TaskRollout(source, envs, seed)
    ro.set_state_readback(True)
    lay = ro.layout()
    model = mr.WorldModel.synthetic(kind, envs, seed)
    cfg = model.default_config()
    assert (lay.environment_count, lay.nq, lay.nv) == (envs, model.max_nq, model.max_nv)
    assert lay.action_count == (model.max_nv - 6 if kind == mr.MR_SYNTHETIC_HUMANOID_TERRAIN else model.max_nv)
    assert lay.world_fingerprint != 0 and lay.submission_count == 0
    rng = np.random.default_rng(seed)
    actions = rng.uniform(-1.3, 1.3, (steps, envs, lay.action_count)).astype(np.float32)  # some clamp
    masks = np.zeros((steps, envs), np.uint32)
    masks[2, 1] = 1
    out = ro.advance(actions, masks)

    st = ol.OracleState(model, cfg)
    cmds = oracle_commands(source, model, actions)
    o_cs = []
    for k in range(steps):
        status, cs, *_ = st.step(cmds[k], reset_mask=masks[k], control_step=k)
        assert np.all(status["code"] == 0)
        o_cs.append(cs)
    o_cs = np.stack(o_cs)
    # final state, bit for bit
    assert np.array_equal(ro.final_q().view(np.uint32), st.q.view(np.uint32))
    scene = ro.final_scene_states()
    o_scene = np.concatenate([st.scene["position"][..., :3], st.scene["orientation"],
                              st.scene["linear_velocity_inv_mass"][..., :3], st.scene["angular_velocity"][..., :3]],
                             axis=-1)
    assert np.array_equal(scene.view(np.uint32), o_scene.astype(np.float32).view(np.uint32))
    # published spans and the advance record
    np.testing.assert_array_equal(ro.status_codes(), np.zeros((steps, envs), np.uint32))
    np.testing.assert_array_equal(ro.active_contacts(), o_cs["constraints"])
    assert out.control_step_count == steps
    assert out.successful_environment_steps == steps * envs and out.failed_environment_steps == 0
    assert out.first_failing_environment == 0xFFFFFFFF and out.first_failing_control_step == 0xFFFFFFFF
    assert out.host_requested_resets == 1
    assert out.maximum_active_contacts == int(o_cs["constraints"].max())
    assert out.maximum_manifolds == int(o_cs["required_manifolds"].max())
    hw = out.high_water
    assert hw.candidate_pairs == int(o_cs["max_active_pairs"].max())
    assert hw.constraint_blocks == int(o_cs["max_constraints"].max())
    assert hw.constraint_rows == 3 * hw.constraint_blocks
    assert hw.raw_contacts == int(o_cs["raw_contacts"].max()) and hw.islands == int(o_cs["islands"].max())
    assert hw.operator_velocity_elements == model.max_nv and hw.ccd_events == 0
    assert out.gpu_milliseconds > 0.0 and out.submission_milliseconds >= out.gpu_milliseconds
    lay = ro.layout()
    assert lay.submitted_control_steps == steps and lay.completed_environment_steps == steps * envs
    assert lay.submission_count == 1
    assert lay.retained_buffer_bytes == (lay.immutable_private_bytes + lay.persistent_state_private_bytes +
                                         lay.transient_private_bytes)
    assert lay.immutable_private_bytes > 0 and lay.persistent_state_private_bytes > 0
    assert lay.total_gpu_milliseconds == pytest.approx(out.gpu_milliseconds)


def test_gpu_task_rollout_rejects_bad_spans_and_reset_restores(gpu):
    ro = mr.TaskRollout(mr.MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES, 3, 5)
    lay = ro.layout()
    good = np.zeros((2, 3, lay.action_count), np.float32)
    with pytest.raises(mr.MetalRoboError, match="action count"):
        ro.advance(good[:, :, :-1])
    with pytest.raises(mr.MetalRoboError, match="reset mask count"):
        ro.advance(good, np.zeros((2, 2), np.uint32))
    bad = good.copy()
    bad[1, 2, 0] = np.nan
    with pytest.raises(mr.MetalRoboError, match="nonfinite"):
        ro.advance(bad)
    assert ro.layout().submission_count == 0
    ro.set_state_readback(True)
    ro.advance(good)
    moved = ro.final_q()
    ro.reset(5)
    ro.advance(good)
    assert np.array_equal(ro.final_q(), moved)  # same seed: the same initial state and trajectory
