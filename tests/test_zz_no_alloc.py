"""mr_world_enqueue_device_static, the entry point that allocates no device
memory: its per-step scratch comes from the library's static device array and
its statuses go to the caller's buffers. Two calls in a row, the second on a
smaller world, each give what mr_world_enqueue_device gives on an identical
world: the same state, manifolds, statuses and advance summary, bit for bit.
A world whose scratch exceeds the array is refused with an error."""

import numpy as np
import pytest

import metalrobo as mr

pytestmark = pytest.mark.gpu


def _snapshot(world):
    q, v, s = world.state()
    h, p, c = world.manifolds()
    for e in range(len(c)):
        h[e, c[e]:] = np.zeros(1, h.dtype)
        p[e, c[e]:] = np.zeros(1, p.dtype)
    return [np.ascontiguousarray(x).view(np.uint8) for x in (q, v, s, h, p, c)]


def _summary(adv):
    return (adv.control_step_count, adv.successful_environment_steps, adv.failed_environment_steps,
            adv.first_failing_environment, adv.first_failing_control_step, adv.first_gpu_status_code,
            adv.maximum_active_pairs, adv.maximum_constraints)


def _compare(kind, envs, steps, seed):
    import torch

    print(f"static scratch: kind {kind}, {envs} environments, {steps} steps, seed {seed}", flush=True)
    model = mr.WorldModel.synthetic(kind, envs, seed)
    cfg = model.default_config()
    own = mr.World(model, cfg)
    free = mr.World(model, cfg)
    cmds = torch.from_numpy(model.commands(0, steps)).to("cuda")
    masks_h = np.zeros((steps, envs), np.int32)
    masks_h[steps - 1, envs - 1] = 1  # one reset on the last step
    masks = torch.from_numpy(masks_h).to("cuda")
    own.enqueue_device(cmds, masks)
    a = own.wait()
    status = torch.zeros(steps * envs * mr.WORLD_STATUS_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    cstatus = torch.zeros(steps * envs * mr.CONTACT_STATUS_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    free.enqueue_device_static(cmds, status, cstatus, masks)
    b = free.wait()
    assert _summary(a) == _summary(b)
    assert b.control_step_count == steps and b.failed_environment_steps == 0
    # the statuses land in the caller's buffers and read back through the C ABI
    own_st, own_cs = own.statuses(), own.contact_statuses()
    assert np.array_equal(status.cpu().numpy(), own_st.view(np.uint8).reshape(-1))
    assert np.array_equal(cstatus.cpu().numpy(), own_cs.view(np.uint8).reshape(-1))
    assert np.array_equal(free.statuses().view(np.uint8), own_st.view(np.uint8))
    assert np.array_equal(free.contact_statuses().view(np.uint8), own_cs.view(np.uint8))
    assert int(own_cs["constraints"].max()) > 0  # the step solved contacts
    for x, y in zip(_snapshot(own), _snapshot(free)):
        assert np.array_equal(x, y)
    own.close()
    free.close()


def test_static_scratch_entry_matches_enqueue_device_twice(gpu):
    _compare(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 8, 4, 3)
    # second call, smaller world: the array keeps the first call's contents
    _compare(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 3, 2, 9)


def test_static_scratch_entry_refuses_a_world_beyond_the_array(gpu):
    import torch

    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 1024, 1)
    world = mr.World(model, model.default_config())
    cmds = torch.from_numpy(model.commands(0, 1)).to("cuda")
    status = torch.zeros(1024 * mr.WORLD_STATUS_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    cstatus = torch.zeros(1024 * mr.CONTACT_STATUS_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    with pytest.raises(mr.MetalRoboError, match="bytes of scratch"):
        world.enqueue_device_static(cmds, status, cstatus)
    world.close()
