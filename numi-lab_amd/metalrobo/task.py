"""Locomotion task programs for the task layer (include/mr_task.h).

humanoid_locomotion_program() builds the velocity-tracking program of the
29-DoF humanoid in the form the reference's G1 locomotion TaskPack takes
(actions = joint-position targets around the nominal pose; actor frame =
root angular velocity, projected gravity, command, joint position error,
joint velocity and previous action; tracking, regularization and
termination operators with rate weights), on this library's synthetic model.
"""

from __future__ import annotations

import numpy as np

from .native import (TASK_ACTION_DTYPE, TASK_OBSERVATION_DTYPE, TASK_REWARD_DTYPE, TASK_TERMINATION_DTYPE)

INVALID = 0xFFFFFFFF

# MRTaskObservationOpcode, MRTaskRewardOpcode, MRTaskTerminationOpcode (include/mr_task.h)
OBS_ANG_VEL, OBS_GRAVITY, OBS_COMMAND, OBS_JOINT_POS, OBS_JOINT_VEL, OBS_PREV_ACTION, OBS_LIN_VEL, OBS_HEIGHT = range(8)
REWARD = {"lin_vel_tracking": 0, "yaw_vel_tracking": 1, "constant": 2, "vertical_velocity": 3,
          "roll_pitch_velocity": 4, "tilt": 5, "height_error": 6, "joint_velocity": 7, "joint_acceleration": 8,
          "action_rate": 9, "joint_limits": 10, "gravity_horizontal": 18}


def observation(op, index=0, component=0, scale=1.0, offset=0.0, noise=0.0, channel=0, flags=0):
    r = np.zeros(1, TASK_OBSERVATION_DTYPE)[0]
    r["source"] = (op, index, component, flags)
    r["transform"] = (scale, offset, noise, 0.0)
    r["auxiliary"] = (INVALID, channel, 0, 0)
    return r


def humanoid_locomotion_program(model, seed: int = 7, noise: bool = True, history: int = 3, delay_slots: int = 4,
                                actuator_delay: int = 1, response_time: float = 0.02) -> dict:
    """The velocity-tracking program for a floating-root model whose joint DoFs follow the 6 root
    DoFs (the synthetic humanoid): one joint-position action per joint DoF, scale 0.25 rad around
    the model's standing pose, targets clamped to the joint limits."""
    dofs = model.dofs()
    h = model.models()[0]
    nv = int(h["nv"])
    q0 = model.initial_state()[0][0].astype(np.float32)
    joints = [j for j in range(nv) if dofs[j]["q_index"] != INVALID and j >= 6]
    acts = np.zeros(len(joints), TASK_ACTION_DTYPE)
    for k, j in enumerate(joints):
        qi = int(dofs[j]["q_index"])
        lo, hi = float(dofs[j]["limits"][0]), float(dofs[j]["limits"][1])
        acts[k]["indices"] = (k, j, qi, j)
        acts[k]["parameters"] = (0.25, lo, hi, response_time)
        acts[k]["actuator"] = (0, 0, 0, 0)
    n = 0.05 if noise else 0.0
    ops = []
    for c in range(3):
        ops.append(observation(OBS_ANG_VEL, 0, c, 0.25, noise=0.2 * n, channel=100 + c))
    for c in range(3):
        ops.append(observation(OBS_GRAVITY, 0, c, 1.0, noise=n, channel=110 + c, flags=1))
    for c, s in enumerate((2.0, 2.0, 0.25)):
        ops.append(observation(OBS_COMMAND, 0, c, s))
    for k in range(len(joints)):
        ops.append(observation(OBS_JOINT_POS, k, 0, 1.0, noise=0.2 * n, channel=200 + k))
    for k in range(len(joints)):
        ops.append(observation(OBS_JOINT_VEL, k, 0, 0.05, noise=n, channel=300 + k))
    for k in range(len(joints)):
        ops.append(observation(OBS_PREV_ACTION, k, 0, 1.0))
    ops.append(observation(OBS_LIN_VEL, 0, 0, 2.0, noise=0.2 * n, channel=120))
    ops.append(observation(OBS_HEIGHT, 0, 0, 1.0, offset=-0.75))
    obs = np.array(ops, TASK_OBSERVATION_DTYPE)
    rewards = []
    for name, w, width in (("lin_vel_tracking", 1.0, 0.25), ("yaw_vel_tracking", 0.5, 0.25),
                           ("vertical_velocity", -2.0, 0.0), ("roll_pitch_velocity", -0.05, 0.0),
                           ("tilt", -1.0, 0.0), ("height_error", -10.0, 0.0), ("joint_velocity", -1.0e-3, 0.0),
                           ("joint_acceleration", -2.5e-7, 0.0), ("action_rate", -0.01, 0.0),
                           ("joint_limits", -5.0, 0.0), ("gravity_horizontal", -0.5, 0.0)):
        r = np.zeros(1, TASK_REWARD_DTYPE)[0]
        r["source"] = (REWARD[name], 0, 0, 0)
        r["parameters"] = (w, width, 0.0, 0.0)
        rewards.append(r)
    terms = np.zeros(2, TASK_TERMINATION_DTYPE)
    terms[0]["source"] = (0, 0, 1, 1)  # minimum root height -> reason height, priority 1
    terms[0]["parameters"] = (0.35, -1.0, 0.0, 0.0)
    terms[1]["source"] = (1, 0, 2, 2)  # maximum tilt -> reason tilt, priority 2
    terms[1]["parameters"] = (1.0, -1.0, 0.0, 0.0)
    return {
        "actions": acts, "observations": obs, "rewards": np.array(rewards, TASK_REWARD_DTYPE), "terminations": terms,
        "default_q": q0, "history_length": history, "delay_slots": delay_slots, "max_episode_steps": 1000,
        "flags": 0, "root_q_offset": 0, "root_v_offset": 0, "actuator_delay": actuator_delay,
        "root_reference": (0.0, 0.0, 0.0, 0.0), "base_height_target": float(q0[2]), "gait_period": 0.8,
        "command_lower": (-1.0, -0.5, -1.0, 0.1), "command_upper": (1.0, 0.5, 1.0, 0.0),
        "command_seconds": (2.0, 6.0), "seed": seed,
    }
