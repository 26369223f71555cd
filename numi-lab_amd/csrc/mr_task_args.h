// Launch record of the task-layer kernels, shared by the host executor and
// the device code.
#pragma once

#include "../../include/mr_task.h"
#include "../../include/mr_world_types.h"

#include <stdint.h>

namespace mr {

struct TaskArgs {
    // program (device copies)
    const MRTaskActionBindingC* actions;             // [A]
    const MRTaskObservationOperatorC* observations;  // [O]
    const MRTaskRewardOperatorC* rewards;            // [R]
    const MRTaskTerminationOperatorC* terminations;  // [T]
    const float* default_q;                          // [max_nq]
    const float* response_fraction;                  // [A]
    const MRDofC* dofs;                              // the world model's DoF table
    uint32_t action_count, observation_count, reward_count, termination_count;
    uint32_t history_length, delay_slots, max_episode_steps, flags;
    uint32_t root_q_offset, root_v_offset, actuator_delay, environments;
    uint32_t max_nq, max_nv;
    float dt;
    float base_height_target, gait_period;
    float root_reference[3];
    float command_lower[4], command_upper[4], command_seconds[2];
    unsigned long long seed;
    unsigned long long policy_revision;  // stamped on every transition
    // world state (device)
    const float* q;        // [E][max_nq]
    const float* v;        // [E][max_nv]
    const float* reset_q;  // [E][max_nq]
    const float* reset_v;  // [E][max_nv]
    // task state (device)
    MRTaskStateC* states;             // [E]
    float* action_history;            // [E][delay_slots][A]
    float* previous_joint;            // [E][2][A]: joint velocity, joint position after the last step
    float* observations_out;          // [E][history][O]
    MRTaskTransitionC* transitions;   // [E]
    float* targets;                   // [E][max_nv] implicit-drive command of the next physics step
    uint32_t* reset_mask;             // [E] reset mask of the next physics step
};

} // namespace mr
