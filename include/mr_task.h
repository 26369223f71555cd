/* Locomotion task layer and policy inference around the batched stepper (C
 * ABI, device pointers, no torch types).
 *
 * The reference runs its task program as Metal kernels either side of the
 * physics step (src/metal/LocomotionTask.metal): mr_locomotion_task_apply_actions
 * (:3442) maps policy actions to implicit-drive targets, mr_locomotion_task_complete
 * (:3904) scores the accepted step (rewards, terminations, episode and command
 * bookkeeping) and mr_locomotion_task_observe (:2184) writes the next actor
 * observation; mr_policy_dense_layer and mr_policy_sample_and_score
 * (src/metal/PolicyInference.metal:78,172) evaluate the policy network. This
 * header carries the records those kernels read, laid out byte for byte as
 * include/metalrobo/task_program_types.h:445-607, and the entry points of
 * this library's HIP kernels for the locomotion subset of the task program:
 *   observations   root angular velocity (local), projected gravity, command,
 *                  joint position error, joint velocity, previous action,
 *                  root linear velocity (local), root height
 *                  (MRTaskObservationOpcode 0-7), optional vector3 normalization;
 *   actions        joint-position bindings with action delay slots and a
 *                  first-order response filter;
 *   rewards        linear/yaw velocity tracking, constant, root vertical and
 *                  roll-pitch velocity, tilt, root height error, joint
 *                  velocity, joint acceleration, action rate, joint limit
 *                  violation, projected gravity horizontal (MRTaskRewardOpcode
 *                  0-10 and 18), integrated over the control step;
 *   terminations   minimum root height, maximum tilt (with penalties and
 *                  priorities), timeout, physics error;
 *   episodes       step and episode counters, command resampling over its
 *                  duration, gait phase, reset of finished episodes.
 * Other opcodes (interaction packs, projectiles, terrain samples, contact
 * groups) are refused when the task is created. */
#ifndef MR_TASK_H
#define MR_TASK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* task_program_types.h:6-12 */
#define MR_TASK_ACTUATOR_JOINT_POSITION 0u
#define MR_TASK_ACTUATOR_GRIPPER_POSITION 4u

/* MRTaskObservationOpcode (task_program_types.h:68-76) */
#define MR_TASK_OBSERVE_ROOT_ANGULAR_VELOCITY_LOCAL 0u
#define MR_TASK_OBSERVE_PROJECTED_GRAVITY 1u
#define MR_TASK_OBSERVE_COMMAND 2u
#define MR_TASK_OBSERVE_JOINT_POSITION_ERROR 3u
#define MR_TASK_OBSERVE_JOINT_VELOCITY 4u
#define MR_TASK_OBSERVE_PREVIOUS_ACTION 5u
#define MR_TASK_OBSERVE_ROOT_LINEAR_VELOCITY_LOCAL 6u
#define MR_TASK_OBSERVE_ROOT_HEIGHT 7u
#define MR_TASK_OBSERVATION_NORMALIZE_VECTOR3 1u

/* MRTaskRewardOpcode (task_program_types.h:164-182) */
#define MR_TASK_REWARD_LINEAR_VELOCITY_TRACKING 0u
#define MR_TASK_REWARD_YAW_VELOCITY_TRACKING 1u
#define MR_TASK_REWARD_CONSTANT 2u
#define MR_TASK_REWARD_ROOT_VERTICAL_VELOCITY_SQUARED 3u
#define MR_TASK_REWARD_ROOT_ROLL_PITCH_VELOCITY_SQUARED 4u
#define MR_TASK_REWARD_TILT_SQUARED 5u
#define MR_TASK_REWARD_ROOT_HEIGHT_ERROR_SQUARED 6u
#define MR_TASK_REWARD_JOINT_VELOCITY_SQUARED 7u
#define MR_TASK_REWARD_JOINT_ACCELERATION_SQUARED 8u
#define MR_TASK_REWARD_ACTION_RATE_SQUARED 9u
#define MR_TASK_REWARD_JOINT_LIMIT_VIOLATION_SQUARED 10u
#define MR_TASK_REWARD_PROJECTED_GRAVITY_HORIZONTAL_SQUARED 18u

/* MRTaskTerminationOpcode / Reason (task_program_types.h:247-263) */
#define MR_TASK_TERMINATE_MINIMUM_ROOT_HEIGHT 0u
#define MR_TASK_TERMINATE_MAXIMUM_TILT 1u
#define MR_TASK_TERMINATION_CONTINUING 0u
#define MR_TASK_TERMINATION_HEIGHT 1u
#define MR_TASK_TERMINATION_TILT 2u
#define MR_TASK_TERMINATION_TIMEOUT 4u
#define MR_TASK_TERMINATION_PHYSICS_ERROR 5u

/* MRTaskProgramFlags (task_program_types.h:14-44) */
#define MR_TASK_PROGRAM_FIXED_ROOT (1u << 8u)

/* MRPolicyActivationOpcode, MRPolicyDenseFlags, MRPolicyProgramFlags (policy_program_types.h:7-26) */
#define MR_POLICY_ACTIVATION_IDENTITY 0u
#define MR_POLICY_ACTIVATION_RELU 1u
#define MR_POLICY_ACTIVATION_TANH 2u
#define MR_POLICY_ACTIVATION_ELU 3u
#define MR_POLICY_ACTIVATION_SILU 4u

/* MRTaskActionBindingGPU (task_program_types.h:445-458), 64 bytes:
 * indices = action index, global DoF index, q index, v index;
 * parameters = normalized scale, lower target, upper target, response time (s);
 * drive = authored stiffness, damping (unused here); actuator.x = kind. */
typedef struct MRTaskActionBindingC {
    uint32_t indices[4];
    float parameters[4];
    float drive[4];
    uint32_t actuator[4];
} MRTaskActionBindingC;

/* MRTaskObservationOperatorGPU (:467-474), 48 bytes: source = opcode, source
 * index, component, flags; transform = scale, offset, uniform-noise amplitude;
 * auxiliary.y = noise channel (auxiliary.x must be MR_INVALID_INDEX: no bias slots). */
typedef struct MRTaskObservationOperatorC {
    uint32_t source[4];
    float transform[4];
    uint32_t auxiliary[4];
} MRTaskObservationOperatorC;

/* MRTaskRewardOperatorGPU (:498-506), 48 bytes: source.x = opcode;
 * parameters.x = weight (a rate, integrated over the control step), .y = width. */
typedef struct MRTaskRewardOperatorC {
    uint32_t source[4];
    float parameters[4];
    float auxiliary[4];
} MRTaskRewardOperatorC;

/* MRTaskTerminationOperatorGPU (:508-515), 48 bytes: source = opcode, index,
 * reason, priority; parameters = threshold, one-shot failure penalty. */
typedef struct MRTaskTerminationOperatorC {
    uint32_t source[4];
    float parameters[4];
    uint32_t schedule[4];
} MRTaskTerminationOperatorC;

/* MRTaskStateGPU (:553-581), 176 bytes. Used here: episode = episode step,
 * episode index; schedule = command steps left, -, actuator delay; status =
 * initialized, pending reset, last termination reason; command_and_phase =
 * commanded x/y/yaw velocity and gait phase. */
typedef struct MRTaskStateC {
    uint32_t episode[4];
    uint32_t schedule[4];
    uint32_t status[4];
    float command_and_phase[4];
    float command_extension[4];
    float air_return_tracking[4];
    float recovery[4];
    uint32_t recovery_stats[4];
    float threat_geometry[4];
    float threat_teacher[4];
    uint32_t threat_metadata[4];
} MRTaskStateC;

/* MRTaskTransitionGPU (:603-623), 128 bytes: reward_and_state = reward,
 * tracking score, root height, tilt; termination = done, timeout, physics
 * error, reason; reward_breakdown0 = task, base, joint velocity, joint
 * acceleration; reward_breakdown1 = action rate, posture/limits, energy, contact. */
typedef struct MRTaskTransitionC {
    float reward_and_state[4];
    uint32_t termination[4];
    float reward_breakdown0[4];
    float reward_breakdown1[4];
    float outcome_channels0[4];
    float outcome_channels1[4];
    uint64_t policy_revision;
    float timeout_bootstrap_value;
    float episode_tracking_score;
    uint32_t task_progress[4];
} MRTaskTransitionC;

/* The locomotion task program (the fields of MRTaskProgramHeaderGPU,
 * task_program_types.h:325-443, that the subset reads; tables by host pointer,
 * copied at creation). */
typedef struct MRLocomotionTaskProgramC {
    const MRTaskActionBindingC* actions;             /* counts0.x */
    const MRTaskObservationOperatorC* observations;  /* one actor frame, counts0.y */
    const MRTaskRewardOperatorC* rewards;            /* counts1.w */
    const MRTaskTerminationOperatorC* terminations;  /* counts2.x */
    const float* default_q;                          /* [nq] nominal pose */
    uint32_t action_count;
    uint32_t observation_count;
    uint32_t reward_count;
    uint32_t termination_count;
    uint32_t history_length;   /* actor frames kept, newest last (layout.y) */
    uint32_t delay_slots;      /* action history slots including the filter slot, >= 2 (layout.w) */
    uint32_t max_episode_steps; /* schedule.x */
    uint32_t flags;            /* MR_TASK_PROGRAM_FIXED_ROOT or 0 */
    uint32_t root_q_offset;    /* root.z */
    uint32_t root_v_offset;    /* root.w */
    uint32_t actuator_delay;   /* MRTaskStateGPU.schedule.z of every episode */
    uint32_t reserved;
    float root_reference[4];   /* link-frame origin relative to the root COM, root axes */
    float base_height_target;  /* locomotion.x */
    float gait_period;         /* locomotion.y, seconds */
    float command_lower[4];    /* xyz lower bounds, w standing-command probability */
    float command_upper[4];    /* xyz upper bounds */
    float command_seconds[2];  /* scheduleSeconds.xy: command duration bounds */
    uint64_t seed;             /* MRTaskDispatchGPU.seed */
} MRLocomotionTaskProgramC;

typedef struct MRLocomotionTaskHandle MRLocomotionTaskHandle;
struct MRWorldHandle;

/* Binds a task program to a world (one task state per environment); the
 * world's reset state is the state a finished episode restarts from. The
 * world must step one workgroup per environment (no work-queue schedule)
 * while the task drives it. Returns NULL with mr_last_error set on failure,
 * including unsupported opcodes and out-of-range indices. */
MRLocomotionTaskHandle* mr_locomotion_task_create(const MRLocomotionTaskProgramC* program,
                                                  struct MRWorldHandle* world);
void mr_locomotion_task_destroy(MRLocomotionTaskHandle* task);
/* Starts episode 0 in every environment from the world's reset state: task
 * states, action and observation histories, commands, and the first actor
 * observation. */
int mr_locomotion_task_reset(MRLocomotionTaskHandle* task);
/* One control step on the world's stream: apply_actions (device actions
 * [E][action_count]) -> physics control step (finished episodes of the step
 * before restart through the reset mask) -> complete -> observe. Enqueues
 * without waiting. */
int mr_locomotion_task_step(MRLocomotionTaskHandle* task, const float* device_actions);
/* Waits, then copies the actor observations [E][history][observation_count],
 * the last step's transitions [E] and task states [E] (any may be null). */
int mr_locomotion_task_read(MRLocomotionTaskHandle* task, float* observations, size_t observation_count,
                            MRTaskTransitionC* transitions, size_t transition_count, MRTaskStateC* states,
                            size_t state_count);
/* Device pointers, valid until destroy: actor observations [E][history][O],
 * transitions [E], the implicit-drive targets of the last step [E][max_nv]. */
const float* mr_locomotion_task_device_observations(const MRLocomotionTaskHandle* task);
const MRTaskTransitionC* mr_locomotion_task_device_transitions(const MRLocomotionTaskHandle* task);
const float* mr_locomotion_task_device_targets(const MRLocomotionTaskHandle* task);

/* mr_policy_dense_layer (PolicyInference.metal:78-170): for every environment
 * e and output neuron n, output[e * output_stride + n] =
 * act(bias[n] + sum_f weights[n * input_width + f] * x[f]) with the sum a chain
 * of fused multiply-adds in feature order, x[f] = input[e * input_stride + f],
 * or with mean/inverse_std non-null, clamp((x - mean) * inverse_std, -clip, clip).
 * All pointers are device pointers; stream is a hipStream_t or null. */
int mr_policy_dense_layer(const float* weights, const float* bias, const float* mean, const float* inverse_std,
                          float observation_clip, const float* input, uint32_t input_stride, float* output,
                          uint32_t output_stride, uint32_t environments, uint32_t input_width, uint32_t output_width,
                          uint32_t activation, void* stream);
/* mr_policy_sample_and_score (PolicyInference.metal:172-268): per environment,
 * latent = mean (+ exp(log_std) * N(0,1) when log_std is non-null, the normal
 * drawn from the SplitMix64 key of (seed, environment, episode index and step
 * from task_states, control_step, revision)), action = clamp(action_scale *
 * latent + action_bias, -action_clip, action_clip), and the diagonal-Gaussian
 * log probability of the latent. Outputs [E][actions], [E][actions], [E]. */
int mr_policy_sample_and_score(const float* actor_mean, uint32_t mean_stride, const float* action_bias,
                               const float* action_scale, const float* log_std, float action_clip, uint64_t seed,
                               uint64_t revision, const MRTaskStateC* task_states, uint32_t environments,
                               uint32_t actions, uint32_t control_step, float* actions_out, float* latents_out,
                               float* log_probabilities_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MR_TASK_H */
