// Locomotion task layer and policy inference on gfx950: the kernels either
// side of the physics control step.
//
// Reference (semantics restated; the code is written independently):
//   mr_task_apply_actions_kernel    src/metal/LocomotionTask.metal:3442-3674
//   mr_task_complete_kernel         src/metal/LocomotionTask.metal:3904-6746 (locomotion subset)
//   mr_task_observe_kernel          src/metal/LocomotionTask.metal:2184-3440, 594-1330 (locomotion subset)
//   mr_policy_dense_tile_kernel     src/metal/PolicyInference.metal:78-170
//   mr_policy_sample_kernel         src/metal/PolicyInference.metal:172-268
//
// complete runs one thread per environment (a few hundred FLOPs per
// environment, beside the physics step's millions); apply_actions and observe
// run one wavefront per environment over its actions and operators; the dense layer
// runs register-tiled over environments x neurons with LDS-staged operands.

#include "mr_device_math.h"
#include "mr_task_args.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

using namespace mrd;

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;

// SplitMix64 finalizer of (value + golden gamma) (LocomotionTask.metal:121-130).
__device__ __forceinline__ unsigned long long mix64(unsigned long long value) {
    value += 0x9e3779b97f4a7c15ull;
    value = (value ^ (value >> 30u)) * 0xbf58476d1ce4e5b9ull;
    value = (value ^ (value >> 27u)) * 0x94d049bb133111ebull;
    return value ^ (value >> 31u);
}

// Counter-keyed uniform in [0, 1) with 24 bits (LocomotionTask.metal:132-151).
__device__ __forceinline__ float random_unit(unsigned long long seed, uint32_t env, uint32_t episode, uint32_t step,
                                             uint32_t channel) {
    unsigned long long key = seed;
    key ^= (unsigned long long)(env + 1u) * 0xd2b74407b1ce6e93ull;
    key ^= (unsigned long long)(episode + 1u) * 0xca5a826395121157ull;
    key ^= (unsigned long long)(step + 1u) * 0x9e3779b185ebca87ull;
    key ^= (unsigned long long)(channel + 1u) * 0x94d049bb133111ebull;
    return (float)(uint32_t)(mix64(key) >> 40u) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float random_range(unsigned long long seed, uint32_t env, uint32_t episode, uint32_t step,
                                              uint32_t channel, float lower, float upper) {
    return lower + (upper - lower) * random_unit(seed, env, episode, step, channel);
}

// Control steps of a duration drawn in [lower, upper] seconds (:1484-1508).
__device__ __forceinline__ uint32_t duration_steps(const mr::TaskArgs& a, uint32_t env, uint32_t episode, uint32_t step,
                                                   uint32_t channel, float lower, float upper) {
    const float s = random_range(a.seed, env, episode, step, channel, lower, upper) / a.dt;
    const uint32_t n = (uint32_t)floorf(s);
    return n > 1u ? n : 1u;
}

// Commanded x/y/yaw velocity of an episode segment (:1570-1615, one difficulty band).
__device__ __forceinline__ V3 sampled_command(const mr::TaskArgs& a, uint32_t env, uint32_t episode, uint32_t step) {
    float c[3];
    for (uint32_t k = 0; k < 3u; ++k)
        c[k] = random_range(a.seed, env, episode, step, 16u + k, a.command_lower[k], a.command_upper[k]);
    if (random_unit(a.seed, env, episode, step, 19u) < a.command_lower[3]) {
        c[0] = 0.0f;
        c[1] = 0.0f;
        c[2] = 0.0f;
    }
    return V3{c[0], c[1], c[2]};
}

__device__ __forceinline__ V3 rotate(Q4 q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = mul(2.0f, cross(u, v));
    return add(add(v, mul(q.w, t)), cross(u, t));
}

__device__ __forceinline__ V3 rotate_inverse(Q4 q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = mul(2.0f, cross(u, v));
    return add(sub(v, mul(q.w, t)), cross(u, t));
}

__device__ __forceinline__ float comp(V3 v, uint32_t i) { return i == 0u ? v.x : (i == 1u ? v.y : v.z); }

// v / |v|, or the fallback when |v|^2 is not a positive finite value (:259-268).
__device__ __forceinline__ V3 normalized_or(V3 v, V3 fallback) {
    const float l2 = dot(v, v);
    if (!(l2 > 1.0e-12f) || !__builtin_isfinite(l2)) return fallback;
    return mul(1.0f / sqrtf(l2), v);
}

// Root kinematics of the task (:396-478).
struct Root {
    Q4 orientation;
    V3 position, linear, angular;
};

__device__ __forceinline__ Root root_of(const mr::TaskArgs& a, const float* q, const float* v) {
    Root r;
    if ((a.flags & MR_TASK_PROGRAM_FIXED_ROOT) != 0u) {
        r.orientation = Q4{0.0f, 0.0f, 0.0f, 1.0f};
        r.position = V3{0.0f, 0.0f, 0.0f};
        r.linear = V3{0.0f, 0.0f, 0.0f};
        r.angular = V3{0.0f, 0.0f, 0.0f};
        return r;
    }
    const uint32_t qo = a.root_q_offset, vo = a.root_v_offset;
    r.orientation = Q4{q[qo + 3u], q[qo + 4u], q[qo + 5u], q[qo + 6u]};
    const V3 offset = rotate(r.orientation, V3{a.root_reference[0], a.root_reference[1], a.root_reference[2]});
    r.position = add(V3{q[qo], q[qo + 1u], q[qo + 2u]}, offset);
    r.angular = V3{v[vo + 3u], v[vo + 4u], v[vo + 5u]};
    r.linear = add(V3{v[vo], v[vo + 1u], v[vo + 2u]}, cross(r.angular, offset));
    return r;
}

// Projected gravity and tilt of the root frame (:4389-4403).
__device__ __forceinline__ V3 projected_gravity(Q4 o) {
    return normalized_or(rotate_inverse(o, V3{0.0f, 0.0f, -1.0f}), V3{0.0f, 0.0f, -1.0f});
}

__device__ __forceinline__ float tilt_of(V3 g) {
    return atan2f(sqrtf(g.x * g.x + g.y * g.y), fmaxf(-g.z, 1.0e-6f));
}

// Unscaled observation value of one operator (cleanObservation, :594-1187),
// followed by transform.x * value + transform.y.
__device__ __forceinline__ float clean_observation(const mr::TaskArgs& a, const MRTaskObservationOperatorC& op,
                                                   const Root& r, const float* q, const float* v,
                                                   const MRTaskStateC& st, const float* previous_action) {
    float value = 0.0f;
    switch (op.source[0]) {
    case MR_TASK_OBSERVE_ROOT_ANGULAR_VELOCITY_LOCAL:
        value = comp(rotate_inverse(r.orientation, r.angular), op.source[2]);
        break;
    case MR_TASK_OBSERVE_PROJECTED_GRAVITY:
        value = comp(projected_gravity(r.orientation), op.source[2]);
        break;
    case MR_TASK_OBSERVE_COMMAND:
        value = op.source[2] < 3u ? st.command_and_phase[op.source[2]] : st.command_extension[op.source[2] - 3u];
        break;
    case MR_TASK_OBSERVE_JOINT_POSITION_ERROR: {
        const uint32_t qi = a.actions[op.source[1]].indices[2];
        value = q[qi] - a.default_q[qi];
        break;
    }
    case MR_TASK_OBSERVE_JOINT_VELOCITY:
        value = v[a.actions[op.source[1]].indices[3]];
        break;
    case MR_TASK_OBSERVE_PREVIOUS_ACTION:
        value = previous_action[op.source[1]];
        break;
    case MR_TASK_OBSERVE_ROOT_LINEAR_VELOCITY_LOCAL:
        value = comp(rotate_inverse(r.orientation, r.linear), op.source[2]);
        break;
    case MR_TASK_OBSERVE_ROOT_HEIGHT:
        value = r.position.z;
        break;
    default:
        value = 0.0f;
        break;
    }
    return op.transform[0] * value + op.transform[1];
}

__device__ __forceinline__ float random_signed(const mr::TaskArgs& a, uint32_t env, uint32_t episode, uint32_t step,
                                               uint32_t channel) {
    return 2.0f * random_unit(a.seed, env, episode, step, channel) - 1.0f;
}

} // namespace

// ------------------------------------------------------------- apply_actions

// Per environment: shift the action delay slots, take the delayed raw action,
// low-pass it with the binding's response fraction (clamp(1 - exp(-dt / response
// time), 0, 1), or 1 without a response time: a constant of the program that the
// host evaluates once), and write the joint-position
// target default_q + scale * filtered, clamped to the binding's bounds, into
// the implicit-drive command [E][max_nv] (every other coordinate 0).
// One wavefront per environment. Lanes over the actions shift the delay line
// and filter the response; after the barrier, lanes over the command's
// coordinates write each coordinate once: the last position-type action bound
// to it (action order), else 0.
extern "C" __global__ void __launch_bounds__(64) mr_task_apply_actions_kernel(mr::TaskArgs a, const float* actions) {
    const uint32_t env = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (env >= a.environments) return;
    const uint32_t A = a.action_count, slots = a.delay_slots;
    float* hist = a.action_history + (size_t)env * slots * A;
    float* cmd = a.targets + (size_t)env * a.max_nv;
    const MRTaskStateC& st = a.states[env];
    const uint32_t filter_slot = slots - 1u;
    const uint32_t raw_last = filter_slot - 1u;
    const uint32_t delay = st.schedule[2] < raw_last ? st.schedule[2] : raw_last;
    const uint32_t selected = raw_last - delay;
    for (uint32_t k = lane; k < A; k += 64u) {
        for (uint32_t d = 0; d < raw_last; ++d) hist[d * A + k] = hist[(d + 1u) * A + k];
        const float requested = actions[(size_t)env * A + k];
        const float previous = hist[filter_slot * A + k];
        const float fraction = a.response_fraction[k];
        hist[raw_last * A + k] = requested;
        const float delayed = hist[selected * A + k];
        hist[filter_slot * A + k] = previous + (delayed - previous) * fraction;
    }
    __syncthreads();
    for (uint32_t j = lane; j < a.max_nv; j += 64u) {
        float value = 0.0f;
        for (uint32_t k = 0; k < A; ++k) {
            const MRTaskActionBindingC& b = a.actions[k];
            if (b.indices[3] != j ||
                (b.actuator[0] != MR_TASK_ACTUATOR_JOINT_POSITION && b.actuator[0] != MR_TASK_ACTUATOR_GRIPPER_POSITION))
                continue;
            const float target = a.default_q[b.indices[2]] + b.parameters[0] * hist[filter_slot * A + k];
            value = fminf(fmaxf(target, b.parameters[1]), b.parameters[2]);
        }
        cmd[j] = value;
    }
}

// ------------------------------------------------------------------ complete

// Per environment, after the physics step: reward operators integrated over
// the control step, terminations, and the episode and command bookkeeping.
extern "C" __global__ void __launch_bounds__(64) mr_task_complete_kernel(mr::TaskArgs a, const MRWorldStatusC* status) {
    const uint32_t env = blockIdx.x * blockDim.x + threadIdx.x;
    if (env >= a.environments) return;
    const uint32_t A = a.action_count, slots = a.delay_slots;
    const float* q = a.q + (size_t)env * a.max_nq;
    const float* v = a.v + (size_t)env * a.max_nv;
    const float* hist = a.action_history + (size_t)env * slots * A;
    float* prev = a.previous_joint + (size_t)env * 2u * A;  // [velocity A][position A]
    MRTaskStateC st = a.states[env];
    const Root r = root_of(a, q, v);
    const V3 base_linear = rotate_inverse(r.orientation, r.linear);
    const V3 base_angular = rotate_inverse(r.orientation, r.angular);
    const V3 g = projected_gravity(r.orientation);
    const float height = r.position.z;
    const float tilt = tilt_of(g);
    const Q4 o = r.orientation;
    float yb0 = 1.0f - 2.0f * (o.y * o.y + o.z * o.z);
    float yb1 = 2.0f * (o.w * o.z + o.x * o.y);
    const float yinv = 1.0f / sqrtf(fmaxf(yb0 * yb0 + yb1 * yb1, 1.0e-12f));
    yb0 *= yinv;
    yb1 *= yinv;
    const float yaw_x = yb0 * r.linear.x + yb1 * r.linear.y;
    const float yaw_y = -yb1 * r.linear.x + yb0 * r.linear.y;
    const float tdx = yaw_x - st.command_and_phase[0], tdy = yaw_y - st.command_and_phase[1];
    const float tracking_error = tdx * tdx + tdy * tdy;
    const float yaw_delta = base_angular.z - st.command_and_phase[2];
    const float yaw_error = yaw_delta * yaw_delta;
    float vel2 = 0.0f, acc2 = 0.0f, rate2 = 0.0f, limit2 = 0.0f;
    const uint32_t raw_last = slots - 2u;
    for (uint32_t k = 0; k < A; ++k) {
        const MRTaskActionBindingC& b = a.actions[k];
        const float da = hist[raw_last * A + k] - hist[(raw_last - 1u) * A + k];
        rate2 += da * da;
        const float position = q[b.indices[2]];
        const float velocity = v[b.indices[3]];
        const float acceleration = (velocity - prev[k]) / a.dt;
        const MRDofC& dof = a.dofs[b.indices[1]];
        const float lower = fmaxf(dof.limits.x - position, 0.0f);
        const float upper = fmaxf(position - dof.limits.y, 0.0f);
        vel2 += velocity * velocity;
        acc2 += acceleration * acceleration;
        limit2 += lower * lower + upper * upper;
    }
    float phase = st.command_and_phase[3] + kTwoPi * a.dt / a.gait_period;
    phase = fmodf(phase, kTwoPi);
    float reward = 0.0f;
    float b0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, b1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < a.reward_count; ++i) {
        const MRTaskRewardOperatorC& op = a.rewards[i];
        float value = 0.0f;
        float* bin = &b0[1];
        switch (op.source[0]) {
        case MR_TASK_REWARD_LINEAR_VELOCITY_TRACKING:
            value = expf(-tracking_error / fmaxf(op.parameters[1], 1.0e-8f));
            bin = &b0[0];
            break;
        case MR_TASK_REWARD_YAW_VELOCITY_TRACKING:
            value = expf(-yaw_error / fmaxf(op.parameters[1], 1.0e-8f));
            bin = &b0[0];
            break;
        case MR_TASK_REWARD_CONSTANT:
            value = 1.0f;
            bin = &b0[0];
            break;
        case MR_TASK_REWARD_ROOT_VERTICAL_VELOCITY_SQUARED:
            value = base_linear.z * base_linear.z;
            break;
        case MR_TASK_REWARD_ROOT_ROLL_PITCH_VELOCITY_SQUARED:
            value = base_angular.x * base_angular.x + base_angular.y * base_angular.y;
            break;
        case MR_TASK_REWARD_TILT_SQUARED:
            value = tilt * tilt;
            break;
        case MR_TASK_REWARD_PROJECTED_GRAVITY_HORIZONTAL_SQUARED:
            value = g.x * g.x + g.y * g.y;
            break;
        case MR_TASK_REWARD_ROOT_HEIGHT_ERROR_SQUARED: {
            const float err = height - a.base_height_target;
            value = err * err;
            break;
        }
        case MR_TASK_REWARD_JOINT_VELOCITY_SQUARED:
            value = vel2;
            bin = &b0[2];
            break;
        case MR_TASK_REWARD_JOINT_ACCELERATION_SQUARED:
            value = acc2;
            bin = &b0[3];
            break;
        case MR_TASK_REWARD_ACTION_RATE_SQUARED:
            value = rate2;
            bin = &b1[0];
            break;
        case MR_TASK_REWARD_JOINT_LIMIT_VIOLATION_SQUARED:
            value = limit2;
            bin = &b1[1];
            break;
        default:
            break;
        }
        const float contribution = op.parameters[0] * value;
        reward += contribution;
        *bin += contribution;
    }
    // weights are rates: integrate over the control step
    reward *= a.dt;
    for (int k = 0; k < 4; ++k) {
        b0[k] *= a.dt;
        b1[k] *= a.dt;
    }
    const float tracking = expf(-tracking_error / 0.25f);
    const float yaw_tracking = expf(-yaw_error / 0.25f);
    const uint32_t episode_steps = st.episode[0] + 1u;
    const bool timeout = episode_steps >= a.max_episode_steps;
    bool done = false;
    uint32_t reason = MR_TASK_TERMINATION_CONTINUING, priority = 0u;
    float penalty = 0.0f;
    for (uint32_t i = 0; i < a.termination_count; ++i) {
        const MRTaskTerminationOperatorC& op = a.terminations[i];
        bool triggered = false;
        if (op.source[0] == MR_TASK_TERMINATE_MINIMUM_ROOT_HEIGHT) triggered = height < op.parameters[0];
        else if (op.source[0] == MR_TASK_TERMINATE_MAXIMUM_TILT) triggered = tilt > op.parameters[0];
        if (triggered && (!done || op.source[3] >= priority)) {
            done = true;
            reason = op.source[2];
            priority = op.source[3];
            penalty = op.parameters[1];
        }
    }
    if (timeout) {
        done = true;
        reason = MR_TASK_TERMINATION_TIMEOUT;
    }
    const bool physics_error = status[env].code != MR_STEP_SUCCESS;
    if (physics_error) {
        done = true;
        reason = MR_TASK_TERMINATION_PHYSICS_ERROR;
    }
    if (done && reason != MR_TASK_TERMINATION_TIMEOUT && reason != MR_TASK_TERMINATION_PHYSICS_ERROR) {
        reward += penalty;
        b0[1] += penalty;
    }
    const float episode_return = st.air_return_tracking[2] + reward;
    const float episode_tracking = st.air_return_tracking[3] + tracking;
    MRTaskTransitionC tr{};
    tr.reward_and_state[0] = reward;
    tr.reward_and_state[1] = 0.5f * (tracking + yaw_tracking);
    tr.reward_and_state[2] = height;
    tr.reward_and_state[3] = tilt;
    tr.termination[0] = done ? 1u : 0u;
    tr.termination[1] = timeout ? 1u : 0u;
    tr.termination[2] = physics_error ? 1u : 0u;
    tr.termination[3] = reason;
    tr.policy_revision = a.policy_revision;
    for (int k = 0; k < 4; ++k) {
        tr.reward_breakdown0[k] = b0[k];
        tr.reward_breakdown1[k] = b1[k];
    }
    tr.episode_tracking_score = done && !physics_error ? episode_tracking / fmaxf((float)episode_steps, 1.0f) : 0.0f;
    a.transitions[env] = tr;
    // command segments
    if (!done && st.schedule[0] <= 1u) {
        const V3 c = sampled_command(a, env, st.episode[1], episode_steps);
        st.command_and_phase[0] = c.x;
        st.command_and_phase[1] = c.y;
        st.command_and_phase[2] = c.z;
        st.schedule[0] = duration_steps(a, env, st.episode[1], episode_steps, 64u, a.command_seconds[0],
                                        a.command_seconds[1]);
    } else if (!done) {
        --st.schedule[0];
    }
    st.episode[0] = episode_steps;
    st.status[1] = done ? 1u : 0u;
    st.status[2] = reason;
    st.command_and_phase[3] = phase;
    st.air_return_tracking[0] = 0.0f;
    st.air_return_tracking[1] = height;
    st.air_return_tracking[2] = done ? 0.0f : episode_return;
    st.air_return_tracking[3] = done ? 0.0f : episode_tracking;
    a.states[env] = st;
    for (uint32_t k = 0; k < A; ++k) {
        prev[k] = v[a.actions[k].indices[3]];
        prev[A + k] = q[a.actions[k].indices[2]];
    }
}

// ------------------------------------------------------------------- observe

// Per environment, before the next action: restart finished (or never
// started) episodes from the reset state, raising the reset mask of the next
// physics step, then write the actor frame into the observation history
// (newest frame last; a new episode fills every frame).
// One wavefront per environment, lanes over the observation operators: every
// lane derives the environment's state, root and normalised gravity the same
// way, lane 0 writes the records, and each lane writes its own columns of the
// history (coalesced across the wavefront).
extern "C" __global__ void __launch_bounds__(64) mr_task_observe_kernel(mr::TaskArgs a) {
    const uint32_t env = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (env >= a.environments) return;
    const uint32_t A = a.action_count, O = a.observation_count, H = a.history_length;
    MRTaskStateC st = a.states[env];
    const bool reset = st.status[0] == 0u || st.status[1] != 0u;
    if (lane == 0u) a.reset_mask[env] = reset ? 1u : 0u;
    const float* q = (reset ? a.reset_q : a.q) + (size_t)env * a.max_nq;
    const float* v = (reset ? a.reset_v : a.v) + (size_t)env * a.max_nv;
    float* prev = a.previous_joint + (size_t)env * 2u * A;
    if (reset) {
        const uint32_t episode = st.episode[1] + 1u;
        const Root r0 = root_of(a, q, v);
        for (int k = 0; k < 4; ++k) {
            st.episode[k] = 0u;
            st.schedule[k] = 0u;
            st.status[k] = 0u;
            st.command_and_phase[k] = 0.0f;
            st.command_extension[k] = 0.0f;
            st.air_return_tracking[k] = 0.0f;
            st.recovery[k] = 0.0f;
            st.recovery_stats[k] = 0u;
            st.threat_geometry[k] = 0.0f;
            st.threat_teacher[k] = 0.0f;
            st.threat_metadata[k] = 0u;
        }
        st.episode[1] = episode;
        st.schedule[0] = duration_steps(a, env, episode, 0u, 32u, a.command_seconds[0], a.command_seconds[1]);
        st.schedule[2] = a.actuator_delay;
        st.status[0] = 1u;
        const V3 c = sampled_command(a, env, episode, 0u);
        st.command_and_phase[0] = c.x;
        st.command_and_phase[1] = c.y;
        st.command_and_phase[2] = c.z;
        st.air_return_tracking[1] = r0.position.z;
        st.threat_metadata[0] = MR_INVALID_INDEX;
        st.threat_metadata[3] = MR_INVALID_INDEX;
        for (uint32_t k = lane; k < A; k += 64u) {
            prev[k] = 0.0f;
            prev[A + k] = q[a.actions[k].indices[2]];
        }
        if (lane == 0u) a.states[env] = st;
    }
    const Root r = root_of(a, q, v);
    const float* previous_action = a.action_history + ((size_t)env * a.delay_slots + (a.delay_slots - 2u)) * A;
    const uint32_t episode = st.episode[1], step = st.episode[0];
    // projected gravity components that are normalized together (writeFrame, :1190-1270)
    V3 gn{0.0f, 0.0f, -1.0f};
    bool have_gn = false;
    for (uint32_t i = 0; i < O; ++i) {
        const MRTaskObservationOperatorC& op = a.observations[i];
        if (op.source[0] != MR_TASK_OBSERVE_PROJECTED_GRAVITY ||
            (op.source[3] & MR_TASK_OBSERVATION_NORMALIZE_VECTOR3) == 0u)
            continue;
        const float value = clean_observation(a, op, r, q, v, st, previous_action) +
            op.transform[2] * random_signed(a, env, episode, step, op.auxiliary[1]);
        if (op.source[2] == 0u) gn.x = value;
        else if (op.source[2] == 1u) gn.y = value;
        else gn.z = value;
        have_gn = true;
    }
    if (have_gn) gn = normalized_or(gn, V3{0.0f, 0.0f, -1.0f});
    float* hist = a.observations_out + (size_t)env * H * O;
    for (uint32_t i = lane; i < O; i += 64u) {
        const MRTaskObservationOperatorC& op = a.observations[i];
        float out;
        if (op.source[0] == MR_TASK_OBSERVE_PROJECTED_GRAVITY &&
            (op.source[3] & MR_TASK_OBSERVATION_NORMALIZE_VECTOR3) != 0u) {
            out = comp(gn, op.source[2]);
        } else {
            out = clean_observation(a, op, r, q, v, st, previous_action) +
                op.transform[2] * random_signed(a, env, episode, step, op.auxiliary[1]);
        }
        if (reset) {
            for (uint32_t h = 0; h < H; ++h) hist[h * O + i] = out;
        } else {
            for (uint32_t h = 0; h + 1u < H; ++h) hist[h * O + i] = hist[(h + 1u) * O + i];
            hist[(H - 1u) * O + i] = out;
        }
    }
}

// ------------------------------------------------------------------- policy

namespace {

__device__ __forceinline__ float activation(uint32_t op, float x) {
    switch (op) {
    case MR_POLICY_ACTIVATION_IDENTITY: return x;
    case MR_POLICY_ACTIVATION_RELU: return fmaxf(x, 0.0f);
    case MR_POLICY_ACTIVATION_TANH: return tanhf(x);
    case MR_POLICY_ACTIVATION_ELU: return x >= 0.0f ? x : expf(x) - 1.0f;
    case MR_POLICY_ACTIVATION_SILU: return x / (1.0f + expf(-x));
    default: return 0.0f;
    }
}

__device__ __forceinline__ float policy_uniform(unsigned long long key) {
    const uint32_t sample = (uint32_t)(mix64(key) >> 40u);
    return ((float)sample + 0.5f) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float policy_normal(unsigned long long key, uint32_t action) {
    const unsigned long long channel = key ^ mix64((unsigned long long)action * 2ull);
    const float first = policy_uniform(channel);
    const float second = policy_uniform(channel ^ 0xd2b74407b1ce6e93ull);
    return sqrtf(-2.0f * logf(first)) * cosf(6.2831853071795864769f * second);
}

} // namespace

// Dense layer over a tile of environments x output neurons per workgroup of
// 256 threads, each thread holding a 4 x 4 block of outputs in registers.
// The features pass in chunks of 16: the chunk's normalised inputs and its
// weights are staged in LDS once per workgroup (transposed, so a thread reads
// its four inputs and four weights of one feature as two 16-byte loads).
// Every output is still one chain of fused multiply-adds in feature order,
// starting from the bias, so the result does not depend on the tile shape.
constexpr uint32_t kDenseChunk = 16u;

template <uint32_t kNeurons>
__global__ void __launch_bounds__(256)
mr_policy_dense_tile_kernel(const float* weights, const float* bias, const float* mean, const float* inverse_std,
                            float clip, const float* input, uint32_t input_stride, float* output,
                            uint32_t output_stride, uint32_t environments, uint32_t in_width, uint32_t out_width,
                            uint32_t act) {
    constexpr uint32_t kEnvs = 4096u / kNeurons;        // (kEnvs / 4) * (kNeurons / 4) = 256 threads
    constexpr uint32_t kPad = 4u;                       // keeps rows 16-byte aligned, spreads the stores
    __shared__ __align__(16) float xs[kDenseChunk][kEnvs + kPad];
    __shared__ __align__(16) float ws[kDenseChunk][kNeurons + kPad];
    const uint32_t t = threadIdx.x;
    const uint32_t env0 = blockIdx.x * kEnvs;
    const uint32_t n0 = blockIdx.y * kNeurons;
    const uint32_t tn = t % (kNeurons / 4u);
    const uint32_t te = t / (kNeurons / 4u);
    float acc[4][4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        const uint32_t n = n0 + tn * 4u + j;
        const float b = n < out_width ? bias[n] : 0.0f;
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) acc[i][j] = b;
    }
    for (uint32_t f0 = 0; f0 < in_width; f0 += kDenseChunk) {
        const uint32_t span = min(kDenseChunk, in_width - f0);
        __syncthreads();
        for (uint32_t q = t; q < kEnvs * kDenseChunk; q += 256u) {
            const uint32_t e = q / kDenseChunk, k = q % kDenseChunk;
            float sample = 0.0f;
            if (env0 + e < environments && k < span) {
                sample = input[(size_t)(env0 + e) * input_stride + f0 + k];
                if (mean != nullptr)
                    sample = fminf(fmaxf((sample - mean[f0 + k]) * inverse_std[f0 + k], -clip), clip);
            }
            xs[k][e] = sample;
        }
        for (uint32_t q = t; q < kNeurons * kDenseChunk; q += 256u) {
            const uint32_t n = q / kDenseChunk, k = q % kDenseChunk;
            ws[k][n] = (n0 + n < out_width && k < span) ? weights[(size_t)(n0 + n) * in_width + f0 + k] : 0.0f;
        }
        __syncthreads();
        const auto feature = [&](uint32_t k) {
            const float4 xv = *reinterpret_cast<const float4*>(&xs[k][te * 4u]);
            const float4 wv = *reinterpret_cast<const float4*>(&ws[k][tn * 4u]);
            const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
            const float wa[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i)
#pragma unroll
                for (uint32_t j = 0; j < 4u; ++j) acc[i][j] = fmaf(wa[j], xa[i], acc[i][j]);
        };
        if (span == kDenseChunk) {
            // full chunk: unrolled, so the LDS reads issue ahead of the FMAs
#pragma unroll
            for (uint32_t k = 0; k < kDenseChunk; ++k) feature(k);
        } else {
            for (uint32_t k = 0; k < span; ++k) feature(k);
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) {
        const uint32_t env = env0 + te * 4u + i;
        if (env >= environments) continue;
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            const uint32_t n = n0 + tn * 4u + j;
            if (n < out_width) output[(size_t)env * output_stride + n] = activation(act, acc[i][j]);
        }
    }
}

// One thread per environment: Gaussian sample of the actor's mean, its log
// probability, and the transformed, clipped action.
extern "C" __global__ void __launch_bounds__(64)
mr_policy_sample_kernel(const float* mean, uint32_t mean_stride, const float* action_bias, const float* action_scale,
                        const float* log_std, float action_clip, unsigned long long seed, unsigned long long revision,
                        const MRTaskStateC* states, uint32_t environments, uint32_t actions, uint32_t control_step,
                        float* actions_out, float* latents_out, float* log_prob_out) {
    const uint32_t env = blockIdx.x * blockDim.x + threadIdx.x;
    if (env >= environments) return;
    const MRTaskStateC st = states[env];
    unsigned long long key = seed;
    key ^= mix64(((unsigned long long)env << 32u) | (unsigned long long)st.episode[1]);
    key ^= mix64(((unsigned long long)st.episode[0] << 32u) | (unsigned long long)control_step);
    key ^= mix64(revision);
    constexpr float kHalfLogTwoPi = 0.91893853320467274178f;
    float logp = 0.0f;
    for (uint32_t k = 0; k < actions; ++k) {
        const float m = mean[(size_t)env * mean_stride + k];
        float latent = m;
        if (log_std != nullptr) {
            const float ls = log_std[k];
            const float n = policy_normal(key, k);
            latent = fmaf(expf(ls), n, m);
            logp += -0.5f * n * n - ls - kHalfLogTwoPi;
        }
        latents_out[(size_t)env * actions + k] = latent;
        actions_out[(size_t)env * actions + k] =
            fminf(fmaxf(fmaf(action_scale[k], latent, action_bias[k]), -action_clip), action_clip);
    }
    log_prob_out[env] = logp;
}

namespace mr {

hipError_t launch_task_apply_actions(const TaskArgs& a, const float* actions, hipStream_t stream) {
    if (a.environments == 0u) return hipSuccess;
    hipLaunchKernelGGL(mr_task_apply_actions_kernel, dim3(a.environments), dim3(64), 0, stream, a, actions);
    return hipGetLastError();
}

hipError_t launch_task_complete(const TaskArgs& a, const MRWorldStatusC* status, hipStream_t stream) {
    hipLaunchKernelGGL(mr_task_complete_kernel, dim3((a.environments + 63u) / 64u), dim3(64), 0, stream, a, status);
    return hipGetLastError();
}

hipError_t launch_task_observe(const TaskArgs& a, hipStream_t stream) {
    if (a.environments == 0u) return hipSuccess;
    hipLaunchKernelGGL(mr_task_observe_kernel, dim3(a.environments), dim3(64), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_policy_dense(const float* weights, const float* bias, const float* mean, const float* inverse_std,
                               float clip, const float* input, uint32_t input_stride, float* output,
                               uint32_t output_stride, uint32_t environments, uint32_t in_width, uint32_t out_width,
                               uint32_t act, hipStream_t stream) {
    if (environments == 0u || out_width == 0u) return hipSuccess;
    // the neuron tile is the smallest of 64 / 32 / 16 / 8 that covers the
    // layer (or 64), so narrow heads do not run mostly empty tiles
    const auto launch = [&](auto kernel, uint32_t neurons) {
        const uint32_t envs = 4096u / neurons;
        const dim3 grid((environments + envs - 1u) / envs, (out_width + neurons - 1u) / neurons);
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, stream, weights, bias, mean, inverse_std, clip, input,
                           input_stride, output, output_stride, environments, in_width, out_width, act);
    };
    if (out_width > 32u) launch(mr_policy_dense_tile_kernel<64u>, 64u);
    else if (out_width > 16u) launch(mr_policy_dense_tile_kernel<32u>, 32u);
    else if (out_width > 8u) launch(mr_policy_dense_tile_kernel<16u>, 16u);
    else launch(mr_policy_dense_tile_kernel<8u>, 8u);
    return hipGetLastError();
}

hipError_t launch_policy_sample(const float* mean, uint32_t mean_stride, const float* action_bias,
                                const float* action_scale, const float* log_std, float action_clip,
                                unsigned long long seed, unsigned long long revision, const MRTaskStateC* states,
                                uint32_t environments, uint32_t actions, uint32_t control_step, float* actions_out,
                                float* latents_out, float* log_prob_out, hipStream_t stream) {
    if (environments == 0u) return hipSuccess;
    hipLaunchKernelGGL(mr_policy_sample_kernel, dim3((environments + 63u) / 64u), dim3(64), 0, stream, mean,
                       mean_stride, action_bias, action_scale, log_std, action_clip, seed, revision, states,
                       environments, actions, control_step, actions_out, latents_out, log_prob_out);
    return hipGetLastError();
}

} // namespace mr
