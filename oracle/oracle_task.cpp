// TEST INFRASTRUCTURE — the CPU oracle of the locomotion task layer and the
// policy layers. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load or run anything under oracle/; the product never
// does.
//
// FP32 restatement, one environment at a time, of the reference kernels
//   apply_actions        src/metal/LocomotionTask.metal:3442-3674
//   complete (subset)    src/metal/LocomotionTask.metal:4375-4500, 4700-4760,
//                        5320-5360, 5960-6160, 6370-6440, 6600-6630
//   observe (subset)     src/metal/LocomotionTask.metal:2330-3010 (reset),
//                        594-1187 (cleanObservation), 1190-1330 (writeFrame)
//   SplitMix64 keys      src/metal/LocomotionTask.metal:121-190, 1484-1615
//   dense layer          src/metal/PolicyInference.metal:78-170
//   sample and score     src/metal/PolicyInference.metal:172-268
// restricted to the opcodes include/mr_task.h lists, written with a fixed
// evaluation order and compiled with -ffp-contract=off.
//
// What pins it (tests/test_task.py, CPU): the SplitMix64 finalizer against
// the published sequence; projected gravity, tilt and local velocities of
// roots at closed-form rotations; the yaw-frame tracking error of a rotated
// root; the first-order response filter and delay-slot selection against
// their closed forms; reward integration, failure penalties, timeout and
// termination priorities on hand-built states; command resampling at the end
// of a segment; the dense layer against numpy float64 with the same fused
// multiply-add order; the Gaussian sample's log probability against the
// closed-form density.

#include "oracle_math.h"

#include "../include/mr_task.h"
#include "../include/mr_world_types.h"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace {

using oracle::Q4;
using oracle::V3;
using oracle::add;
using oracle::cross;
using oracle::dot;
using oracle::mul;
using oracle::sub;

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;

uint64_t mix64(uint64_t value) {
    value += 0x9e3779b97f4a7c15ull;
    value = (value ^ (value >> 30u)) * 0xbf58476d1ce4e5b9ull;
    value = (value ^ (value >> 27u)) * 0x94d049bb133111ebull;
    return value ^ (value >> 31u);
}

float random_unit(uint64_t seed, uint32_t env, uint32_t episode, uint32_t step, uint32_t channel) {
    uint64_t key = seed;
    key ^= (uint64_t)(env + 1u) * 0xd2b74407b1ce6e93ull;
    key ^= (uint64_t)(episode + 1u) * 0xca5a826395121157ull;
    key ^= (uint64_t)(step + 1u) * 0x9e3779b185ebca87ull;
    key ^= (uint64_t)(channel + 1u) * 0x94d049bb133111ebull;
    return (float)(uint32_t)(mix64(key) >> 40u) * (1.0f / 16777216.0f);
}

float random_range(uint64_t seed, uint32_t env, uint32_t episode, uint32_t step, uint32_t channel, float lo,
                   float hi) {
    return lo + (hi - lo) * random_unit(seed, env, episode, step, channel);
}

float random_signed(uint64_t seed, uint32_t env, uint32_t episode, uint32_t step, uint32_t channel) {
    return 2.0f * random_unit(seed, env, episode, step, channel) - 1.0f;
}

uint32_t duration_steps(const MRLocomotionTaskProgramC& p, float dt, uint32_t env, uint32_t episode, uint32_t step,
                        uint32_t channel) {
    const float s = random_range(p.seed, env, episode, step, channel, p.command_seconds[0], p.command_seconds[1]) / dt;
    const uint32_t n = (uint32_t)std::floor(s);
    return n > 1u ? n : 1u;
}

V3 sampled_command(const MRLocomotionTaskProgramC& p, uint32_t env, uint32_t episode, uint32_t step) {
    float c[3];
    for (uint32_t k = 0; k < 3u; ++k)
        c[k] = random_range(p.seed, env, episode, step, 16u + k, p.command_lower[k], p.command_upper[k]);
    if (random_unit(p.seed, env, episode, step, 19u) < p.command_lower[3]) c[0] = c[1] = c[2] = 0.0f;
    return V3{c[0], c[1], c[2]};
}

V3 rotate(Q4 q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = mul(2.0f, cross(u, v));
    return add(add(v, mul(q.w, t)), cross(u, t));
}

V3 rotate_inverse(Q4 q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = mul(2.0f, cross(u, v));
    return add(sub(v, mul(q.w, t)), cross(u, t));
}

float comp(V3 v, uint32_t i) { return i == 0u ? v.x : (i == 1u ? v.y : v.z); }

V3 normalized_or(V3 v, V3 fallback) {
    const float l2 = dot(v, v);
    if (!(l2 > 1.0e-12f) || !std::isfinite(l2)) return fallback;
    return mul(1.0f / std::sqrt(l2), v);
}

struct Root {
    Q4 orientation;
    V3 position, linear, angular;
};

Root root_of(const MRLocomotionTaskProgramC& p, const float* q, const float* v) {
    Root r;
    if ((p.flags & MR_TASK_PROGRAM_FIXED_ROOT) != 0u) {
        r.orientation = Q4{0.0f, 0.0f, 0.0f, 1.0f};
        r.position = r.linear = r.angular = V3{0.0f, 0.0f, 0.0f};
        return r;
    }
    const uint32_t qo = p.root_q_offset, vo = p.root_v_offset;
    r.orientation = Q4{q[qo + 3u], q[qo + 4u], q[qo + 5u], q[qo + 6u]};
    const V3 offset = rotate(r.orientation, V3{p.root_reference[0], p.root_reference[1], p.root_reference[2]});
    r.position = add(V3{q[qo], q[qo + 1u], q[qo + 2u]}, offset);
    r.angular = V3{v[vo + 3u], v[vo + 4u], v[vo + 5u]};
    r.linear = add(V3{v[vo], v[vo + 1u], v[vo + 2u]}, cross(r.angular, offset));
    return r;
}

V3 projected_gravity(Q4 o) {
    return normalized_or(rotate_inverse(o, V3{0.0f, 0.0f, -1.0f}), V3{0.0f, 0.0f, -1.0f});
}

float tilt_of(V3 g) { return std::atan2(std::sqrt(g.x * g.x + g.y * g.y), std::max(-g.z, 1.0e-6f)); }

float clean_observation(const MRLocomotionTaskProgramC& p, const MRTaskObservationOperatorC& op, const Root& r,
                        const float* q, const float* v, const MRTaskStateC& st, const float* previous_action) {
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
        const uint32_t qi = p.actions[op.source[1]].indices[2];
        value = q[qi] - p.default_q[qi];
        break;
    }
    case MR_TASK_OBSERVE_JOINT_VELOCITY:
        value = v[p.actions[op.source[1]].indices[3]];
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
        break;
    }
    return op.transform[0] * value + op.transform[1];
}

float response_fraction(float dt, float response) {
    if (!(response > 0.0f)) return 1.0f;
    const double f = 1.0 - std::exp(-(double)dt / (double)response);
    return (float)std::min(1.0, std::max(0.0, f));
}

float activation(uint32_t op, float x) {
    switch (op) {
    case MR_POLICY_ACTIVATION_IDENTITY: return x;
    case MR_POLICY_ACTIVATION_RELU: return std::max(x, 0.0f);
    case MR_POLICY_ACTIVATION_TANH: return std::tanh(x);
    case MR_POLICY_ACTIVATION_ELU: return x >= 0.0f ? x : std::exp(x) - 1.0f;
    case MR_POLICY_ACTIVATION_SILU: return x / (1.0f + std::exp(-x));
    default: return 0.0f;
    }
}

float policy_uniform(uint64_t key) {
    const uint32_t sample = (uint32_t)(mix64(key) >> 40u);
    return ((float)sample + 0.5f) * (1.0f / 16777216.0f);
}

float policy_normal(uint64_t key, uint32_t action) {
    const uint64_t channel = key ^ mix64((uint64_t)action * 2ull);
    const float first = policy_uniform(channel);
    const float second = policy_uniform(channel ^ 0xd2b74407b1ce6e93ull);
    return std::sqrt(-2.0f * std::log(first)) * std::cos(6.2831853071795864769f * second);
}

} // namespace

extern "C" {

uint64_t mr_oracle_mix64(uint64_t value) { return mix64(value); }

float mr_oracle_task_random_unit(uint64_t seed, uint32_t env, uint32_t episode, uint32_t step, uint32_t channel) {
    return random_unit(seed, env, episode, step, channel);
}

// apply_actions for environments [0, E): action history [E][slots][A], actions
// [E][A], targets [E][max_nv] (overwritten), dt the control timestep.
void mr_oracle_task_apply_actions(const MRLocomotionTaskProgramC* pp, uint32_t E, uint32_t max_nv, float dt,
                                  const MRTaskStateC* states, float* history, const float* actions, float* targets) {
    const MRLocomotionTaskProgramC& p = *pp;
    const uint32_t A = p.action_count, slots = p.delay_slots;
    for (uint32_t env = 0; env < E; ++env) {
        float* hist = history + (size_t)env * slots * A;
        float* cmd = targets + (size_t)env * max_nv;
        for (uint32_t j = 0; j < max_nv; ++j) cmd[j] = 0.0f;
        const uint32_t filter_slot = slots - 1u, raw_last = filter_slot - 1u;
        const uint32_t delay = std::min(states[env].schedule[2], raw_last);
        const uint32_t selected = raw_last - delay;
        for (uint32_t k = 0; k < A; ++k) {
            for (uint32_t d = 0; d < raw_last; ++d) hist[d * A + k] = hist[(d + 1u) * A + k];
            const float requested = actions[(size_t)env * A + k];
            const float previous = hist[filter_slot * A + k];
            const MRTaskActionBindingC& b = p.actions[k];
            const float fraction = response_fraction(dt, b.parameters[3]);
            hist[raw_last * A + k] = requested;
            const float delayed = hist[selected * A + k];
            hist[filter_slot * A + k] = previous + (delayed - previous) * fraction;
            const float filtered = hist[filter_slot * A + k];
            if (b.actuator[0] != MR_TASK_ACTUATOR_JOINT_POSITION && b.actuator[0] != MR_TASK_ACTUATOR_GRIPPER_POSITION)
                continue;
            const float target = p.default_q[b.indices[2]] + b.parameters[0] * filtered;
            cmd[b.indices[3]] = std::min(std::max(target, b.parameters[1]), b.parameters[2]);
        }
    }
}

// complete for environments [0, E): q [E][max_nq], v [E][max_nv] after the
// physics step, dofs the model's DoF table, status_codes [E] of the step;
// policy_revision is stamped on every transition.
void mr_oracle_task_complete(const MRLocomotionTaskProgramC* pp, uint32_t E, uint32_t max_nq, uint32_t max_nv,
                             float dt, const float* qs, const float* vs, const MRDofC* dofs,
                             const uint32_t* status_codes, MRTaskStateC* states, const float* history,
                             float* previous_joint, MRTaskTransitionC* transitions, uint64_t policy_revision) {
    const MRLocomotionTaskProgramC& p = *pp;
    const uint32_t A = p.action_count, slots = p.delay_slots;
    for (uint32_t env = 0; env < E; ++env) {
        const float* q = qs + (size_t)env * max_nq;
        const float* v = vs + (size_t)env * max_nv;
        const float* hist = history + (size_t)env * slots * A;
        float* prev = previous_joint + (size_t)env * 2u * A;
        MRTaskStateC st = states[env];
        const Root r = root_of(p, q, v);
        const V3 base_linear = rotate_inverse(r.orientation, r.linear);
        const V3 base_angular = rotate_inverse(r.orientation, r.angular);
        const V3 g = projected_gravity(r.orientation);
        const float height = r.position.z;
        const float tilt = tilt_of(g);
        const Q4 o = r.orientation;
        float yb0 = 1.0f - 2.0f * (o.y * o.y + o.z * o.z);
        float yb1 = 2.0f * (o.w * o.z + o.x * o.y);
        const float yinv = 1.0f / std::sqrt(std::max(yb0 * yb0 + yb1 * yb1, 1.0e-12f));
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
            const MRTaskActionBindingC& b = p.actions[k];
            const float da = hist[raw_last * A + k] - hist[(raw_last - 1u) * A + k];
            rate2 += da * da;
            const float position = q[b.indices[2]];
            const float velocity = v[b.indices[3]];
            const float acceleration = (velocity - prev[k]) / dt;
            const MRDofC& dof = dofs[b.indices[1]];
            const float lower = std::max(dof.limits.x - position, 0.0f);
            const float upper = std::max(position - dof.limits.y, 0.0f);
            vel2 += velocity * velocity;
            acc2 += acceleration * acceleration;
            limit2 += lower * lower + upper * upper;
        }
        float phase = st.command_and_phase[3] + kTwoPi * dt / p.gait_period;
        phase = std::fmod(phase, kTwoPi);
        float reward = 0.0f;
        float b0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, b1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t i = 0; i < p.reward_count; ++i) {
            const MRTaskRewardOperatorC& op = p.rewards[i];
            float value = 0.0f;
            float* bin = &b0[1];
            switch (op.source[0]) {
            case MR_TASK_REWARD_LINEAR_VELOCITY_TRACKING:
                value = std::exp(-tracking_error / std::max(op.parameters[1], 1.0e-8f));
                bin = &b0[0];
                break;
            case MR_TASK_REWARD_YAW_VELOCITY_TRACKING:
                value = std::exp(-yaw_error / std::max(op.parameters[1], 1.0e-8f));
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
                const float err = height - p.base_height_target;
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
        reward *= dt;
        for (int k = 0; k < 4; ++k) {
            b0[k] *= dt;
            b1[k] *= dt;
        }
        const float tracking = std::exp(-tracking_error / 0.25f);
        const float yaw_tracking = std::exp(-yaw_error / 0.25f);
        const uint32_t episode_steps = st.episode[0] + 1u;
        const bool timeout = episode_steps >= p.max_episode_steps;
        bool done = false;
        uint32_t reason = MR_TASK_TERMINATION_CONTINUING, priority = 0u;
        float penalty = 0.0f;
        for (uint32_t i = 0; i < p.termination_count; ++i) {
            const MRTaskTerminationOperatorC& op = p.terminations[i];
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
        const bool physics_error = status_codes[env] != MR_STEP_SUCCESS;
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
        tr.policy_revision = policy_revision;
        for (int k = 0; k < 4; ++k) {
            tr.reward_breakdown0[k] = b0[k];
            tr.reward_breakdown1[k] = b1[k];
        }
        tr.episode_tracking_score =
            done && !physics_error ? episode_tracking / std::max((float)episode_steps, 1.0f) : 0.0f;
        transitions[env] = tr;
        if (!done && st.schedule[0] <= 1u) {
            const V3 c = sampled_command(p, env, st.episode[1], episode_steps);
            st.command_and_phase[0] = c.x;
            st.command_and_phase[1] = c.y;
            st.command_and_phase[2] = c.z;
            st.schedule[0] = duration_steps(p, dt, env, st.episode[1], episode_steps, 64u);
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
        states[env] = st;
        for (uint32_t k = 0; k < A; ++k) {
            prev[k] = v[p.actions[k].indices[3]];
            prev[A + k] = q[p.actions[k].indices[2]];
        }
    }
}

// observe for environments [0, E): restarts finished episodes (reset_mask[e]
// = 1 and reads reset_q/reset_v), then writes the actor history
// [E][history][O], newest frame last.
void mr_oracle_task_observe(const MRLocomotionTaskProgramC* pp, uint32_t E, uint32_t max_nq, uint32_t max_nv,
                            float dt, const float* qs, const float* vs, const float* reset_qs, const float* reset_vs,
                            MRTaskStateC* states, const float* history, float* previous_joint, float* observations,
                            uint32_t* reset_mask) {
    const MRLocomotionTaskProgramC& p = *pp;
    const uint32_t A = p.action_count, O = p.observation_count, H = p.history_length;
    for (uint32_t env = 0; env < E; ++env) {
        MRTaskStateC st = states[env];
        const bool reset = st.status[0] == 0u || st.status[1] != 0u;
        reset_mask[env] = reset ? 1u : 0u;
        const float* q = (reset ? reset_qs : qs) + (size_t)env * max_nq;
        const float* v = (reset ? reset_vs : vs) + (size_t)env * max_nv;
        float* prev = previous_joint + (size_t)env * 2u * A;
        if (reset) {
            const uint32_t episode = st.episode[1] + 1u;
            const Root r0 = root_of(p, q, v);
            st = MRTaskStateC{};
            st.episode[1] = episode;
            st.schedule[0] = duration_steps(p, dt, env, episode, 0u, 32u);
            st.schedule[2] = p.actuator_delay;
            st.status[0] = 1u;
            const V3 c = sampled_command(p, env, episode, 0u);
            st.command_and_phase[0] = c.x;
            st.command_and_phase[1] = c.y;
            st.command_and_phase[2] = c.z;
            st.air_return_tracking[1] = r0.position.z;
            st.threat_metadata[0] = MR_INVALID_INDEX;
            st.threat_metadata[3] = MR_INVALID_INDEX;
            for (uint32_t k = 0; k < A; ++k) {
                prev[k] = 0.0f;
                prev[A + k] = q[p.actions[k].indices[2]];
            }
            states[env] = st;
        }
        const Root r = root_of(p, q, v);
        const float* previous_action = history + ((size_t)env * p.delay_slots + (p.delay_slots - 2u)) * A;
        const uint32_t episode = st.episode[1], step = st.episode[0];
        V3 gn{0.0f, 0.0f, -1.0f};
        bool have_gn = false;
        for (uint32_t i = 0; i < O; ++i) {
            const MRTaskObservationOperatorC& op = p.observations[i];
            if (op.source[0] != MR_TASK_OBSERVE_PROJECTED_GRAVITY ||
                (op.source[3] & MR_TASK_OBSERVATION_NORMALIZE_VECTOR3) == 0u)
                continue;
            const float value = clean_observation(p, op, r, q, v, st, previous_action) +
                op.transform[2] * random_signed(p.seed, env, episode, step, op.auxiliary[1]);
            if (op.source[2] == 0u) gn.x = value;
            else if (op.source[2] == 1u) gn.y = value;
            else gn.z = value;
            have_gn = true;
        }
        if (have_gn) gn = normalized_or(gn, V3{0.0f, 0.0f, -1.0f});
        float* hist = observations + (size_t)env * H * O;
        for (uint32_t i = 0; i < O; ++i) {
            const MRTaskObservationOperatorC& op = p.observations[i];
            float out;
            if (op.source[0] == MR_TASK_OBSERVE_PROJECTED_GRAVITY &&
                (op.source[3] & MR_TASK_OBSERVATION_NORMALIZE_VECTOR3) != 0u) {
                out = comp(gn, op.source[2]);
            } else {
                out = clean_observation(p, op, r, q, v, st, previous_action) +
                    op.transform[2] * random_signed(p.seed, env, episode, step, op.auxiliary[1]);
            }
            if (reset) {
                for (uint32_t h = 0; h < H; ++h) hist[h * O + i] = out;
            } else {
                for (uint32_t h = 0; h + 1u < H; ++h) hist[h * O + i] = hist[(h + 1u) * O + i];
                hist[(H - 1u) * O + i] = out;
            }
        }
    }
}

void mr_oracle_policy_dense(const float* weights, const float* bias, const float* mean, const float* inverse_std,
                            float clip, const float* input, uint32_t input_stride, float* output,
                            uint32_t output_stride, uint32_t environments, uint32_t in_width, uint32_t out_width,
                            uint32_t act) {
    for (uint32_t env = 0; env < environments; ++env)
        for (uint32_t n = 0; n < out_width; ++n) {
            const float* x = input + (size_t)env * input_stride;
            const float* w = weights + (size_t)n * in_width;
            float value = bias[n];
            for (uint32_t f = 0; f < in_width; ++f) {
                float sample = x[f];
                if (mean != nullptr) sample = std::min(std::max((sample - mean[f]) * inverse_std[f], -clip), clip);
                value = std::fma(w[f], sample, value);
            }
            output[(size_t)env * output_stride + n] = activation(act, value);
        }
}

void mr_oracle_policy_sample(const float* mean, uint32_t mean_stride, const float* action_bias,
                             const float* action_scale, const float* log_std, float action_clip, uint64_t seed,
                             uint64_t revision, const MRTaskStateC* states, uint32_t environments, uint32_t actions,
                             uint32_t control_step, float* actions_out, float* latents_out, float* log_prob_out) {
    constexpr float kHalfLogTwoPi = 0.91893853320467274178f;
    for (uint32_t env = 0; env < environments; ++env) {
        const MRTaskStateC& st = states[env];
        uint64_t key = seed;
        key ^= mix64(((uint64_t)env << 32u) | (uint64_t)st.episode[1]);
        key ^= mix64(((uint64_t)st.episode[0] << 32u) | (uint64_t)control_step);
        key ^= mix64(revision);
        float logp = 0.0f;
        for (uint32_t k = 0; k < actions; ++k) {
            const float m = mean[(size_t)env * mean_stride + k];
            float latent = m;
            if (log_std != nullptr) {
                const float ls = log_std[k];
                const float n = policy_normal(key, k);
                latent = std::fma(std::exp(ls), n, m);
                logp += -0.5f * n * n - ls - kHalfLogTwoPi;
            }
            latents_out[(size_t)env * actions + k] = latent;
            actions_out[(size_t)env * actions + k] =
                std::min(std::max(std::fma(action_scale[k], latent, action_bias[k]), -action_clip), action_clip);
        }
        log_prob_out[env] = logp;
    }
}

} // extern "C"
