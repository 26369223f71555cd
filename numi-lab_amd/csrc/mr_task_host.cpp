// Host executor of the locomotion task layer: program validation, device
// tables and the per-control-step sequence apply_actions -> physics ->
// complete -> observe on the world's stream.
#include "mr_task_host.h"
#include "mr_policy.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>

namespace mr {

namespace {

template <typename T>
void upload(DeviceBuffer& b, const T* data, size_t count) {
    b.reserve(std::max<size_t>(count * sizeof(T), 16u));
    if (count > 0u) hip_check(hipMemcpy(b.ptr, data, count * sizeof(T), hipMemcpyHostToDevice), "task upload");
}

void require(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error("locomotion task: " + what);
}

bool supported_reward(uint32_t op) {
    return op <= MR_TASK_REWARD_JOINT_LIMIT_VIOLATION_SQUARED || op == MR_TASK_REWARD_PROJECTED_GRAVITY_HORIZONTAL_SQUARED;
}

} // namespace

float task_response_fraction(float dt, float response_seconds) {
    if (!(response_seconds > 0.0f)) return 1.0f;
    const double f = 1.0 - std::exp(-(double)dt / (double)response_seconds);
    return (float)std::min(1.0, std::max(0.0, f));
}

LocomotionTask::LocomotionTask(const MRLocomotionTaskProgramC& p, World& world) : world_(world) {
    const ModelData& m = world.model();
    const uint32_t E = world.environment_count();
    for (uint32_t e = 0; e < E; ++e)
        require(m.environment_model[e] == m.environment_model[0], "every environment must share one model");
    const MRModelHeaderC& h = m.models[m.environment_model[0]];
    const uint32_t nq = h.nq, nv = h.nv;
    require(p.action_count >= 1u && p.action_count <= 64u && p.actions != nullptr, "action_count in [1, 64]");
    require(p.observation_count >= 1u && p.observation_count <= 1024u && p.observations != nullptr,
            "observation_count in [1, 1024]");
    require(p.reward_count <= 64u && (p.reward_count == 0u || p.rewards != nullptr), "reward_count in [0, 64]");
    require(p.termination_count <= 16u && (p.termination_count == 0u || p.terminations != nullptr),
            "termination_count in [0, 16]");
    require(p.default_q != nullptr, "default_q is required");
    require(p.history_length >= 1u && p.history_length <= 64u, "history_length in [1, 64]");
    require(p.delay_slots >= 3u && p.delay_slots <= 64u, "delay_slots in [3, 64] (raw slots plus the filter slot)");
    require(p.max_episode_steps >= 1u, "max_episode_steps must be positive");
    require(p.flags == 0u || p.flags == MR_TASK_PROGRAM_FIXED_ROOT, "unsupported program flags");
    require(std::isfinite(p.gait_period) && p.gait_period > 0.0f, "gait_period must be positive");
    require(p.command_seconds[0] > 0.0f && p.command_seconds[1] >= p.command_seconds[0],
            "command_seconds must be positive and ordered");
    for (int k = 0; k < 4; ++k)
        require(std::isfinite(p.command_lower[k]) && std::isfinite(p.command_upper[k]), "commands must be finite");
    const bool fixed = (p.flags & MR_TASK_PROGRAM_FIXED_ROOT) != 0u;
    if (!fixed) {
        require(h.root_type == MR_ROOT_FLOATING, "a floating-root program needs a floating-root model");
        require(p.root_q_offset + 7u <= nq && p.root_v_offset + 6u <= nv, "root offsets out of range");
    }
    for (uint32_t k = 0; k < p.action_count; ++k) {
        const MRTaskActionBindingC& b = p.actions[k];
        require(b.actuator[0] == MR_TASK_ACTUATOR_JOINT_POSITION || b.actuator[0] == MR_TASK_ACTUATOR_GRIPPER_POSITION,
                "action " + std::to_string(k) + ": only joint-position actuators are supported");
        require(b.indices[1] < nv && b.indices[2] < nq && b.indices[3] < nv,
                "action " + std::to_string(k) + ": DoF/q/v index out of range");
        require(m.dofs[h.dof_offset + b.indices[1]].q_index == b.indices[2],
                "action " + std::to_string(k) + ": q index is not the DoF's coordinate");
        require(b.parameters[1] <= b.parameters[2], "action " + std::to_string(k) + ": lower target above upper");
    }
    for (uint32_t i = 0; i < p.observation_count; ++i) {
        const MRTaskObservationOperatorC& o = p.observations[i];
        const uint32_t op = o.source[0];
        require(op <= MR_TASK_OBSERVE_ROOT_HEIGHT, "observation " + std::to_string(i) + ": unsupported opcode " +
                                                       std::to_string(op));
        if (op == MR_TASK_OBSERVE_JOINT_POSITION_ERROR || op == MR_TASK_OBSERVE_JOINT_VELOCITY ||
            op == MR_TASK_OBSERVE_PREVIOUS_ACTION)
            require(o.source[1] < p.action_count, "observation " + std::to_string(i) + ": action index out of range");
        if (op == MR_TASK_OBSERVE_ROOT_ANGULAR_VELOCITY_LOCAL || op == MR_TASK_OBSERVE_PROJECTED_GRAVITY ||
            op == MR_TASK_OBSERVE_ROOT_LINEAR_VELOCITY_LOCAL)
            require(o.source[2] < 3u, "observation " + std::to_string(i) + ": component out of range");
        if (op == MR_TASK_OBSERVE_COMMAND)
            require(o.source[2] < 7u, "observation " + std::to_string(i) + ": command lane out of range");
        require(o.auxiliary[0] == MR_INVALID_INDEX, "observation " + std::to_string(i) + ": sensor bias slots are "
                                                                                           "not supported");
    }
    for (uint32_t i = 0; i < p.reward_count; ++i)
        require(supported_reward(p.rewards[i].source[0]), "reward " + std::to_string(i) + ": unsupported opcode " +
                                                              std::to_string(p.rewards[i].source[0]));
    for (uint32_t i = 0; i < p.termination_count; ++i)
        require(p.terminations[i].source[0] <= MR_TASK_TERMINATE_MAXIMUM_TILT,
                "termination " + std::to_string(i) + ": unsupported opcode");
    for (uint32_t j = 0; j < nq; ++j) require(std::isfinite(p.default_q[j]), "default_q must be finite");

    const float dt = world.control_timestep();
    std::vector<float> fraction(p.action_count);
    for (uint32_t k = 0; k < p.action_count; ++k) fraction[k] = task_response_fraction(dt, p.actions[k].parameters[3]);
    std::vector<float> dq(m.max_nq, 0.0f);
    std::copy(p.default_q, p.default_q + nq, dq.begin());
    upload(d_actions_, p.actions, p.action_count);
    upload(d_observations_, p.observations, p.observation_count);
    upload(d_rewards_, p.rewards, p.reward_count);
    upload(d_terminations_, p.terminations, p.termination_count);
    upload(d_default_q_, dq.data(), dq.size());
    upload(d_fraction_, fraction.data(), fraction.size());

    // the program's content hash: scalar fields (pointers zeroed), then each table and the nominal pose
    {
        MRLocomotionTaskProgramC scalars = p;
        scalars.actions = nullptr;
        scalars.observations = nullptr;
        scalars.rewards = nullptr;
        scalars.terminations = nullptr;
        scalars.default_q = nullptr;
        std::vector<uint8_t> bytes;
        const auto add = [&](const void* data, size_t n) {
            const auto* b = static_cast<const uint8_t*>(data);
            bytes.insert(bytes.end(), b, b + n);
        };
        add(&scalars, sizeof(scalars));
        add(p.actions, p.action_count * sizeof(MRTaskActionBindingC));
        add(p.observations, p.observation_count * sizeof(MRTaskObservationOperatorC));
        add(p.rewards, p.reward_count * sizeof(MRTaskRewardOperatorC));
        add(p.terminations, p.termination_count * sizeof(MRTaskTerminationOperatorC));
        add(p.default_q, nq * sizeof(float));
        fingerprint_ = learning_pack_hash(bytes.data(), bytes.size());
    }

    TaskArgs& a = args_;
    a.actions = static_cast<const MRTaskActionBindingC*>(d_actions_.ptr);
    a.observations = static_cast<const MRTaskObservationOperatorC*>(d_observations_.ptr);
    a.rewards = static_cast<const MRTaskRewardOperatorC*>(d_rewards_.ptr);
    a.terminations = static_cast<const MRTaskTerminationOperatorC*>(d_terminations_.ptr);
    a.default_q = static_cast<const float*>(d_default_q_.ptr);
    a.response_fraction = static_cast<const float*>(d_fraction_.ptr);
    a.action_count = p.action_count;
    a.observation_count = p.observation_count;
    a.reward_count = p.reward_count;
    a.termination_count = p.termination_count;
    a.history_length = p.history_length;
    a.delay_slots = p.delay_slots;
    a.max_episode_steps = p.max_episode_steps;
    a.flags = p.flags;
    a.root_q_offset = p.root_q_offset;
    a.root_v_offset = p.root_v_offset;
    a.actuator_delay = p.actuator_delay;
    a.environments = E;
    a.max_nq = m.max_nq;
    a.max_nv = m.max_nv;
    a.dt = dt;
    a.base_height_target = p.base_height_target;
    a.gait_period = p.gait_period;
    for (int k = 0; k < 3; ++k) a.root_reference[k] = p.root_reference[k];
    for (int k = 0; k < 4; ++k) {
        a.command_lower[k] = p.command_lower[k];
        a.command_upper[k] = p.command_upper[k];
    }
    a.command_seconds[0] = p.command_seconds[0];
    a.command_seconds[1] = p.command_seconds[1];
    a.seed = p.seed;

    d_states_.reserve((size_t)E * sizeof(MRTaskStateC));
    d_history_.reserve((size_t)E * p.delay_slots * p.action_count * sizeof(float));
    d_previous_.reserve((size_t)E * 2u * p.action_count * sizeof(float));
    d_obs_.reserve((size_t)E * p.history_length * p.observation_count * sizeof(float));
    d_transitions_.reserve((size_t)E * sizeof(MRTaskTransitionC));
    d_targets_.reserve((size_t)E * std::max(1u, m.max_nv) * sizeof(float));
    d_mask_.reserve((size_t)E * sizeof(uint32_t));
    a.states = static_cast<MRTaskStateC*>(d_states_.ptr);
    a.action_history = static_cast<float*>(d_history_.ptr);
    a.previous_joint = static_cast<float*>(d_previous_.ptr);
    a.observations_out = static_cast<float*>(d_obs_.ptr);
    a.transitions = static_cast<MRTaskTransitionC*>(d_transitions_.ptr);
    a.targets = static_cast<float*>(d_targets_.ptr);
    a.reset_mask = static_cast<uint32_t*>(d_mask_.ptr);
    refresh_world_pointers();

    // a finished episode restarts from the nominal pose at rest (LocomotionTask.metal:2350-2380)
    std::vector<float> rq((size_t)E * m.max_nq), rv((size_t)E * m.max_nv, 0.0f);
    for (uint32_t e = 0; e < E; ++e) std::copy(dq.begin(), dq.end(), rq.begin() + (size_t)e * m.max_nq);
    std::vector<MRBodyStateC> rs = m.initial_scene;
    rs.resize((size_t)E * m.max_scene_bodies);
    world.set_reset_state(rq.data(), rv.data(), rs.data());
    reset();
}

void LocomotionTask::refresh_world_pointers() {
    const World::DeviceView d = world_.device_view();
    args_.q = d.q;
    args_.v = d.v;
    args_.reset_q = d.reset_q;
    args_.reset_v = d.reset_v;
    const ModelData& m = world_.model();
    args_.dofs = d.dofs + m.models[m.environment_model[0]].dof_offset;
}

void LocomotionTask::reset() {
    hipStream_t s = world_.stream();
    const uint32_t E = args_.environments;
    hip_check(hipStreamSynchronize(s), "sync");
    hip_check(hipMemset(d_states_.ptr, 0, (size_t)E * sizeof(MRTaskStateC)), "clear task states");
    hip_check(hipMemset(d_history_.ptr, 0, (size_t)E * args_.delay_slots * args_.action_count * sizeof(float)),
              "clear action history");
    hip_check(hipMemset(d_previous_.ptr, 0, (size_t)E * 2u * args_.action_count * sizeof(float)), "clear joints");
    hip_check(hipMemset(d_transitions_.ptr, 0, (size_t)E * sizeof(MRTaskTransitionC)), "clear transitions");
    // every state is uninitialized: observe starts episode 1 everywhere and raises every reset mask
    hip_check(launch_task_observe(args_, s), "task observe");
}

LocomotionTask::~LocomotionTask() = default;

void LocomotionTask::set_policy(const PolicyPack& pack) {
    auto compiled = std::make_unique<Policy>(pack, args_.environments, observation_floats(), args_.action_count);
    const size_t n = (size_t)args_.environments * args_.action_count * sizeof(float);
    d_policy_actions_.reserve(n);
    d_policy_latents_.reserve(n);
    d_policy_log_probabilities_.reserve((size_t)args_.environments * sizeof(float));
    policy_ = std::move(compiled);
    args_.policy_revision = policy_->revision();
}

void LocomotionTask::clear_policy() {
    policy_.reset();
    args_.policy_revision = 0;
}

uint64_t LocomotionTask::policy_revision() const { return policy_ ? policy_->revision() : 0u; }

const float* LocomotionTask::device_policy_actions() const {
    return policy_ ? static_cast<const float*>(d_policy_actions_.ptr) : nullptr;
}

const float* LocomotionTask::device_policy_latents() const {
    return policy_ ? static_cast<const float*>(d_policy_latents_.ptr) : nullptr;
}

const float* LocomotionTask::device_policy_log_probabilities() const {
    return policy_ ? static_cast<const float*>(d_policy_log_probabilities_.ptr) : nullptr;
}

void LocomotionTask::read_policy_outputs(float* actions, size_t action_count, float* latents, size_t latent_count,
                                         float* log_probabilities, size_t log_probability_count) const {
    if (!policy_) throw std::runtime_error("locomotion task: no policy is installed");
    const size_t E = args_.environments, EA = E * args_.action_count;
    const auto copy = [&](float* out, size_t count, const DeviceBuffer& b, size_t need, const char* what) {
        if (out == nullptr) return;
        if (count != need)
            throw std::invalid_argument(std::string("locomotion task: ") + what + " count must be " +
                                        std::to_string(need));
        hip_check(hipMemcpy(out, b.ptr, need * sizeof(float), hipMemcpyDeviceToHost), what);
    };
    hip_check(hipStreamSynchronize(world_.stream()), "sync");
    copy(actions, action_count, d_policy_actions_, EA, "policy actions");
    copy(latents, latent_count, d_policy_latents_, EA, "policy latents");
    copy(log_probabilities, log_probability_count, d_policy_log_probabilities_, E, "policy log probabilities");
}

void LocomotionTask::step(const float* device_actions) {
    if (world_.queue_grid() != 0u)
        throw std::runtime_error("locomotion task: the world must step one workgroup per environment");
    hipStream_t s = world_.stream();
    if (policy_) {
        if (device_actions != nullptr)
            throw std::runtime_error("locomotion task: a policy is installed, so the step takes no actions");
        auto* actions = static_cast<float*>(d_policy_actions_.ptr);
        policy_->act(args_.observations_out, nullptr, args_.states, 0u, args_.seed, actions,
                     static_cast<float*>(d_policy_latents_.ptr),
                     static_cast<float*>(d_policy_log_probabilities_.ptr), nullptr, s);
        device_actions = actions;
    } else if (device_actions == nullptr) {
        throw std::runtime_error("locomotion task: actions are required");
    }
    hip_check(launch_task_apply_actions(args_, device_actions, s), "task apply_actions");
    world_.enqueue_chained(args_.targets, args_.reset_mask, 1u);
    hip_check(launch_task_complete(args_, world_.last_status_device(), s), "task complete");
    hip_check(launch_task_observe(args_, s), "task observe");
}

void LocomotionTask::read(float* observations, MRTaskTransitionC* transitions, MRTaskStateC* states) const {
    const uint32_t E = args_.environments;
    hip_check(hipStreamSynchronize(world_.stream()), "sync");
    if (observations != nullptr)
        hip_check(hipMemcpy(observations, args_.observations_out, (size_t)E * observation_floats() * sizeof(float),
                            hipMemcpyDeviceToHost),
                  "read observations");
    if (transitions != nullptr)
        hip_check(hipMemcpy(transitions, args_.transitions, (size_t)E * sizeof(MRTaskTransitionC),
                            hipMemcpyDeviceToHost),
                  "read transitions");
    if (states != nullptr)
        hip_check(hipMemcpy(states, args_.states, (size_t)E * sizeof(MRTaskStateC), hipMemcpyDeviceToHost),
                  "read task states");
}

static_assert(sizeof(MRTaskActionBindingC) == 64, "MRTaskActionBindingGPU layout");
static_assert(sizeof(MRTaskObservationOperatorC) == 48, "MRTaskObservationOperatorGPU layout");
static_assert(sizeof(MRTaskRewardOperatorC) == 48, "MRTaskRewardOperatorGPU layout");
static_assert(sizeof(MRTaskTerminationOperatorC) == 48, "MRTaskTerminationOperatorGPU layout");
static_assert(sizeof(MRTaskStateC) == 176, "MRTaskStateGPU layout");
static_assert(sizeof(MRTaskTransitionC) == 128, "MRTaskTransitionGPU layout");

} // namespace mr

static_assert(sizeof(MRLocomotionTaskProgramC) == 160 && offsetof(MRLocomotionTaskProgramC, seed) == 152,
              "MRLocomotionTaskProgramC layout (mirrored by the ctypes binding)");

namespace mr {

void LocomotionTask::write_policy_rollout_pack(const MRPolicyRolloutBatchC& batch, const std::string& id,
                                               const std::string& path) const {
    if (!policy_) throw std::invalid_argument("a compiled policy must be installed before publishing a rollout pack");
    if (batch.transition_count != 0 && batch.transitions == nullptr)
        throw std::invalid_argument("rollout transition pointer is null");
    // MRTaskTransitionC -> MRLearningTransitionC as c_api.cpp:4420-4447 maps them
    std::vector<MRLearningTransitionC> transitions(batch.transition_count);
    for (size_t i = 0; i < batch.transition_count; ++i) {
        const MRTaskTransitionC& s = batch.transitions[i];
        MRLearningTransitionC& d = transitions[i];
        d = MRLearningTransitionC{};
        d.reward_and_bootstrap[0] = s.reward_and_state[0];
        d.reward_and_bootstrap[1] = s.timeout_bootstrap_value;
        for (int k = 0; k < 4; ++k) {
            d.termination[k] = s.termination[k];
            d.context[k] = s.task_progress[k];
        }
        d.policy_revision = s.policy_revision;
    }
    MRPolicyRolloutPackC k{};
    k.id = id.c_str();
    k.task_fingerprint = fingerprint_;
    k.policy_fingerprint = policy_->fingerprint();
    k.policy_revision = policy_->revision();
    k.environment_count = args_.environments;
    k.control_step_count = batch.control_step_count;
    k.actor_observation_count = observation_floats();
    // without a critic network the critic reads the actor observation
    k.critic_observation_count = policy_->critic_input() != 0 ? policy_->critic_input() : observation_floats();
    k.action_count = args_.action_count;
    k.motion_feature_count = 0;
    k.actor_observations = batch.actor_observations;
    k.actor_observation_count_total = batch.actor_observation_count;
    k.critic_observations = batch.critic_observations;
    k.critic_observation_count_total = batch.critic_observation_count;
    k.motion_features = batch.motion_features;
    k.motion_feature_count_total = batch.motion_feature_count;
    k.teacher_actions = batch.teacher_actions;
    k.teacher_action_count = batch.teacher_action_count;
    k.latents = batch.latents;
    k.latent_count = batch.latent_count;
    k.log_probabilities = batch.log_probabilities;
    k.log_probability_count = batch.log_probability_count;
    k.values = batch.values;
    k.value_count = batch.value_count;
    k.bootstrap_values = batch.bootstrap_values;
    k.bootstrap_value_count = batch.bootstrap_value_count;
    k.transitions = transitions.data();
    k.transition_count = transitions.size();
    mr::write_policy_rollout_pack(k, path);
}

} // namespace mr
