// Host executor of the locomotion task layer around one World.
#pragma once

#include "mr_packs.h"
#include "mr_task_args.h"
#include "mr_world.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

namespace mr {

class Policy;

class LocomotionTask {
public:
    // Validates the program against the world's model (single tree, indices in
    // range, supported opcodes), copies its tables to the device and sets the
    // world's reset state to the program's nominal pose at rest.
    LocomotionTask(const MRLocomotionTaskProgramC& program, World& world);
    ~LocomotionTask();
    LocomotionTask(const LocomotionTask&) = delete;
    LocomotionTask& operator=(const LocomotionTask&) = delete;

    void reset();
    // With a policy installed, device_actions is null and the policy acts on
    // the current observations first.
    void step(const float* device_actions);
    void read(float* observations, MRTaskTransitionC* transitions, MRTaskStateC* states) const;

    uint32_t environments() const { return args_.environments; }
    uint32_t observation_floats() const { return args_.history_length * args_.observation_count; }
    const float* device_observations() const { return args_.observations_out; }
    const MRTaskTransitionC* device_transitions() const { return args_.transitions; }
    const float* device_targets() const { return args_.targets; }

    // Compiles and installs a policy whose actor reads this task's observation
    // floats and writes its action count; a failed install keeps the old one.
    void set_policy(const PolicyPack& pack);
    void clear_policy();
    uint64_t policy_revision() const;
    const Policy* policy() const { return policy_.get(); }
    const float* device_policy_actions() const;
    const float* device_policy_latents() const;
    const float* device_policy_log_probabilities() const;
    void read_policy_outputs(float* actions, size_t action_count, float* latents, size_t latent_count,
                             float* log_probabilities, size_t log_probability_count) const;
    // FNV-1a content hash over the program's scalar fields and tables.
    uint64_t fingerprint() const { return fingerprint_; }
    // One rollout pack of the batch with this task's dimensions and the
    // installed policy's revision and fingerprint.
    void write_policy_rollout_pack(const MRPolicyRolloutBatchC& batch, const std::string& id,
                                   const std::string& path) const;

private:
    World& world_;
    TaskArgs args_{};
    DeviceBuffer d_actions_, d_observations_, d_rewards_, d_terminations_, d_default_q_, d_fraction_;
    DeviceBuffer d_states_, d_history_, d_previous_, d_obs_, d_transitions_, d_targets_, d_mask_;
    std::unique_ptr<Policy> policy_;
    DeviceBuffer d_policy_actions_, d_policy_latents_, d_policy_log_probabilities_;
    uint64_t fingerprint_ = 0;
    void refresh_world_pointers();
};

hipError_t launch_task_apply_actions(const TaskArgs& a, const float* actions, hipStream_t stream);
hipError_t launch_task_complete(const TaskArgs& a, const MRWorldStatusC* status, hipStream_t stream);
hipError_t launch_task_observe(const TaskArgs& a, hipStream_t stream);
hipError_t launch_policy_dense(const float* weights, const float* bias, const float* mean, const float* inverse_std,
                               float clip, const float* input, uint32_t input_stride, float* output,
                               uint32_t output_stride, uint32_t environments, uint32_t in_width, uint32_t out_width,
                               uint32_t act, hipStream_t stream);
hipError_t launch_policy_sample(const float* mean, uint32_t mean_stride, const float* action_bias,
                                const float* action_scale, const float* log_std, float action_clip,
                                unsigned long long seed, unsigned long long revision, const MRTaskStateC* states,
                                uint32_t environments, uint32_t actions, uint32_t control_step, float* actions_out,
                                float* latents_out, float* log_prob_out, hipStream_t stream);

// The response fraction of a binding: clamp(1 - exp(-dt / response), 0, 1),
// 1 without a response time (LocomotionTask.metal:3566-3576), evaluated once on
// the host in double precision and rounded to float.
float task_response_fraction(float dt, float response_seconds);

} // namespace mr
