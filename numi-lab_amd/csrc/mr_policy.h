// A policy pack compiled into device memory: the tables of the actor and
// critic networks in one arena, two ping-pong scratch buffers and the actor
// mean, sized for a fixed environment count.
#pragma once

#include "mr_packs.h"
#include "mr_world.h"

#include <hip/hip_runtime.h>

namespace mr {

class Policy {
public:
    // Compiles the pack (compilePolicyProgram's shape rules,
    // PolicyProgram.cpp:225-420): layer widths chain from the actor input to
    // the action count (critic: to 1), final layers are identity, the
    // normalisation tables match the input width or are empty (mean 0,
    // inverse standard deviation 1, both positive), a stochastic policy has
    // one log standard deviation in [-5, 2] per action, bias and scale match
    // the action count or are empty (0 and 1). actor_input / actions of 0
    // take the widths from the pack. Throws std::invalid_argument.
    Policy(const PolicyPack& pack, uint32_t environments, uint32_t actor_input = 0, uint32_t actions = 0);
    Policy(const Policy&) = delete;
    Policy& operator=(const Policy&) = delete;

    uint64_t revision() const { return revision_; }
    uint64_t fingerprint() const { return fingerprint_; }
    uint32_t environments() const { return environments_; }
    uint32_t actor_input() const { return actor_input_; }
    uint32_t critic_input() const { return critic_input_; }
    uint32_t actions() const { return actions_; }

    // Actor network -> critic network (when critic_observations and values
    // are given and the pack has one) -> sample and score. All device
    // pointers; enqueued on stream.
    void act(const float* actor_observations, const float* critic_observations, const MRTaskStateC* states,
             uint32_t control_step, uint64_t seed, float* actions, float* latents, float* log_probabilities,
             float* values, hipStream_t stream);

private:
    struct Layer {
        size_t weights, bias;  // float offsets in the arena
        uint32_t input, output, activation;
    };
    void network(const std::vector<Layer>& layers, size_t mean, size_t inverse_std, const float* input,
                 float* output, uint32_t output_stride, hipStream_t stream);
    uint64_t revision_ = 0, fingerprint_ = 0;
    uint32_t environments_ = 0, actor_input_ = 0, critic_input_ = 0, actions_ = 0, widest_ = 0;
    float observation_clip_ = 0.0f, action_clip_ = 0.0f;
    bool stochastic_ = false;
    std::vector<Layer> actor_, critic_;
    size_t mean_ = 0, inverse_std_ = 0, critic_mean_ = 0, critic_inverse_std_ = 0;
    size_t bias_ = 0, scale_ = 0, log_std_ = 0;
    DeviceBuffer arena_, scratch_a_, scratch_b_, actor_mean_;
    const float* table(size_t offset) const { return static_cast<const float*>(arena_.ptr) + offset; }
};

}  // namespace mr
