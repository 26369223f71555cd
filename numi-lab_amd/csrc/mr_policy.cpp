// Device policy: compile a policy pack into one arena of float tables and run
// inference as the reference encodes it (MetalWorld.mm:9425-9640): one dense
// launch per layer, ping-ponging between two scratch buffers, the first layer
// normalising and clipping its input, then the sample-and-score launch.
#include "mr_policy.h"
#include "mr_task_host.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace mr {
namespace {

void reject(const std::string& what) { throw std::invalid_argument("policy: " + what); }

}  // namespace

Policy::Policy(const PolicyPack& pack, uint32_t environments, uint32_t actor_input, uint32_t actions)
    : revision_(pack.revision), environments_(environments) {
    if (environments == 0) reject("environment count must be positive");
    if (pack.id.empty() || pack.revision == 0 || pack.layers.empty()) reject("identity, revision and layers are required");
    if (!std::isfinite(pack.observation_clip) || !(pack.observation_clip > 0.0f) || !std::isfinite(pack.action_clip) ||
        !(pack.action_clip > 0.0f))
        reject("observation and action clips must be finite and positive");
    actor_input_ = actor_input != 0 ? actor_input : pack.layers.front().input_count;
    actions_ = actions != 0 ? actions : pack.layers.back().output_count;
    critic_input_ = pack.critic_layers.empty() ? 0u : pack.critic_layers.front().input_count;
    stochastic_ = !pack.action_log_std.empty();
    if (stochastic_ && pack.critic_layers.empty()) reject("a stochastic policy requires a critic");
    observation_clip_ = pack.observation_clip;
    action_clip_ = pack.action_clip;

    std::vector<float> arena;
    const auto append = [&](const std::vector<float>& v) {
        const size_t at = arena.size();
        arena.insert(arena.end(), v.begin(), v.end());
        return at;
    };
    const auto compile = [&](const std::vector<DenseLayer>& layers, uint32_t input, uint32_t output, const char* name,
                             std::vector<Layer>& out) {
        uint32_t expected = input;
        for (size_t i = 0; i < layers.size(); ++i) {
            const DenseLayer& l = layers[i];
            const std::string where = std::string(name) + " layer " + std::to_string(i);
            if (l.input_count != expected || l.output_count == 0 ||
                (uint64_t)l.input_count * l.output_count != l.weights.size() || l.bias.size() != l.output_count ||
                l.activation > MR_POLICY_ACTIVATION_SILU)
                reject(where + ": shape or activation is invalid");
            for (float x : l.weights)
                if (!std::isfinite(x)) reject(where + ": non-finite weight");
            for (float x : l.bias)
                if (!std::isfinite(x)) reject(where + ": non-finite bias");
            const bool last = i + 1 == layers.size();
            if (last && (l.output_count != output || l.activation != MR_POLICY_ACTIVATION_IDENTITY))
                reject(where + ": the final layer must be identity with " + std::to_string(output) + " outputs");
            if (!last) widest_ = std::max(widest_, l.output_count);
            out.push_back(Layer{append(l.weights), append(l.bias), l.input_count, l.output_count, l.activation});
            expected = l.output_count;
        }
    };
    const auto normalisation = [&](const std::vector<float>& mean, const std::vector<float>& inverse_std, uint32_t width,
                                   const char* name, size_t& mean_at, size_t& inverse_at) {
        if ((!mean.empty() && mean.size() != width) || (!inverse_std.empty() && inverse_std.size() != width))
            reject(std::string(name) + " normalisation does not match the input width " + std::to_string(width));
        std::vector<float> m = mean.empty() ? std::vector<float>(width, 0.0f) : mean;
        std::vector<float> s = inverse_std.empty() ? std::vector<float>(width, 1.0f) : inverse_std;
        for (float x : m)
            if (!std::isfinite(x)) reject(std::string(name) + " mean is not finite");
        for (float x : s)
            if (!std::isfinite(x) || !(x > 0.0f))
                reject(std::string(name) + " inverse standard deviations must be finite and positive");
        mean_at = append(m);
        inverse_at = append(s);
    };
    compile(pack.layers, actor_input_, actions_, "actor", actor_);
    normalisation(pack.observation_mean, pack.observation_inverse_std, actor_input_, "actor observation", mean_,
                  inverse_std_);
    if (!pack.critic_layers.empty()) {
        compile(pack.critic_layers, critic_input_, 1u, "critic", critic_);
        normalisation(pack.critic_observation_mean, pack.critic_observation_inverse_std, critic_input_,
                      "critic observation", critic_mean_, critic_inverse_std_);
    } else if (!pack.critic_observation_mean.empty() || !pack.critic_observation_inverse_std.empty()) {
        reject("critic normalisation without a critic network");
    }
    if (stochastic_) {
        if (pack.action_log_std.size() != actions_) reject("one log standard deviation per action is required");
        for (float x : pack.action_log_std)
            if (!(x >= -5.0f && x <= 2.0f)) reject("Gaussian log standard deviations must be within [-5, 2]");
        log_std_ = append(pack.action_log_std);
    }
    if ((!pack.action_bias.empty() && pack.action_bias.size() != actions_) ||
        (!pack.action_scale.empty() && pack.action_scale.size() != actions_))
        reject("action bias and scale must match the action count");
    for (float x : pack.action_bias)
        if (!std::isfinite(x)) reject("action bias is not finite");
    for (float x : pack.action_scale)
        if (!std::isfinite(x)) reject("action scale is not finite");
    bias_ = append(pack.action_bias.empty() ? std::vector<float>(actions_, 0.0f) : pack.action_bias);
    scale_ = append(pack.action_scale.empty() ? std::vector<float>(actions_, 1.0f) : pack.action_scale);

    const std::vector<uint8_t> payload = serialize_policy(pack);
    fingerprint_ = learning_pack_hash(payload.data(), payload.size());

    arena_.reserve(arena.size() * sizeof(float));
    hip_check(hipMemcpy(arena_.ptr, arena.data(), arena.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload policy tables");
    const size_t hidden = (size_t)environments * std::max(widest_, 1u) * sizeof(float);
    scratch_a_.reserve(hidden);
    scratch_b_.reserve(hidden);
    actor_mean_.reserve((size_t)environments * actions_ * sizeof(float));
}

void Policy::network(const std::vector<Layer>& layers, size_t mean, size_t inverse_std, const float* input,
                     float* output, uint32_t output_stride, hipStream_t stream) {
    float* scratch[2] = {static_cast<float*>(scratch_a_.ptr), static_cast<float*>(scratch_b_.ptr)};
    for (size_t i = 0; i < layers.size(); ++i) {
        const Layer& l = layers[i];
        const bool first = i == 0, last = i + 1 == layers.size();
        const float* src = first ? input : scratch[(i - 1) & 1u];
        float* dst = last ? output : scratch[i & 1u];
        hip_check(launch_policy_dense(table(l.weights), table(l.bias), first ? table(mean) : nullptr,
                                      first ? table(inverse_std) : nullptr, observation_clip_, src, l.input,
                                      dst, last ? output_stride : l.output, environments_, l.input, l.output,
                                      l.activation, stream),
                  "policy dense layer");
    }
}

void Policy::act(const float* actor_observations, const float* critic_observations, const MRTaskStateC* states,
                 uint32_t control_step, uint64_t seed, float* actions, float* latents, float* log_probabilities,
                 float* values, hipStream_t stream) {
    if (actor_observations == nullptr || states == nullptr || actions == nullptr || latents == nullptr ||
        log_probabilities == nullptr)
        reject("actor observations, task states and action outputs are required");
    if ((critic_observations == nullptr) != (values == nullptr))
        reject("critic observations and values go together");
    float* mean = static_cast<float*>(actor_mean_.ptr);
    network(actor_, mean_, inverse_std_, actor_observations, mean, actions_, stream);
    if (critic_observations != nullptr && !critic_.empty())
        network(critic_, critic_mean_, critic_inverse_std_, critic_observations, values, 1u, stream);
    hip_check(launch_policy_sample(mean, actions_, table(bias_), table(scale_), stochastic_ ? table(log_std_) : nullptr,
                                   action_clip_, seed, revision_, states, environments_, actions_, control_step,
                                   actions, latents, log_probabilities, stream),
              "policy sample and score");
}

}  // namespace mr
