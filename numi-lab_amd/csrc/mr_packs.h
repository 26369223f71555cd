// Learning-pack wire format (policy and rollout packs) and the host form of a
// policy pack.
#pragma once

#include "../../include/mr_packs.h"

#include <cstdint>
#include <string>
#include <vector>

namespace mr {

struct DenseLayer {
    uint32_t input_count = 0, output_count = 0, activation = 0;
    std::vector<float> weights;  // [output_count][input_count]
    std::vector<float> bias;     // [output_count]
};

// PolicyPack (PolicyProgram.hpp:39-76) as read from or written to a pack.
struct PolicyPack {
    std::string id;
    uint64_t revision = 0;
    uint64_t contract[5] = {0, 0, 0, 0, 0};  // version, world, task, observation, action fingerprints
    std::vector<float> observation_mean, observation_inverse_std;
    std::vector<DenseLayer> layers;
    std::vector<float> critic_observation_mean, critic_observation_inverse_std;
    std::vector<DenseLayer> critic_layers;
    std::vector<float> action_log_std, action_bias, action_scale;
    float observation_clip = 100.0f;
    float action_clip = 3.4028234663852886e38f;
    // set by read_policy_pack / serialisation
    uint32_t format_version = MR_POLICY_PACK_FORMAT_VERSION;
    uint64_t content_hash = 0;
};

// 64-bit FNV-1a over the bytes, 0 mapped to 1 (LearningPacks.cpp:920-930).
uint64_t learning_pack_hash(const void* bytes, size_t count);

// Copies a C policy record (policyPackFromC, c_api.cpp:1752-1840); throws on
// null tables with non-zero counts.
PolicyPack policy_pack_from_c(const MRPolicyPackC& c);
// validatePolicyArtifact (LearningPacks.cpp:399-502); throws std::invalid_argument.
void validate_policy_artifact(const PolicyPack& p);
// serializePolicy (LearningPacks.cpp:1709-1756), wire version 4.
std::vector<uint8_t> serialize_policy(const PolicyPack& p);
// Validates and publishes; returns the content hash.
uint64_t write_policy_pack(const PolicyPack& p, const std::string& path);
// readPolicyPack (LearningPacks.cpp:2458-2470): versions 3 and 4.
PolicyPack read_policy_pack(const std::string& path);

// validatePolicyRolloutArtifact + serializePolicyRollout + writePack; returns the content hash.
uint64_t write_policy_rollout_pack(const MRPolicyRolloutPackC& pack, const std::string& path);

}  // namespace mr
