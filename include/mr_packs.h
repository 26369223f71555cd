/* On-disk learning packs and compiled policies (C ABI, plain pointers and
 * sizes, no torch types).
 *
 * A learning pack is the reference's deterministic little-endian artifact
 * (src/core/LearningPacks.cpp:29-61, 2074-2295): a 32-byte header (magic
 * "MRLEARN\0", wire version, kind, payload bytes, content hash) followed by
 * the payload. The content hash is 64-bit FNV-1a over the payload bytes with
 * 0 mapped to 1 (:920-930). Payload scalars are raw little-endian values;
 * strings and float tables are a u64 count followed by the elements; a table
 * of records is a u64 count followed by each record (:933-1125).
 *
 * Two kinds are read and written here, byte-compatible with the reference:
 *   PolicyPack        kind 2, wire version 4 (3 is read too): the actor and
 *                     critic dense networks, observation normalisation,
 *                     action distribution and the semantic contract
 *                     (serializePolicy, :1709-1816);
 *   PolicyRolloutPack kind 3, wire version 7: one batch of rollout samples
 *                     (serializePolicyRollout, :1818-1904).
 * The records below are laid out as include/metalrobo/c_api.h:334-394 and
 * task_program_types.h:628-637. */
#ifndef MR_PACKS_H
#define MR_PACKS_H

#include <stddef.h>
#include <stdint.h>

#include "mr_task.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MR_LEARNING_PACK_HEADER_BYTES 32u
#define MR_LEARNING_PACK_KIND_POLICY 2u
#define MR_LEARNING_PACK_KIND_POLICY_ROLLOUT 3u
#define MR_POLICY_PACK_FORMAT_VERSION 4u
#define MR_POLICY_ROLLOUT_PACK_FORMAT_VERSION 7u

/* c_api.h:334-342 */
typedef struct MRPolicyDenseLayerC {
    uint32_t input_count;
    uint32_t output_count;
    uint32_t activation; /* MR_POLICY_ACTIVATION_* */
    const float* weights; /* [output_count][input_count] */
    size_t weight_count;
    const float* bias;    /* [output_count] */
    size_t bias_count;
} MRPolicyDenseLayerC;

/* c_api.h:344-372 */
typedef struct MRPolicyPackC {
    const char* id;
    uint64_t revision;
    uint64_t contract_version;
    uint64_t world_fingerprint;
    uint64_t task_fingerprint;
    uint64_t observation_fingerprint;
    uint64_t action_fingerprint;
    const float* observation_mean;
    size_t observation_mean_count;
    const float* observation_inverse_standard_deviation;
    size_t observation_inverse_standard_deviation_count;
    const MRPolicyDenseLayerC* layers;
    size_t layer_count;
    const float* critic_observation_mean;
    size_t critic_observation_mean_count;
    const float* critic_observation_inverse_standard_deviation;
    size_t critic_observation_inverse_standard_deviation_count;
    const MRPolicyDenseLayerC* critic_layers;
    size_t critic_layer_count;
    const float* action_log_standard_deviation;
    size_t action_log_standard_deviation_count;
    const float* action_bias;
    size_t action_bias_count;
    const float* action_scale;
    size_t action_scale_count;
    float observation_clip;
    float action_clip;
} MRPolicyPackC;

/* c_api.h:374-394 */
typedef struct MRPolicyRolloutBatchC {
    uint32_t control_step_count;
    const float* actor_observations;
    size_t actor_observation_count;
    const float* critic_observations;
    size_t critic_observation_count;
    const float* motion_features;
    size_t motion_feature_count;
    const float* teacher_actions;
    size_t teacher_action_count;
    const float* latents;
    size_t latent_count;
    const float* log_probabilities;
    size_t log_probability_count;
    const float* values;
    size_t value_count;
    const float* bootstrap_values;
    size_t bootstrap_value_count;
    const MRTaskTransitionC* transitions;
    size_t transition_count;
} MRPolicyRolloutBatchC;

/* MRLearningTransitionGPU (task_program_types.h:628-637): the per-sample
 * record of a rollout pack. */
typedef struct MRLearningTransitionC {
    float reward_and_bootstrap[4]; /* reward, timeout bootstrap value, 0, 0 */
    uint32_t termination[4];       /* done, timeout, physics error, termination reason */
    uint32_t context[4];           /* difficulty band, terrain level, impact sequence, impact flags */
    uint64_t policy_revision;
    uint64_t reserved;
} MRLearningTransitionC;

/* One outcome channel of a rollout pack (PolicyOutcomeDescriptor,
 * LearningPacks.hpp:37-41). */
typedef struct MRPolicyOutcomeDescriptorC {
    const char* id;
    const char* unit;
    uint32_t direction;
} MRPolicyOutcomeDescriptorC;

/* A whole rollout pack (PolicyRolloutPack, LearningPacks.hpp:56-82): sample
 * tables are step-major [control_step][environment][...]. No counterpart in
 * the reference's C API, which publishes rollout packs only from a task
 * handle; mr_locomotion_task_write_policy_rollout_pack is that form. */
typedef struct MRPolicyRolloutPackC {
    const char* id;
    uint64_t task_fingerprint;
    uint64_t policy_fingerprint;
    uint64_t policy_revision;
    uint32_t environment_count;
    uint32_t control_step_count;
    uint32_t actor_observation_count;
    uint32_t critic_observation_count;
    uint32_t action_count;
    uint32_t motion_feature_count;
    const float* actor_observations;
    size_t actor_observation_count_total;
    const float* critic_observations;
    size_t critic_observation_count_total;
    const float* motion_features;
    size_t motion_feature_count_total;
    const float* teacher_actions;
    size_t teacher_action_count;
    const float* latents;
    size_t latent_count;
    const float* log_probabilities;
    size_t log_probability_count;
    const float* values;
    size_t value_count;
    const float* bootstrap_values;
    size_t bootstrap_value_count;
    const MRPolicyOutcomeDescriptorC* outcomes;
    size_t outcome_count;
    const float* outcome_values; /* [sample][outcome] */
    size_t outcome_value_count;
    const MRLearningTransitionC* transitions;
    size_t transition_count;
} MRPolicyRolloutPackC;

/* mr_learning_pack_content_hash (c_api.h:568-573): the wire hash of one
 * payload. Returns 0 only for a null payload with a non-zero byte count. */
uint64_t mr_learning_pack_content_hash(const void* payload, size_t byte_count);

/* mr_write_policy_pack (c_api.h:562-567): validates the policy as
 * validatePolicyArtifact does (LearningPacks.cpp:399-502) and publishes it at
 * the path through a sibling temporary file and one rename. 0 or -1. */
int mr_write_policy_pack(const MRPolicyPackC* policy, const char* policy_pack_path);

/* readPolicyPack (LearningPacks.cpp:2458-2470): reads and validates a policy
 * pack of wire version 3 or 4 (header, kind, length, content hash, every
 * table exactly consumed). The handle owns the tables; its view stays valid
 * until destroy. No C counterpart in the reference (its C API reads packs
 * only into a task handle, mr_task_rollout_load_policy_pack). NULL with
 * mr_last_error set on failure. */
typedef struct MRPolicyPackHandle MRPolicyPackHandle;
MRPolicyPackHandle* mr_read_policy_pack(const char* policy_pack_path);
const MRPolicyPackC* mr_policy_pack_view(const MRPolicyPackHandle* pack);
uint64_t mr_policy_pack_content_hash(const MRPolicyPackHandle* pack);
uint32_t mr_policy_pack_format_version(const MRPolicyPackHandle* pack);
void mr_policy_pack_destroy(MRPolicyPackHandle* pack);

/* writePolicyRolloutPack (LearningPacks.cpp:2472-2533): validates the sample
 * counts against the dimensions (validatePolicyRolloutArtifact, :726-918) and
 * publishes the pack. 0 or -1. */
int mr_write_policy_rollout_pack(const MRPolicyRolloutPackC* pack, const char* output_path);

/* A policy compiled into device memory for a fixed environment count: actor
 * and critic layers, normalisation and the action distribution. Inference
 * (encodePolicyInference, MetalWorld.mm:9425-9640) runs the actor network over
 * the actor observations (the first layer normalises and clips its input),
 * the critic network over the critic observations when both exist, then
 * samples and scores the actions. Layers ping-pong between two scratch
 * buffers of [E][widest layer] floats. */
typedef struct MRPolicyHandle MRPolicyHandle;
MRPolicyHandle* mr_policy_create(const MRPolicyPackC* policy, uint32_t environments);
void mr_policy_destroy(MRPolicyHandle* policy);
uint64_t mr_policy_revision(const MRPolicyHandle* policy);
uint32_t mr_policy_actor_input_count(const MRPolicyHandle* policy);
uint32_t mr_policy_critic_input_count(const MRPolicyHandle* policy); /* 0 without a critic */
uint32_t mr_policy_action_count(const MRPolicyHandle* policy);
/* Device pointers; critic_observations and values may be null (no critic
 * pass). task_states [E] key the sampling noise with seed and the policy's
 * revision; without a log standard deviation the action is the mean. Outputs
 * actions and latents [E][actions], log probabilities [E], values [E].
 * Enqueued on stream (a hipStream_t or null). */
int mr_policy_act(MRPolicyHandle* policy, const float* actor_observations, const float* critic_observations,
                  const MRTaskStateC* task_states, uint32_t control_step, uint64_t seed, float* actions,
                  float* latents, float* log_probabilities, float* values, void* stream);

/* The reference's task-rollout policy entry points, for the locomotion task
 * (c_api.h:598-617). With a policy installed, mr_locomotion_task_step takes no
 * action stream (pass NULL): inference runs between the previous step's
 * observation and action application. The actor's input width must equal the
 * task's observation floats (history x observation_count) and its output the
 * action count; the task has no critic observations, so the critic is not
 * evaluated. Sampling uses the task program's seed and control step 0 of a
 * one-step submission. A failed install leaves the installed policy. */
int mr_locomotion_task_set_policy(MRLocomotionTaskHandle* task, const MRPolicyPackC* policy);
int mr_locomotion_task_load_policy_pack(MRLocomotionTaskHandle* task, const char* policy_pack_path);
uint64_t mr_locomotion_task_policy_revision(const MRLocomotionTaskHandle* task);
int mr_locomotion_task_clear_policy(MRLocomotionTaskHandle* task);
/* Device pointers of the last policy step: actions, latents [E][A], log
 * probabilities [E]; null without a policy. */
const float* mr_locomotion_task_device_policy_actions(const MRLocomotionTaskHandle* task);
const float* mr_locomotion_task_device_policy_latents(const MRLocomotionTaskHandle* task);
const float* mr_locomotion_task_device_policy_log_probabilities(const MRLocomotionTaskHandle* task);
/* Waits, then copies the same three tables to host memory (any may be null). */
int mr_locomotion_task_read_policy_outputs(MRLocomotionTaskHandle* task, float* actions, size_t action_count,
                                           float* latents, size_t latent_count, float* log_probabilities,
                                           size_t log_probability_count);
/* mr_task_rollout_write_policy_rollout_pack (c_api.h:747-755): publishes one
 * aggregated batch with the dimensions of the task and the installed policy.
 * Each MRTaskTransitionC becomes an MRLearningTransitionC as the reference
 * maps it (c_api.cpp:4420-4447). The fingerprints are this library's: the task
 * program's content hash over its records and the policy's pack content
 * hash. The task defines no outcome channels. Requires an installed policy. */
int mr_locomotion_task_write_policy_rollout_pack(const MRLocomotionTaskHandle* task,
                                                 const MRPolicyRolloutBatchC* batch, const char* batch_id,
                                                 const char* output_path);

#ifdef __cplusplus
}
#endif

#endif /* MR_PACKS_H */
