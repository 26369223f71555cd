/*
 * Flat C ABI of the MI355X-native batched temporalCone stepper.
 *
 * Drop-in boundary for the stepping subset of the reference engine ABI
 * (include/metalrobo/c_api.h, implemented in src/c_api.cpp). Conventions
 * follow the reference:
 *   - int return: 0 success, -1 failure; constructors return NULL on failure;
 *   - every entry point catches all exceptions and stores the message in a
 *     thread-local error string read with mr_last_error (c_api.cpp:183, 321-332);
 *   - caller spans are copied before return; returned pointers alias
 *     handle-owned memory valid until the next advance/reset/destroy
 *     (c_api.h:562-563, 680-681);
 *   - a handle is not internally serialized for concurrent advance calls.
 * Only plain pointers and sizes cross this boundary; device pointers are
 * accepted by the *_device entry points for inputs already resident in HBM.
 */
#ifndef MR_C_API_H
#define MR_C_API_H

#include "mr_mlp.h"
#include "mr_world_types.h"

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- library --- */

/* c_api.h:553 mr_version */
const char* mr_version(void);
/* c_api.h:554 mr_runtime_abi_fingerprint */
uint64_t mr_runtime_abi_fingerprint(void);
/* c_api.h:555-557 mr_last_error (thread-local, empty string when none) */
const char* mr_last_error(void);

/* ------------------------------------------------------------- models --- */

typedef struct MRWorldModelHandle MRWorldModelHandle;
typedef struct MRWorldHandle MRWorldHandle;

enum MRSyntheticWorldKindC {
    MR_SYNTHETIC_CHAIN_BOXES = 1u,      /* 7-DoF chain + 20 boxes on a plane */
    MR_SYNTHETIC_HUMANOID_TERRAIN = 2u, /* floating 29-DoF humanoid on a heightfield */
    /* heterogeneous batch: every world its own random tree of 6-40 revolute
     * DoF (branch probability 0.3, link masses log-uniform 0.1-10 kg) on a
     * fixed root, with 0 to (63 - links) free boxes or spheres (a world holds
     * at most 64 bodies) over a ground plane */
    MR_SYNTHETIC_RANDOM_TREES = 3u
};

typedef struct MRSyntheticWorldConfigC {
    uint32_t kind;              /* MRSyntheticWorldKindC */
    uint32_t environment_count;
    uint64_t seed;
} MRSyntheticWorldConfigC;

/* Builds a seeded synthetic world set (one model per environment for the
 * randomized chain worlds, one shared model for the humanoid). Replaces the
 * robot-factory sources of MRRunManifestC (c_api.h:160-166) for the
 * benchmark worlds. */
MRWorldModelHandle* mr_world_model_create_synthetic(const MRSyntheticWorldConfigC* config);
/* Copies a caller-described model (all arrays are copied). `microstep` is
 * the physics timestep used to precompute per-body damping scales. */
MRWorldModelHandle* mr_world_model_create(const MRWorldModelDescC* desc, float microstep);
void mr_world_model_destroy(MRWorldModelHandle* model);
/* Pointers in *out alias handle-owned arrays valid until destroy. */
int mr_world_model_describe(const MRWorldModelHandle* model, MRWorldModelDescC* out);
/* Initial state arrays: q [E][max_nq], v [E][max_nv], scene [E][max_scene_bodies]. */
int mr_world_model_initial_state(const MRWorldModelHandle* model, float* q, size_t q_count, float* v,
                                 size_t v_count, MRBodyStateC* scene, size_t scene_count);
/* Seeded synthetic command stream [steps][E][max_nv] for control steps
 * [first_step, first_step + step_count): efforts N(0, 5 N m) for the chain
 * worlds, PD targets standing pose + U[-0.3, 0.3] rad for the humanoid. */
int mr_world_model_commands(const MRWorldModelHandle* model, uint32_t first_step, uint32_t step_count, float* out,
                            size_t count);
int mr_world_model_default_config(const MRWorldModelHandle* model, MRWorldStepConfigC* out);

/* ------------------------------------------------- world (executor) --- */

/* Physics-only mirror of MetalWorldContext (MetalWorld.hpp:759-818):
 * compileMetalWorld + initializeResidentState + submitResident + wait. */
typedef struct MRWorldAdvanceC {
    uint32_t control_step_count;
    uint32_t successful_environment_steps;
    uint32_t failed_environment_steps;
    uint32_t first_failing_environment;
    uint32_t first_failing_control_step;
    uint32_t first_gpu_status_code;
    uint32_t maximum_active_pairs;
    uint32_t maximum_constraints;
    double gpu_milliseconds;
} MRWorldAdvanceC;

typedef struct MRWorldLayoutC {
    uint32_t environment_count;
    uint32_t max_nq;
    uint32_t max_nv;
    uint32_t max_scene_bodies;
    uint32_t pair_capacity;
    uint32_t manifold_capacity;
    uint32_t constraint_capacity;
    uint32_t lds_bytes_per_environment;
    uint64_t retained_device_bytes;
} MRWorldLayoutC;

/* compileMetalWorld (MetalWorld.mm:14302): validates the model, allocates the
 * resident arena and uploads the model's initial state. */
MRWorldHandle* mr_world_create(const MRWorldModelHandle* model, const MRWorldStepConfigC* config);
void mr_world_destroy(MRWorldHandle* world);
int mr_world_layout(const MRWorldHandle* world, MRWorldLayoutC* out);
/* initializeResidentState (MetalWorld.mm:16409): replaces q/v/scene, clears manifolds. */
int mr_world_initialize_resident_state(MRWorldHandle* world, const float* q, size_t q_count, const float* v,
                                       size_t v_count, const MRBodyStateC* scene, size_t scene_count);
/* Reset state applied where a reset mask is set (MetalWorldBatch resetQ/resetV). */
int mr_world_set_reset_state(MRWorldHandle* world, const float* q, size_t q_count, const float* v, size_t v_count,
                             const MRBodyStateC* scene, size_t scene_count);
/* Replaces the persistent manifolds ([E][manifold_capacity] headers, [E][capacity][4] points, [E] counts). */
int mr_world_write_manifolds(MRWorldHandle* world, const MRManifoldHeaderC* headers, const MRManifoldPointC* points,
                             const uint32_t* counts);
/* submitResident + wait (MetalWorld.mm:16409-16441, 15976): commands
 * [steps][E][max_nv] (efforts or PD targets), optional reset masks [steps][E]. */
int mr_world_advance(MRWorldHandle* world, const float* commands, size_t command_count, const uint32_t* reset_masks,
                     size_t mask_count, uint32_t control_step_count, MRWorldAdvanceC* out);
/* Same with device pointers already resident in HBM; enqueues without waiting. */
int mr_world_enqueue_device(MRWorldHandle* world, const float* device_commands, const uint32_t* device_reset_masks,
                            uint32_t control_step_count);
int mr_world_wait(MRWorldHandle* world, MRWorldAdvanceC* out);
/* Persistent horizon (no counterpart in the reference, whose encoder submits
 * every control step for all environments): needs a work-queue schedule
 * (mr_world_set_schedule with grid > 0); every scheduled world runs all
 * control_step_count steps inside one launch, its next step starting as soon
 * as its previous one ends, so no step waits for the slowest world of the
 * step before it. Inputs as mr_world_enqueue_device ([steps][E][max_nv],
 * [steps][E]); each world's results are those of one launch per control step.
 * With a reordered schedule the queue is re-sorted once, after the launch, by
 * each world's mean sweeps per control step over the horizon, most first.
 * Wait with mr_world_wait. */
int mr_world_enqueue_device_horizon(MRWorldHandle* world, const float* device_commands,
                                    const uint32_t* device_reset_masks, uint32_t control_step_count);
/* With step timing on (mr_world_set_step_timing), the wall-clock duration of
 * every control step of every world in the last eager advance, [steps][E] in
 * microseconds, measured in the kernel on the GPU's constant-rate clock
 * (s_memrealtime, whose rate is measured once per device against HIP events
 * when step timing is enabled) from the start of the world's step to the
 * store of its statuses. */
int mr_world_read_world_step_microseconds(const MRWorldHandle* world, float* out, size_t count);
/* Largest per-step scratch, in bytes, that mr_world_enqueue_device_static can serve. */
#define MR_WORLD_STATIC_SCRATCH_BYTES ((size_t)64u << 20)
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mr_world_enqueue_device that calls no hipMalloc, hipFree
 * or allocating library routine. The per-step scratch of the control-step
 * kernel (staged manifolds, Jacobian and response rows, constraint records)
 * and the advance summary live in one library-owned device array of
 * MR_WORLD_STATIC_SCRATCH_BYTES per GPU, held until the process ends, instead
 * of the world's own scratch; a world whose scratch exceeds it is refused
 * with an error. The statuses go to the caller's device buffers:
 * device_status [steps][E] MRWorldStatusC and device_contact_status [steps][E]
 * MRContactStatusC, which mr_world_read_statuses / _contact_statuses then read.
 * Call with the world's device current. Every call on one GPU shares that
 * GPU's array, which keeps the last call's contents: calls that overlap on one
 * GPU (another world's stream included) race, so order them (mr_world_wait).
 * Per-launch step timing must be off. Wait with mr_world_wait. */
int mr_world_enqueue_device_static(MRWorldHandle* world, const float* device_commands,
                                   const uint32_t* device_reset_masks, uint32_t control_step_count,
                                   MRWorldStatusC* device_status, MRContactStatusC* device_contact_status);
/* Launch schedule (no counterpart in the reference, whose encoder dispatches
 * one threadgroup per environment; for heterogeneous batches and the step
 * server). grid = 0 restores one workgroup per environment. grid > 0 runs the
 * persistent work-queue kernel with `grid` workgroups taking the environments
 * of `order` (distinct indices; a subset is allowed) in that order; unlisted
 * environments are not stepped and keep their previous status records. */
int mr_world_set_schedule(MRWorldHandle* world, const uint32_t* order, size_t count, uint32_t grid);
/* With a work-queue schedule over every environment (mr_world_set_schedule
 * with all E indices): after each control step, a one-workgroup kernel
 * re-sorts the queue by the cone sweeps each world took in that step, most
 * first, for the next step (no counterpart in the reference; for adaptive
 * residual solves, whose per-world cost varies from world to world). Results
 * do not depend on the order. enabled = 0 restores the fixed order; a new
 * mr_world_set_schedule also clears it. */
int mr_world_set_schedule_reorder(MRWorldHandle* world, uint32_t enabled);
/* Workgroups of the control-step kernel the GPU keeps resident for this world. */
uint32_t mr_world_resident_workgroups(const MRWorldHandle* world);
/* Graph-captured horizon: control_step_count steps on device commands
 * [steps][E][max_nv] at a fixed address (no reset masks), captured into a HIP
 * graph on first use and replayed while the commands pointer, the step count,
 * the schedule and the world's buffers are unchanged; wait with mr_world_wait. */
int mr_world_enqueue_graph(MRWorldHandle* world, const float* device_commands, uint32_t control_step_count);
int mr_world_read_state(const MRWorldHandle* world, float* q, size_t q_count, float* v, size_t v_count,
                        MRBodyStateC* scene, size_t scene_count);
int mr_world_read_manifolds(const MRWorldHandle* world, MRManifoldHeaderC* headers, MRManifoldPointC* points,
                            uint32_t* counts);
/* Statuses of the last advance: [steps][E] each. */
int mr_world_read_statuses(const MRWorldHandle* world, MRWorldStatusC* out, size_t count);
int mr_world_read_contact_statuses(const MRWorldHandle* world, MRContactStatusC* out, size_t count);
/* ABA acceleration of the final microstep of the last control step, zero for failed environments. */
int mr_world_read_acceleration(const MRWorldHandle* world, float* out, size_t count);
/* Active broadphase pairs of the captured microstep: [E][pair_capacity], counts [E]. */
int mr_world_read_evidence(const MRWorldHandle* world, MRPairRecordC* pairs, size_t pair_count, uint32_t* counts,
                           size_t count_count);
/* Stream the world launches on (hipStream_t), for callers that order work around it. */
void* mr_world_stream(const MRWorldHandle* world);
/* Average duration of the last advance's control-step launches (ms) measured with HIP events. */
double mr_world_last_kernel_milliseconds(const MRWorldHandle* world);
/* Per-launch HIP-event timing (off by default); durations of the last advance, [steps] in ms. */
int mr_world_set_step_timing(MRWorldHandle* world, uint32_t enabled);
int mr_world_read_step_milliseconds(const MRWorldHandle* world, double* out, size_t count);
/* Per-stage shader-clock cycle counters of the fused kernel (diagnostic, off by
 * default): [E][16] summed over the last advance (stage order: kinematics, ABA,
 * projection, broadphase, narrowphase, emission, responses, evaluation,
 * warm start, sweeps, residuals, integration). */
int mr_world_set_stage_profiling(MRWorldHandle* world, uint32_t enabled);
int mr_world_read_stage_cycles(const MRWorldHandle* world, uint64_t* out, size_t count);
/* Learned warm start (no counterpart in the reference; see mr_mlp.h): with
 * non-null device weights, later advances run the MLP instance of the
 * control-step kernel, which predicts each contact's initial impulse from
 * its features at the start of the cone solve and warm-starts from it
 * (contacts whose normalized features lie beyond 8 training standard
 * deviations keep the plain warm start); a microstep whose sweeps then miss
 * residual_tolerance is solved again from the plain warm start (contact
 * status diagnostics.z counts these). Needs
 * the adaptive solve (residual_tolerance > 0). The weights must stay valid
 * while set; null restores the plain path. */
int mr_world_set_warm_start_mlp(MRWorldHandle* world, const MRMlpWeightsC* weights);
/* Training-sample capture at microstep capture_substep of each control step:
 * per constraint, the 8 features, the converged impulse (3), a valid flag
 * (1.0 for the step's constraints, 0.0 for unused rows), the warm start the
 * cone solve began from (3) and whether the network replaced it. */
int mr_world_set_mlp_capture(MRWorldHandle* world, uint32_t enabled);
/* [E][constraint_capacity][20] floats of the last captured control step: the
 * 8 features, the converged impulse (3), a valid flag, the warm start the
 * solve began from (3), 1.0 where the network replaced the plain warm start
 * (0.0 otherwise), and with the learned warm start on, the network's scaled
 * correction (3) and 1.0 where every normalized feature was within bounds. */
int mr_world_read_mlp_samples(const MRWorldHandle* world, float* out, size_t count);

/* ------------------------------------------------- task rollout --- */

/* The task-rollout records below are laid out byte for byte as the
 * reference's (include/metalrobo/c_api.h:55-235), so a caller compiled
 * against that header passes and receives the same bytes here; sizes and
 * offsets are asserted at the end of this section. Fields that have no
 * meaning for this library's physics-only synthetic sources must be zero (or
 * null) and are rejected otherwise; output fields with no counterpart here
 * are written as zero. */

typedef struct MRTaskRolloutHandle MRTaskRolloutHandle;

/* c_api.h:55-61 */
typedef struct MRTaskRolloutDynamicSphereC {
    float position[3];
    float linear_velocity[3];
    float radius;
    float mass;
    uint32_t launch_step;
} MRTaskRolloutDynamicSphereC;

/* c_api.h:62-88. Read here: environment_count, physics_substeps,
 * velocity_iterations, final_velocity_iterations, control_timestep_seconds
 * (0 keeps each default), seed. materialize_articulated_contact_responses
 * and disable_task_terminations are accepted with any value (this library
 * always materializes the response columns and has no task terminations);
 * every other field must be zero. */
typedef struct MRTaskRolloutConfigC {
    uint32_t environment_count;
    uint32_t physics_substeps;
    uint32_t velocity_iterations;
    uint32_t final_velocity_iterations;
    float control_timestep_seconds;
    uint64_t seed;
    const MRTaskRolloutDynamicSphereC* dynamic_spheres;
    uint32_t dynamic_sphere_count;
    uint32_t disable_task_terminations;
    uint32_t interaction_reference_mode;
    float interaction_student_authority;
    uint32_t override_interaction_student_authority;
    float interaction_reset_phase_fraction;
    uint32_t override_interaction_reset_phase_fraction;
    uint32_t materialize_articulated_contact_responses;
    uint32_t override_difficulty_band_range;
    uint32_t minimum_difficulty_band;
    uint32_t maximum_difficulty_band;
    float interaction_reset_phase_probability;
    uint32_t override_interaction_reset_phase_probability;
    float interaction_reset_maximum_phase;
    uint32_t override_interaction_reset_maximum_phase;
} MRTaskRolloutConfigC;

/* c_api.h:110-127; referenced by pointer only (must be null here). */
typedef struct MRTaskVisualObservationConfigC MRTaskVisualObservationConfigC;

/* c_api.h:128-158. Written here: environment_count, nq, nv, action_count,
 * scene_body_count, world_fingerprint (FNV-1a of the compiled model
 * tables), the submission counters, retained_buffer_bytes with its split
 * into immutable (model tables), persistent state (q, v, scene bodies,
 * manifolds) and transient (per-step scratch, commands, statuses) bytes,
 * and the GPU / submission milliseconds. Observation, motion-feature,
 * episode and the other fingerprints are zero: this library runs no task
 * program. */
typedef struct MRTaskRolloutLayoutC {
    uint32_t environment_count;
    uint32_t nq;
    uint32_t nv;
    uint32_t action_count;
    uint32_t actor_observation_count;
    uint32_t critic_observation_count;
    uint32_t scene_body_count;
    uint32_t motion_feature_count;
    uint32_t maximum_episode_steps;
    uint64_t world_fingerprint;
    uint64_t task_fingerprint;
    uint64_t observation_fingerprint;
    uint64_t action_fingerprint;
    uint64_t run_fingerprint;
    uint64_t robot_fingerprint;
    uint64_t sensor_fingerprint;
    uint64_t reality_fingerprint;
    uint64_t teacher_fingerprint;
    uint64_t submitted_control_steps;
    uint64_t completed_environment_steps;
    uint64_t submission_count;
    size_t retained_buffer_bytes;
    size_t immutable_private_bytes;
    size_t persistent_state_private_bytes;
    size_t transient_private_bytes;
    size_t shared_boundary_bytes;
    size_t peak_aliased_bytes;
    double total_gpu_milliseconds;
    double total_submission_milliseconds;
} MRTaskRolloutLayoutC;

/* c_api.h:160-166 robot-factory sources (not built here: creating a
 * rollout from them fails with an error), plus this library's synthetic
 * physics worlds. */
enum MRRunManifestSourceC {
    MR_RUN_SOURCE_UNITREE_G1 = 0u,
    MR_RUN_SOURCE_FRANKA_PICK_PLACE = 1u,
    MR_RUN_SOURCE_IMPORTED_URDF = 2u,
    MR_RUN_SOURCE_WORLD_PACK = 3u,
    MR_RUN_SOURCE_PX4_X500 = 4u,
    /* 7-DoF effort-driven chain + 20 boxes on a ground plane; actions are
     * joint efforts of 5 N m per unit */
    MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES = 100u,
    /* floating 29-DoF humanoid on a 256 x 256 heightfield; actions are PD
     * targets of 0.3 rad per unit around the standing pose */
    MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN = 101u,
    /* per-environment random 6-40 DoF revolute trees with free bodies;
     * actions are joint efforts of 1 N m per unit */
    MR_RUN_SOURCE_SYNTHETIC_RANDOM_TREES = 102u
};

/* c_api.h:29-32 */
enum MRLocomotionSurfaceC { MR_LOCOMOTION_SURFACE_GROUND = 0u, MR_LOCOMOTION_SURFACE_TERRAIN = 1u };

/* c_api.h:171-190. Read here: profile, source, surface (must name the
 * synthetic source's own ground: TERRAIN for the humanoid, GROUND for the
 * others). task and every pack path / visual program must be zero or null;
 * metallib_path names a Metal shader library and is ignored. */
typedef struct MRRunManifestC {
    MRTaskRolloutConfigC profile;
    uint32_t source;
    uint32_t surface;
    uint32_t task;
    const char* urdf_path;
    const char* srdf_path;
    const char* world_pack_path;
    const char* task_pack_path;
    const char* robot_actuator_pack_path;
    const char* sensor_program_pack_path;
    const char* reality_program_pack_path;
    const char* teacher_pack_path;
    const char* teacher_clip_id;
    const MRTaskVisualObservationConfigC* visual_sensor_program;
    const MRTaskVisualObservationConfigC* inspection_visual_program;
    const char* metallib_path;
} MRRunManifestC;

/* c_api.h:192-220. Maxima over the advance's published per-environment,
 * per-control-step records: candidate_pairs (active broadphase pairs),
 * raw_contacts, manifolds, constraint_blocks (contacts), constraint_rows
 * (3 per contact), islands, articulation_point_queries (2 per contact)
 * and operator_velocity_elements (nv). Stages this library does not have
 * (convex, mesh, CCD, rods, quality solver, tiles) and the island
 * reference counts stay zero. */
typedef struct MRTaskRolloutStageHighWaterC {
    uint32_t candidate_pairs;
    uint32_t raw_contacts;
    uint32_t manifolds;
    uint32_t constraint_blocks;
    uint32_t constraint_rows;
    uint32_t islands;
    uint32_t hard_convex_pairs;
    uint32_t mesh_triangle_candidates;
    uint32_t solver_tiles;
    uint32_t spill_rows;
    uint32_t ccd_candidates;
    uint32_t ccd_events;
    uint32_t endpoint_runtime_records;
    uint32_t articulation_point_queries;
    uint32_t rod_candidate_pairs;
    uint32_t rod_raw_contacts;
    uint32_t rod_manifolds;
    uint32_t rod_ccd_events;
    uint32_t quality_generalized_velocities;
    uint32_t quality_rows;
    uint32_t quality_krylov_vectors;
    uint32_t quality_direct_tiles;
    uint32_t dynamic_nodes;
    uint32_t island_node_references;
    uint32_t island_constraint_references;
    uint32_t rod_factor_blocks;
    uint32_t operator_velocity_elements;
} MRTaskRolloutStageHighWaterC;

/* c_api.h:222-235 */
typedef struct MRTaskRolloutAdvanceC {
    uint32_t control_step_count;
    uint32_t successful_environment_steps;
    uint32_t failed_environment_steps;
    uint32_t first_failing_environment;
    uint32_t first_failing_control_step;
    uint32_t first_gpu_status_code;
    uint32_t host_requested_resets;
    uint32_t maximum_active_contacts;
    uint32_t maximum_manifolds;
    MRTaskRolloutStageHighWaterC high_water;
    double gpu_milliseconds;
    double submission_milliseconds;
} MRTaskRolloutAdvanceC;

#ifdef __cplusplus
#define MR_TR_ASSERT(c, m) static_assert(c, m)
#else
#define MR_TR_ASSERT(c, m) _Static_assert(c, m)
#endif
/* Sizes and offsets of the reference's records on LP64 (x86-64 / arm64). */
MR_TR_ASSERT(sizeof(MRTaskRolloutDynamicSphereC) == 36, "MRTaskRolloutDynamicSphereC size");
MR_TR_ASSERT(sizeof(MRTaskRolloutConfigC) == 104, "MRTaskRolloutConfigC size");
MR_TR_ASSERT(offsetof(MRTaskRolloutConfigC, seed) == 24, "MRTaskRolloutConfigC.seed");
MR_TR_ASSERT(offsetof(MRTaskRolloutConfigC, dynamic_spheres) == 32, "MRTaskRolloutConfigC.dynamic_spheres");
MR_TR_ASSERT(offsetof(MRTaskRolloutConfigC, override_interaction_reset_maximum_phase) == 96,
             "MRTaskRolloutConfigC tail");
MR_TR_ASSERT(sizeof(MRTaskRolloutLayoutC) == 200, "MRTaskRolloutLayoutC size");
MR_TR_ASSERT(offsetof(MRTaskRolloutLayoutC, world_fingerprint) == 40, "MRTaskRolloutLayoutC.world_fingerprint");
MR_TR_ASSERT(offsetof(MRTaskRolloutLayoutC, retained_buffer_bytes) == 136, "MRTaskRolloutLayoutC.retained_buffer_bytes");
MR_TR_ASSERT(offsetof(MRTaskRolloutLayoutC, total_gpu_milliseconds) == 184, "MRTaskRolloutLayoutC.total_gpu_milliseconds");
MR_TR_ASSERT(sizeof(MRRunManifestC) == 216, "MRRunManifestC size");
MR_TR_ASSERT(offsetof(MRRunManifestC, source) == 104, "MRRunManifestC.source");
MR_TR_ASSERT(offsetof(MRRunManifestC, urdf_path) == 120, "MRRunManifestC.urdf_path");
MR_TR_ASSERT(offsetof(MRRunManifestC, metallib_path) == 208, "MRRunManifestC.metallib_path");
MR_TR_ASSERT(sizeof(MRTaskRolloutStageHighWaterC) == 108, "MRTaskRolloutStageHighWaterC size");
MR_TR_ASSERT(sizeof(MRTaskRolloutAdvanceC) == 160, "MRTaskRolloutAdvanceC size");
MR_TR_ASSERT(offsetof(MRTaskRolloutAdvanceC, high_water) == 36, "MRTaskRolloutAdvanceC.high_water");
MR_TR_ASSERT(offsetof(MRTaskRolloutAdvanceC, gpu_milliseconds) == 144, "MRTaskRolloutAdvanceC.gpu_milliseconds");

/* c_api.h:583-585 */
MRTaskRolloutHandle* mr_create_task_rollout(const MRRunManifestC* manifest);
/* c_api.h:586 */
void mr_task_rollout_destroy(MRTaskRolloutHandle* handle);
/* c_api.h:587-590: restores the seeded initial state of every environment. */
int mr_task_rollout_reset(MRTaskRolloutHandle* handle, uint64_t seed);
/* c_api.h:631-634 */
int mr_task_rollout_set_state_readback(MRTaskRolloutHandle* handle, uint32_t enabled);
/* c_api.h:635-645: normalized actions [steps][envs][actions] in [-1, 1],
 * optional reset masks [steps][envs]. */
int mr_task_rollout_advance(MRTaskRolloutHandle* handle, const float* normalized_actions, size_t action_count,
                            const uint32_t* reset_masks, size_t mask_count, uint32_t control_step_count,
                            uint64_t policy_revision, uint32_t evaluate_final_policy, MRTaskRolloutAdvanceC* out);
/* c_api.h:646-648 */
MRTaskRolloutLayoutC mr_task_rollout_layout(const MRTaskRolloutHandle* handle);
/* c_api.h:682-687: [steps][envs] of the last advance. */
const uint32_t* mr_task_rollout_status_codes(const MRTaskRolloutHandle* handle);
const uint32_t* mr_task_rollout_active_contacts(const MRTaskRolloutHandle* handle);
/* c_api.h:736-738: [envs][nq] after the last advance (state readback enabled). */
const float* mr_task_rollout_final_q(const MRTaskRolloutHandle* handle);
/* c_api.h:739-746: [env][body][pos3, quat4, linv3, angv3]. */
int mr_task_rollout_copy_final_scene_states(const MRTaskRolloutHandle* handle, float* out, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* MR_C_API_H */
