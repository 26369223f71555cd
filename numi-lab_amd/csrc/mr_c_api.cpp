// extern "C" boundary: handle ownership, argument validation and exception
// translation (src/c_api.cpp:98-134, 183, 321-332 conventions).
#include "../../include/mr_c_api.h"

#include "mr_builder.h"
#include "mr_mlp_host.h"
#include "mr_packs.h"
#include "mr_policy.h"
#include "mr_task_host.h"
#include "mr_world.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct MRWorldModelHandle {
    mr::ModelData model;
};

struct MRWorldHandle {
    std::unique_ptr<mr::World> world;
};

struct MRLocomotionTaskHandle {
    std::unique_ptr<mr::LocomotionTask> task;
};

struct MRPolicyPackHandle {
    mr::PolicyPack pack;
    std::vector<MRPolicyDenseLayerC> layers, critic_layers;
    MRPolicyPackC view{};
};

struct MRPolicyHandle {
    std::unique_ptr<mr::Policy> policy;
};

struct MRTaskRolloutHandle {
    mr::ModelData model;
    std::unique_ptr<mr::World> world;
    uint32_t action_count = 0;
    float action_scale = 1.0f;
    bool state_readback = true;
    std::vector<uint32_t> status_codes;
    std::vector<uint32_t> active_contacts;
    std::vector<float> final_q;
    std::vector<MRBodyStateC> final_scene;
    uint64_t submitted_control_steps = 0;
    uint64_t completed_environment_steps = 0;
    uint64_t submission_count = 0;
    double total_gpu_ms = 0.0;
    double total_submission_ms = 0.0;
};

namespace {

thread_local std::string gLastError;

template <typename F>
int translate(F&& f) {
    try {
        gLastError.clear();
        f();
        return 0;
    } catch (const std::exception& e) {
        gLastError = e.what();
    } catch (...) {
        gLastError = "unknown native exception";
    }
    return -1;
}

template <typename T, typename F>
T* construct(F&& f) {
    try {
        gLastError.clear();
        return f();
    } catch (const std::exception& e) {
        gLastError = e.what();
    } catch (...) {
        gLastError = "unknown native exception";
    }
    return nullptr;
}

void require(bool c, const char* msg) {
    if (!c) throw std::runtime_error(msg);
}

void require_count(size_t got, size_t want, const char* what) {
    if (got != want)
        throw std::runtime_error(std::string(what) + ": expected " + std::to_string(want) + " elements, got " +
                                 std::to_string(got));
}

void fill_advance(const mr::AdvanceSummary& s, MRWorldAdvanceC* out) {
    if (out == nullptr) return;
    out->control_step_count = s.control_step_count;
    out->successful_environment_steps = s.successful_environment_steps;
    out->failed_environment_steps = s.failed_environment_steps;
    out->first_failing_environment = s.first_failing_environment;
    out->first_failing_control_step = s.first_failing_control_step;
    out->first_gpu_status_code = s.first_gpu_status_code;
    out->maximum_active_pairs = s.maximum_active_pairs;
    out->maximum_constraints = s.maximum_constraints;
    out->gpu_milliseconds = s.gpu_milliseconds;
}

} // namespace

extern "C" {

const char* mr_version(void) { return "numi-lab_amd 0.1.0 (gfx950 temporalCone)"; }

uint64_t mr_runtime_abi_fingerprint(void) {
    uint64_t h = 1469598103934665603ull;
    const uint64_t parts[] = {MR_WORLD_ABI_VERSION, sizeof(MRBodyC), sizeof(MRShapeC), sizeof(MRBodyStateC),
                              sizeof(MRManifoldHeaderC), sizeof(MRManifoldPointC), sizeof(MRWorldStatusC),
                              sizeof(MRContactStatusC), sizeof(MRWorldStepConfigC)};
    for (uint64_t p : parts) {
        h ^= p;
        h *= 1099511628211ull;
    }
    return h;
}

const char* mr_last_error(void) { return gLastError.c_str(); }

MRWorldModelHandle* mr_world_model_create_synthetic(const MRSyntheticWorldConfigC* config) {
    return construct<MRWorldModelHandle>([&]() {
        require(config != nullptr, "config is null");
        auto* h = new MRWorldModelHandle();
        try {
            h->model = mr::build_synthetic(config->kind, config->environment_count, config->seed);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    });
}

MRWorldModelHandle* mr_world_model_create(const MRWorldModelDescC* desc, float microstep) {
    return construct<MRWorldModelHandle>([&]() {
        require(desc != nullptr, "desc is null");
        require(microstep > 0.0f, "microstep must be positive");
        auto* h = new MRWorldModelHandle();
        try {
            h->model = mr::copy_model(*desc, microstep);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    });
}

void mr_world_model_destroy(MRWorldModelHandle* model) { delete model; }

int mr_world_model_describe(const MRWorldModelHandle* model, MRWorldModelDescC* out) {
    return translate([&]() {
        require(model != nullptr && out != nullptr, "null argument");
        *out = model->model.describe();
    });
}

int mr_world_model_initial_state(const MRWorldModelHandle* model, float* q, size_t q_count, float* v,
                                 size_t v_count, MRBodyStateC* scene, size_t scene_count) {
    return translate([&]() {
        require(model != nullptr, "model is null");
        const mr::ModelData& m = model->model;
        require(m.initial_scene.size() == (size_t)m.environment_count() * m.max_scene_bodies,
                "model has no initial state");
        require_count(q_count, m.initial_q.size(), "q");
        require_count(v_count, m.initial_v.size(), "v");
        require_count(scene_count, m.initial_scene.size(), "scene");
        std::memcpy(q, m.initial_q.data(), q_count * sizeof(float));
        std::memcpy(v, m.initial_v.data(), v_count * sizeof(float));
        if (scene_count) std::memcpy(scene, m.initial_scene.data(), scene_count * sizeof(MRBodyStateC));
    });
}

int mr_world_model_commands(const MRWorldModelHandle* model, uint32_t first_step, uint32_t step_count, float* out,
                            size_t count) {
    return translate([&]() {
        require(model != nullptr && out != nullptr, "null argument");
        const mr::ModelData& m = model->model;
        require_count(count, (size_t)step_count * m.environment_count() * m.max_nv, "commands");
        m.commands(first_step, step_count, out);
    });
}

int mr_world_model_default_config(const MRWorldModelHandle* model, MRWorldStepConfigC* out) {
    return translate([&]() {
        require(model != nullptr && out != nullptr, "null argument");
        *out = model->model.default_config;
    });
}

MRWorldHandle* mr_world_create(const MRWorldModelHandle* model, const MRWorldStepConfigC* config) {
    return construct<MRWorldHandle>([&]() {
        require(model != nullptr && config != nullptr, "null argument");
        auto* h = new MRWorldHandle();
        try {
            h->world = std::make_unique<mr::World>(model->model, *config);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    });
}

void mr_world_destroy(MRWorldHandle* world) { delete world; }

int mr_world_layout(const MRWorldHandle* world, MRWorldLayoutC* out) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        const mr::World& w = *world->world;
        out->environment_count = w.environment_count();
        out->max_nq = w.model().max_nq;
        out->max_nv = w.model().max_nv;
        out->max_scene_bodies = w.model().max_scene_bodies;
        out->pair_capacity = w.config().pair_capacity;
        out->manifold_capacity = w.config().manifold_capacity;
        out->constraint_capacity = w.config().constraint_capacity;
        out->lds_bytes_per_environment = w.lds().total;
        out->retained_device_bytes = w.retained_bytes();
    });
}

int mr_world_initialize_resident_state(MRWorldHandle* world, const float* q, size_t q_count, const float* v,
                                       size_t v_count, const MRBodyStateC* scene, size_t scene_count) {
    return translate([&]() {
        require(world != nullptr && q != nullptr && v != nullptr, "null argument");
        const mr::ModelData& m = world->world->model();
        const size_t E = world->world->environment_count();
        require_count(q_count, E * m.max_nq, "q");
        require_count(v_count, E * m.max_nv, "v");
        require_count(scene_count, E * m.max_scene_bodies, "scene");
        world->world->set_state(q, v, scene);
    });
}

int mr_world_set_reset_state(MRWorldHandle* world, const float* q, size_t q_count, const float* v, size_t v_count,
                             const MRBodyStateC* scene, size_t scene_count) {
    return translate([&]() {
        require(world != nullptr && q != nullptr && v != nullptr, "null argument");
        const mr::ModelData& m = world->world->model();
        const size_t E = world->world->environment_count();
        require_count(q_count, E * m.max_nq, "q");
        require_count(v_count, E * m.max_nv, "v");
        require_count(scene_count, E * m.max_scene_bodies, "scene");
        world->world->set_reset_state(q, v, scene);
    });
}

int mr_world_write_manifolds(MRWorldHandle* world, const MRManifoldHeaderC* headers, const MRManifoldPointC* points,
                             const uint32_t* counts) {
    return translate([&]() {
        require(world != nullptr && headers != nullptr && points != nullptr && counts != nullptr, "null argument");
        world->world->set_manifolds(headers, points, counts);
    });
}

int mr_world_advance(MRWorldHandle* world, const float* commands, size_t command_count, const uint32_t* reset_masks,
                     size_t mask_count, uint32_t control_step_count, MRWorldAdvanceC* out) {
    return translate([&]() {
        require(world != nullptr && commands != nullptr, "null argument");
        mr::World& w = *world->world;
        const size_t E = w.environment_count();
        require(control_step_count > 0u, "control_step_count must be positive");
        require_count(command_count, (size_t)control_step_count * E * w.model().max_nv, "commands");
        if (reset_masks != nullptr) require_count(mask_count, (size_t)control_step_count * E, "reset masks");
        for (size_t i = 0; i < command_count; ++i)
            if (!std::isfinite(commands[i])) throw std::runtime_error("nonfinite command input");
        w.enqueue(commands, reset_masks, control_step_count, false);
        fill_advance(w.wait(), out);
    });
}

int mr_world_enqueue_device(MRWorldHandle* world, const float* device_commands, const uint32_t* device_reset_masks,
                            uint32_t control_step_count) {
    return translate([&]() {
        require(world != nullptr && device_commands != nullptr, "null argument");
        world->world->enqueue(device_commands, device_reset_masks, control_step_count, true);
    });
}

int mr_world_enqueue_device_horizon(MRWorldHandle* world, const float* device_commands,
                                    const uint32_t* device_reset_masks, uint32_t control_step_count) {
    return translate([&]() {
        require(world != nullptr && device_commands != nullptr, "null argument");
        world->world->enqueue_horizon(device_commands, device_reset_masks, control_step_count);
    });
}

int mr_world_read_world_step_microseconds(const MRWorldHandle* world, float* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        const mr::World& w = *world->world;
        require_count(count, (size_t)w.last_steps() * w.environment_count(), "world step times");
        w.read_world_step_microseconds(out, w.last_steps());
    });
}

int mr_world_enqueue_device_static(MRWorldHandle* world, const float* device_commands,
                                   const uint32_t* device_reset_masks, uint32_t control_step_count,
                                   MRWorldStatusC* device_status, MRContactStatusC* device_contact_status) {
    return translate([&]() {
        require(world != nullptr && device_commands != nullptr, "null argument");
        world->world->enqueue_static(device_commands, device_reset_masks, control_step_count, device_status,
                                     device_contact_status);
    });
}

int mr_world_set_schedule(MRWorldHandle* world, const uint32_t* order, size_t count, uint32_t grid) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        require(count <= world->world->environment_count(), "set_schedule: more entries than environments");
        world->world->set_schedule(order, (uint32_t)count, grid);
    });
}

int mr_world_set_schedule_reorder(MRWorldHandle* world, uint32_t enabled) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        world->world->set_schedule_reorder(enabled != 0u);
    });
}

uint32_t mr_world_resident_workgroups(const MRWorldHandle* world) {
    uint32_t out = 0u;
    (void)translate([&]() {
        require(world != nullptr, "null argument");
        out = world->world->resident_workgroups();
    });
    return out;
}

int mr_world_enqueue_graph(MRWorldHandle* world, const float* device_commands, uint32_t control_step_count) {
    return translate([&]() {
        require(world != nullptr && device_commands != nullptr, "null argument");
        world->world->enqueue_graph(device_commands, control_step_count);
    });
}

int mr_world_wait(MRWorldHandle* world, MRWorldAdvanceC* out) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        fill_advance(world->world->wait(), out);
    });
}

int mr_world_read_state(const MRWorldHandle* world, float* q, size_t q_count, float* v, size_t v_count,
                        MRBodyStateC* scene, size_t scene_count) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        const mr::ModelData& m = world->world->model();
        const size_t E = world->world->environment_count();
        if (q) require_count(q_count, E * m.max_nq, "q");
        if (v) require_count(v_count, E * m.max_nv, "v");
        if (scene) require_count(scene_count, E * m.max_scene_bodies, "scene");
        world->world->get_state(q, v, scene);
    });
}

int mr_world_read_manifolds(const MRWorldHandle* world, MRManifoldHeaderC* headers, MRManifoldPointC* points,
                            uint32_t* counts) {
    return translate([&]() {
        require(world != nullptr && headers != nullptr && points != nullptr && counts != nullptr, "null argument");
        world->world->get_manifolds(headers, points, counts);
    });
}

int mr_world_read_statuses(const MRWorldHandle* world, MRWorldStatusC* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        const mr::World& w = *world->world;
        require_count(count, (size_t)w.last_steps() * w.environment_count(), "statuses");
        w.read_statuses(out, w.last_steps());
    });
}

int mr_world_read_contact_statuses(const MRWorldHandle* world, MRContactStatusC* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        const mr::World& w = *world->world;
        require_count(count, (size_t)w.last_steps() * w.environment_count(), "contact statuses");
        w.read_contact_statuses(out, w.last_steps());
    });
}

int mr_world_read_acceleration(const MRWorldHandle* world, float* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        const mr::World& w = *world->world;
        require_count(count, (size_t)w.environment_count() * w.model().max_nv, "acceleration");
        w.read_acceleration(out);
    });
}

int mr_world_read_evidence(const MRWorldHandle* world, MRPairRecordC* pairs, size_t pair_count, uint32_t* counts,
                           size_t count_count) {
    return translate([&]() {
        require(world != nullptr && pairs != nullptr && counts != nullptr, "null argument");
        const mr::World& w = *world->world;
        require(w.config().capture_evidence != 0u, "world was created without evidence capture");
        require_count(pair_count, (size_t)w.environment_count() * w.config().pair_capacity, "evidence pairs");
        require_count(count_count, (size_t)w.environment_count(), "evidence counts");
        w.read_evidence(pairs, counts);
    });
}

void* mr_world_stream(const MRWorldHandle* world) {
    return world != nullptr ? static_cast<void*>(world->world->stream()) : nullptr;
}

double mr_world_last_kernel_milliseconds(const MRWorldHandle* world) {
    if (world == nullptr || world->world->last_steps() == 0u) return 0.0;
    return (double)world->world->last_kernel_milliseconds() / (double)world->world->last_steps();
}

int mr_world_set_step_timing(MRWorldHandle* world, uint32_t enabled) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        world->world->set_step_timing(enabled != 0u);
    });
}

int mr_world_read_step_milliseconds(const MRWorldHandle* world, double* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        require_count(count, world->world->last_steps(), "step durations");
        world->world->read_step_milliseconds(out, world->world->last_steps());
    });
}

int mr_world_set_stage_profiling(MRWorldHandle* world, uint32_t enabled) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        world->world->set_stage_profiling(enabled != 0u);
    });
}

int mr_world_read_stage_cycles(const MRWorldHandle* world, uint64_t* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        require_count(count, (size_t)world->world->environment_count() * mr::kStageCount, "stage cycles");
        world->world->read_stage_cycles(reinterpret_cast<unsigned long long*>(out));
    });
}

int mr_world_set_warm_start_mlp(MRWorldHandle* world, const MRMlpWeightsC* weights) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        world->world->set_warm_start_mlp(weights);
    });
}

int mr_world_set_mlp_capture(MRWorldHandle* world, uint32_t enabled) {
    return translate([&]() {
        require(world != nullptr, "null argument");
        world->world->set_mlp_capture(enabled != 0u);
    });
}

int mr_world_read_mlp_samples(const MRWorldHandle* world, float* out, size_t count) {
    return translate([&]() {
        require(world != nullptr && out != nullptr, "null argument");
        require_count(count,
                      (size_t)world->world->environment_count() * world->world->config().constraint_capacity *
                          mr::kMlpSampleFloats,
                      "MLP samples");
        world->world->read_mlp_samples(out);
    });
}

// ------------------------------------------------------------ task rollout

MRTaskRolloutHandle* mr_create_task_rollout(const MRRunManifestC* manifest) {
    return construct<MRTaskRolloutHandle>([&]() {
        require(manifest != nullptr, "manifest is null");
        const MRTaskRolloutConfigC& p = manifest->profile;
        uint32_t kind = 0, surface = MR_LOCOMOTION_SURFACE_GROUND;
        switch (manifest->source) {
        case MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES: kind = mr::kChainBoxes; break;
        case MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN:
            kind = mr::kHumanoidTerrain;
            surface = MR_LOCOMOTION_SURFACE_TERRAIN;
            break;
        case MR_RUN_SOURCE_SYNTHETIC_RANDOM_TREES: kind = mr::kRandomTrees; break;
        case MR_RUN_SOURCE_UNITREE_G1:
        case MR_RUN_SOURCE_FRANKA_PICK_PLACE:
        case MR_RUN_SOURCE_IMPORTED_URDF:
        case MR_RUN_SOURCE_WORLD_PACK:
        case MR_RUN_SOURCE_PX4_X500:
            throw std::runtime_error("unsupported run source " + std::to_string(manifest->source) +
                                     ": robot-factory and pack sources are not built by this engine; use a "
                                     "synthetic source (100-102)");
        default: throw std::runtime_error("unsupported run source " + std::to_string(manifest->source));
        }
        require(p.environment_count > 0u, "profile.environment_count must be positive");
        require(manifest->surface == surface,
                "manifest.surface must name the synthetic source's own ground (TERRAIN for the humanoid, "
                "GROUND otherwise)");
        require(manifest->task == 0u, "manifest.task must be 0: this library runs no task program");
        require(manifest->urdf_path == nullptr && manifest->srdf_path == nullptr &&
                    manifest->world_pack_path == nullptr && manifest->task_pack_path == nullptr &&
                    manifest->robot_actuator_pack_path == nullptr && manifest->sensor_program_pack_path == nullptr &&
                    manifest->reality_program_pack_path == nullptr && manifest->teacher_pack_path == nullptr &&
                    manifest->teacher_clip_id == nullptr,
                "manifest pack paths must be null for a synthetic source");
        require(manifest->visual_sensor_program == nullptr && manifest->inspection_visual_program == nullptr,
                "visual programs are not supported by this library");
        require(p.dynamic_spheres == nullptr && p.dynamic_sphere_count == 0u,
                "profile.dynamic_spheres are not supported by this library");
        require(p.interaction_reference_mode == 0u && p.interaction_student_authority == 0.0f &&
                    p.override_interaction_student_authority == 0u && p.interaction_reset_phase_fraction == 0.0f &&
                    p.override_interaction_reset_phase_fraction == 0u && p.override_difficulty_band_range == 0u &&
                    p.minimum_difficulty_band == 0u && p.maximum_difficulty_band == 0u &&
                    p.interaction_reset_phase_probability == 0.0f &&
                    p.override_interaction_reset_phase_probability == 0u &&
                    p.interaction_reset_maximum_phase == 0.0f && p.override_interaction_reset_maximum_phase == 0u,
                "profile interaction and difficulty-band fields must be zero: this library runs no task program");
        auto* h = new MRTaskRolloutHandle();
        try {
            h->model = mr::build_synthetic(kind, p.environment_count, p.seed);
            MRWorldStepConfigC cfg = h->model.default_config;
            if (p.physics_substeps) cfg.physics_substeps = p.physics_substeps;
            if (p.velocity_iterations) cfg.velocity_iterations = p.velocity_iterations;
            if (p.final_velocity_iterations) cfg.final_velocity_iterations = p.final_velocity_iterations;
            if (p.control_timestep_seconds > 0.0f) cfg.control_timestep = p.control_timestep_seconds;
            h->world = std::make_unique<mr::World>(h->model, cfg);
            const MRModelHeaderC& m0 = h->model.models[0];
            h->action_count = m0.root_type == MR_ROOT_FLOATING ? h->model.max_nv - 6u : h->model.max_nv;
            h->action_scale = kind == mr::kChainBoxes ? 5.0f : (kind == mr::kHumanoidTerrain ? 0.3f : 1.0f);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    });
}

void mr_task_rollout_destroy(MRTaskRolloutHandle* handle) { delete handle; }

int mr_task_rollout_reset(MRTaskRolloutHandle* handle, uint64_t seed) {
    return translate([&]() {
        require(handle != nullptr, "null handle");
        if (seed != handle->model.seed) {
            const uint32_t E = handle->model.environment_count();
            mr::ModelData rebuilt = mr::build_synthetic(handle->model.kind, E, seed);
            handle->model = rebuilt;
            handle->world = std::make_unique<mr::World>(handle->model, handle->world->config());
        } else {
            handle->world->set_state(handle->model.initial_q.data(), handle->model.initial_v.data(),
                                     handle->model.initial_scene.data());
        }
        handle->status_codes.clear();
        handle->active_contacts.clear();
        handle->final_q.clear();
        handle->final_scene.clear();
    });
}

int mr_task_rollout_set_state_readback(MRTaskRolloutHandle* handle, uint32_t enabled) {
    return translate([&]() {
        require(handle != nullptr, "null handle");
        handle->state_readback = enabled != 0u;
    });
}

int mr_task_rollout_advance(MRTaskRolloutHandle* handle, const float* normalized_actions, size_t action_count,
                            const uint32_t* reset_masks, size_t mask_count, uint32_t control_step_count,
                            uint64_t policy_revision, uint32_t evaluate_final_policy, MRTaskRolloutAdvanceC* out) {
    (void)policy_revision;  // no policy program: the host action stream drives every step
    if (out == nullptr) {
        gLastError = "task-rollout advance result is null.";
        return -1;
    }
    std::memset(out, 0, sizeof(*out));
    out->first_failing_environment = MR_INVALID_INDEX;
    out->first_failing_control_step = MR_INVALID_INDEX;
    return translate([&]() {
        const auto t0 = std::chrono::steady_clock::now();
        require(handle != nullptr, "null handle");
        require(control_step_count > 0u, "task-rollout control_step_count must be positive.");
        require(evaluate_final_policy == 0u, "this library has no policy program to evaluate");
        mr::World& w = *handle->world;
        const mr::ModelData& m = handle->model;
        const size_t E = w.environment_count(), A = handle->action_count, nv = m.max_nv;
        require(normalized_actions != nullptr && action_count == (size_t)control_step_count * E * A,
                "normalized task action count must equal step_count * environment_count * compiled action_count");
        require((reset_masks == nullptr && mask_count == 0u) ||
                    (reset_masks != nullptr && mask_count == (size_t)control_step_count * E),
                "task reset mask count must be zero or step_count * environment_count");
        std::vector<float> commands((size_t)control_step_count * E * nv, 0.0f);
        const uint32_t first = m.models[0].root_type == MR_ROOT_FLOATING ? 6u : 0u;
        uint32_t resets = 0;
        for (size_t s = 0; s < control_step_count; ++s)
            for (size_t e = 0; e < E; ++e) {
                if (reset_masks != nullptr && reset_masks[s * E + e] != 0u) ++resets;
                const uint32_t env_nv = m.models[m.environment_model[e]].nv;
                for (size_t a = 0; a < A; ++a) {
                    float x = normalized_actions[(s * E + e) * A + a];
                    if (!std::isfinite(x)) throw std::runtime_error("nonfinite action input");
                    const size_t j = first + a;
                    if (j >= env_nv) continue;  // random trees: beyond this world's DoFs
                    x = x < -1.0f ? -1.0f : (x > 1.0f ? 1.0f : x);
                    const float base = m.kind == mr::kHumanoidTerrain ? m.standing_pose[j] : 0.0f;
                    commands[(s * E + e) * nv + j] = base + handle->action_scale * x;
                }
            }
        out->host_requested_resets = resets;
        w.enqueue(commands.data(), reset_masks, control_step_count, false);
        const mr::AdvanceSummary sum = w.wait();
        const size_t n = (size_t)control_step_count * E;
        std::vector<MRWorldStatusC> st(n);
        std::vector<MRContactStatusC> cs(n);
        w.read_statuses(st.data(), control_step_count);
        w.read_contact_statuses(cs.data(), control_step_count);
        handle->status_codes.resize(n);
        handle->active_contacts.resize(n);
        MRTaskRolloutStageHighWaterC& hw = out->high_water;
        for (size_t i = 0; i < n; ++i) {
            handle->status_codes[i] = st[i].code != MR_STEP_SUCCESS ? st[i].code : cs[i].code;
            handle->active_contacts[i] = cs[i].constraints;
            out->maximum_active_contacts = std::max(out->maximum_active_contacts, cs[i].constraints);
            out->maximum_manifolds = std::max(out->maximum_manifolds, cs[i].required_manifolds);
            hw.candidate_pairs = std::max(hw.candidate_pairs, cs[i].max_active_pairs);
            hw.raw_contacts = std::max(hw.raw_contacts, cs[i].raw_contacts);
            hw.manifolds = std::max(hw.manifolds, cs[i].manifolds);
            hw.constraint_blocks = std::max(hw.constraint_blocks, cs[i].max_constraints);
            hw.islands = std::max(hw.islands, cs[i].islands);
        }
        hw.constraint_rows = 3u * hw.constraint_blocks;
        hw.articulation_point_queries = 2u * hw.constraint_blocks;
        hw.operator_velocity_elements = m.max_nv;
        if (handle->state_readback) {
            handle->final_q.resize(E * m.max_nq);
            handle->final_scene.resize(E * m.max_scene_bodies);
            w.get_state(handle->final_q.data(), nullptr, handle->final_scene.data());
        }
        const double wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        handle->submitted_control_steps += control_step_count;
        handle->completed_environment_steps += sum.successful_environment_steps;
        handle->submission_count += 1u;
        handle->total_gpu_ms += sum.gpu_milliseconds;
        handle->total_submission_ms += wall_ms;
        out->control_step_count = control_step_count;
        out->successful_environment_steps = sum.successful_environment_steps;
        out->failed_environment_steps = sum.failed_environment_steps;
        out->first_failing_environment = sum.first_failing_environment;
        out->first_failing_control_step = sum.first_failing_control_step;
        out->first_gpu_status_code = sum.first_gpu_status_code;
        out->gpu_milliseconds = sum.gpu_milliseconds;
        out->submission_milliseconds = wall_ms;
        if (sum.failed_environment_steps > 0u)
            throw std::runtime_error("task-rollout GPU step failed with status " +
                                     std::to_string(sum.first_gpu_status_code) + " in environment " +
                                     std::to_string(sum.first_failing_environment) + " at control step " +
                                     std::to_string(sum.first_failing_control_step));
    });
}

namespace {
// FNV-1a over the compiled model tables: equal for equal worlds.
uint64_t world_fingerprint(const mr::ModelData& m) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* data, size_t bytes) {
        const auto* b = static_cast<const unsigned char*>(data);
        for (size_t i = 0; i < bytes; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    auto mix_vec = [&](const auto& v) { mix(v.data(), v.size() * sizeof(v[0])); };
    mix_vec(m.models);
    mix_vec(m.environment_model);
    mix_vec(m.bodies);
    mix_vec(m.dofs);
    mix_vec(m.shapes);
    mix_vec(m.materials);
    mix_vec(m.pairs);
    mix_vec(m.heights);
    mix(&m.heightfield, sizeof(m.heightfield));
    mix(&m.gravity, sizeof(m.gravity));
    return h;
}
} // namespace

MRTaskRolloutLayoutC mr_task_rollout_layout(const MRTaskRolloutHandle* handle) {
    MRTaskRolloutLayoutC out;
    std::memset(&out, 0, sizeof(out));
    if (handle == nullptr) return out;
    const mr::World& w = *handle->world;
    out.environment_count = w.environment_count();
    out.nq = handle->model.max_nq;
    out.nv = handle->model.max_nv;
    out.action_count = handle->action_count;
    out.scene_body_count = handle->model.max_scene_bodies;
    out.world_fingerprint = world_fingerprint(handle->model);
    out.submitted_control_steps = handle->submitted_control_steps;
    out.completed_environment_steps = handle->completed_environment_steps;
    out.submission_count = handle->submission_count;
    out.immutable_private_bytes = w.immutable_bytes();
    out.persistent_state_private_bytes = w.persistent_bytes();
    out.transient_private_bytes = w.transient_bytes();
    out.retained_buffer_bytes = out.immutable_private_bytes + out.persistent_state_private_bytes +
                                out.transient_private_bytes;
    out.total_gpu_milliseconds = handle->total_gpu_ms;
    out.total_submission_milliseconds = handle->total_submission_ms;
    return out;
}

const uint32_t* mr_task_rollout_status_codes(const MRTaskRolloutHandle* handle) {
    return handle != nullptr && !handle->status_codes.empty() ? handle->status_codes.data() : nullptr;
}

const uint32_t* mr_task_rollout_active_contacts(const MRTaskRolloutHandle* handle) {
    return handle != nullptr && !handle->active_contacts.empty() ? handle->active_contacts.data() : nullptr;
}

const float* mr_task_rollout_final_q(const MRTaskRolloutHandle* handle) {
    return handle != nullptr && !handle->final_q.empty() ? handle->final_q.data() : nullptr;
}

int mr_task_rollout_copy_final_scene_states(const MRTaskRolloutHandle* handle, float* out, size_t count) {
    return translate([&]() {
        require(handle != nullptr && out != nullptr, "null argument");
        require(!handle->final_scene.empty(), "no final scene state (enable state readback and advance first)");
        require_count(count, handle->final_scene.size() * 13u, "scene states");
        for (size_t i = 0; i < handle->final_scene.size(); ++i) {
            const MRBodyStateC& s = handle->final_scene[i];
            float* o = out + i * 13u;
            o[0] = s.position.x; o[1] = s.position.y; o[2] = s.position.z;
            o[3] = s.orientation.x; o[4] = s.orientation.y; o[5] = s.orientation.z; o[6] = s.orientation.w;
            o[7] = s.linear_velocity_inv_mass.x; o[8] = s.linear_velocity_inv_mass.y; o[9] = s.linear_velocity_inv_mass.z;
            o[10] = s.angular_velocity.x; o[11] = s.angular_velocity.y; o[12] = s.angular_velocity.z;
        }
    });
}

// ------------------------------------------------------- warm-start MLP

namespace {
void require_weights(const MRMlpWeightsC* w) {
    require(w != nullptr, "null MLP weights");
    require(w->w1 && w->b1 && w->w2 && w->b2 && w->w3 && w->b3 && w->in_shift && w->in_scale && w->out_scale,
            "MLP weights: every pointer must be set");
}
} // namespace

int mr_mlp_forward(const MRMlpWeightsC* weights, const float* x, uint32_t rows, float* y_hat,
                   const MRMlpTraceC* trace, void* stream) {
    return translate([&] {
        require_weights(weights);
        require(x != nullptr && y_hat != nullptr, "mr_mlp_forward: null input or output");
        require(rows > 0u, "mr_mlp_forward: rows must be positive");
        if (trace != nullptr)
            require(trace->xq && trace->z1 && trace->h1 && trace->z2 && trace->h2,
                    "mr_mlp_forward: a trace needs every buffer");
        mr::hip_check(mr::mlp_forward(*weights, x, rows, y_hat, trace, static_cast<hipStream_t>(stream)),
                  "mr_mlp_forward launch");
    });
}

uint32_t mr_mlp_loss_scratch_floats(uint32_t rows) { return mr::mlp_loss_scratch_floats(rows); }

int mr_mlp_loss(const float* y_hat, const float* target, const float* out_scale, uint32_t rows, float huber_delta,
                float* dy, float* loss, float* scratch, void* stream) {
    return translate([&] {
        require(y_hat && target && out_scale && dy && loss && scratch, "mr_mlp_loss: null pointer");
        require(rows > 0u && rows <= 0x7fffffffu / MR_MLP_OUTPUTS, "mr_mlp_loss: rows out of range");
        require(huber_delta >= 0.0f && std::isfinite(huber_delta), "mr_mlp_loss: huber_delta must be >= 0");
        mr::hip_check(mr::mlp_loss(y_hat, target, out_scale, rows, huber_delta, dy, loss, scratch,
                                   static_cast<hipStream_t>(stream)),
                  "mr_mlp_loss launch");
    });
}

int mr_mlp_backward(const MRMlpWeightsC* weights, const MRMlpTraceC* trace, const float* dy, uint32_t rows,
                    float* dz2, float* dz1, const MRMlpGradsC* grads, void* stream) {
    return translate([&] {
        require_weights(weights);
        require(trace != nullptr && trace->xq && trace->z1 && trace->h1 && trace->z2 && trace->h2,
                "mr_mlp_backward: a trace needs every buffer");
        require(grads != nullptr && grads->w1 && grads->b1 && grads->w2 && grads->b2 && grads->w3 && grads->b3,
                "mr_mlp_backward: every gradient buffer must be set");
        require(dy && dz2 && dz1, "mr_mlp_backward: null pointer");
        require(rows > 0u && rows <= 65535u * 16u, "mr_mlp_backward: rows must be in 1..1048560");
        mr::hip_check(mr::mlp_backward(*weights, *trace, dy, rows, dz2, dz1, *grads, static_cast<hipStream_t>(stream)),
                  "mr_mlp_backward launch");
    });
}

int mr_mlp_adam(float* param, uint16_t* param_bf16, const float* grad, float* m1, float* m2, uint32_t count,
                float lr, float beta1, float beta2, float eps, uint32_t step, void* stream) {
    return translate([&] {
        require(param && grad && m1 && m2, "mr_mlp_adam: null pointer");
        require(count > 0u && step >= 1u, "mr_mlp_adam: count and step must be positive");
        require(beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps > 0.0f && lr >= 0.0f,
                "mr_mlp_adam: hyper-parameters out of range");
        mr::hip_check(mr::mlp_adam(param, param_bf16, grad, m1, m2, count, lr, beta1, beta2, eps, step,
                               static_cast<hipStream_t>(stream)),
                  "mr_mlp_adam launch");
    });
}

/* ------------------------------------------------- locomotion task --- */

MRLocomotionTaskHandle* mr_locomotion_task_create(const MRLocomotionTaskProgramC* program, MRWorldHandle* world) {
    return construct<MRLocomotionTaskHandle>([&]() {
        require(program != nullptr && world != nullptr, "null argument");
        auto* h = new MRLocomotionTaskHandle();
        try {
            h->task = std::make_unique<mr::LocomotionTask>(*program, *world->world);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    });
}

void mr_locomotion_task_destroy(MRLocomotionTaskHandle* task) { delete task; }

int mr_locomotion_task_reset(MRLocomotionTaskHandle* task) {
    return translate([&]() {
        require(task != nullptr, "null argument");
        task->task->reset();
    });
}

int mr_locomotion_task_step(MRLocomotionTaskHandle* task, const float* device_actions) {
    return translate([&]() {
        require(task != nullptr, "null argument");
        task->task->step(device_actions);
    });
}

int mr_locomotion_task_read(MRLocomotionTaskHandle* task, float* observations, size_t observation_count,
                            MRTaskTransitionC* transitions, size_t transition_count, MRTaskStateC* states,
                            size_t state_count) {
    return translate([&]() {
        require(task != nullptr, "null argument");
        const mr::LocomotionTask& t = *task->task;
        const size_t E = t.environments();
        if (observations != nullptr) require_count(observation_count, E * t.observation_floats(), "observations");
        if (transitions != nullptr) require_count(transition_count, E, "transitions");
        if (states != nullptr) require_count(state_count, E, "task states");
        t.read(observations, transitions, states);
    });
}

const float* mr_locomotion_task_device_observations(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->device_observations() : nullptr;
}

const MRTaskTransitionC* mr_locomotion_task_device_transitions(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->device_transitions() : nullptr;
}

const float* mr_locomotion_task_device_targets(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->device_targets() : nullptr;
}

int mr_policy_dense_layer(const float* weights, const float* bias, const float* mean, const float* inverse_std,
                          float observation_clip, const float* input, uint32_t input_stride, float* output,
                          uint32_t output_stride, uint32_t environments, uint32_t input_width, uint32_t output_width,
                          uint32_t activation, void* stream) {
    return translate([&]() {
        require(weights && bias && input && output, "mr_policy_dense_layer: null pointer");
        require((mean == nullptr) == (inverse_std == nullptr), "mr_policy_dense_layer: mean and inverse_std go together");
        require(input_width > 0u && output_width > 0u, "mr_policy_dense_layer: widths must be positive");
        require(input_stride >= input_width && output_stride >= output_width, "mr_policy_dense_layer: strides below widths");
        require(activation <= MR_POLICY_ACTIVATION_SILU, "mr_policy_dense_layer: unknown activation");
        require(mean == nullptr || observation_clip > 0.0f, "mr_policy_dense_layer: observation_clip must be positive");
        mr::hip_check(mr::launch_policy_dense(weights, bias, mean, inverse_std, observation_clip, input, input_stride,
                                              output, output_stride, environments, input_width, output_width,
                                              activation, static_cast<hipStream_t>(stream)),
                      "mr_policy_dense_layer launch");
    });
}

int mr_policy_sample_and_score(const float* actor_mean, uint32_t mean_stride, const float* action_bias,
                               const float* action_scale, const float* log_std, float action_clip, uint64_t seed,
                               uint64_t revision, const MRTaskStateC* task_states, uint32_t environments,
                               uint32_t actions, uint32_t control_step, float* actions_out, float* latents_out,
                               float* log_probabilities_out, void* stream) {
    return translate([&]() {
        require(actor_mean && action_bias && action_scale && task_states && actions_out && latents_out &&
                    log_probabilities_out,
                "mr_policy_sample_and_score: null pointer");
        require(actions > 0u && mean_stride >= actions, "mr_policy_sample_and_score: action width");
        require(action_clip > 0.0f, "mr_policy_sample_and_score: action_clip must be positive");
        mr::hip_check(mr::launch_policy_sample(actor_mean, mean_stride, action_bias, action_scale, log_std, action_clip,
                                               seed, revision, task_states, environments, actions, control_step,
                                               actions_out, latents_out, log_probabilities_out,
                                               static_cast<hipStream_t>(stream)),
                      "mr_policy_sample_and_score launch");
    });
}

// ---- learning packs and compiled policies ----

uint64_t mr_learning_pack_content_hash(const void* payload, size_t byte_count) {
    if (payload == nullptr && byte_count != 0u) return 0u;
    return mr::learning_pack_hash(payload, byte_count);
}

int mr_write_policy_pack(const MRPolicyPackC* policy, const char* policy_pack_path) {
    return translate([&]() {
        require(policy != nullptr && policy_pack_path != nullptr && policy_pack_path[0] != '\0',
                "PolicyPack and output path are required");
        (void)mr::write_policy_pack(mr::policy_pack_from_c(*policy), policy_pack_path);
    });
}

MRPolicyPackHandle* mr_read_policy_pack(const char* policy_pack_path) {
    return construct<MRPolicyPackHandle>([&]() {
        require(policy_pack_path != nullptr && policy_pack_path[0] != '\0', "PolicyPack path is empty");
        auto h = std::make_unique<MRPolicyPackHandle>();
        h->pack = mr::read_policy_pack(policy_pack_path);
        const mr::PolicyPack& p = h->pack;
        const auto layers = [](const std::vector<mr::DenseLayer>& in, std::vector<MRPolicyDenseLayerC>& out) {
            out.resize(in.size());
            for (size_t i = 0; i < in.size(); ++i)
                out[i] = MRPolicyDenseLayerC{in[i].input_count, in[i].output_count, in[i].activation,
                                             in[i].weights.data(), in[i].weights.size(), in[i].bias.data(),
                                             in[i].bias.size()};
        };
        layers(p.layers, h->layers);
        layers(p.critic_layers, h->critic_layers);
        MRPolicyPackC& v = h->view;
        v.id = p.id.c_str();
        v.revision = p.revision;
        v.contract_version = p.contract[0];
        v.world_fingerprint = p.contract[1];
        v.task_fingerprint = p.contract[2];
        v.observation_fingerprint = p.contract[3];
        v.action_fingerprint = p.contract[4];
        v.observation_mean = p.observation_mean.data();
        v.observation_mean_count = p.observation_mean.size();
        v.observation_inverse_standard_deviation = p.observation_inverse_std.data();
        v.observation_inverse_standard_deviation_count = p.observation_inverse_std.size();
        v.layers = h->layers.data();
        v.layer_count = h->layers.size();
        v.critic_observation_mean = p.critic_observation_mean.data();
        v.critic_observation_mean_count = p.critic_observation_mean.size();
        v.critic_observation_inverse_standard_deviation = p.critic_observation_inverse_std.data();
        v.critic_observation_inverse_standard_deviation_count = p.critic_observation_inverse_std.size();
        v.critic_layers = h->critic_layers.data();
        v.critic_layer_count = h->critic_layers.size();
        v.action_log_standard_deviation = p.action_log_std.data();
        v.action_log_standard_deviation_count = p.action_log_std.size();
        v.action_bias = p.action_bias.data();
        v.action_bias_count = p.action_bias.size();
        v.action_scale = p.action_scale.data();
        v.action_scale_count = p.action_scale.size();
        v.observation_clip = p.observation_clip;
        v.action_clip = p.action_clip;
        return h.release();
    });
}

const MRPolicyPackC* mr_policy_pack_view(const MRPolicyPackHandle* pack) {
    return pack != nullptr ? &pack->view : nullptr;
}

uint64_t mr_policy_pack_content_hash(const MRPolicyPackHandle* pack) {
    return pack != nullptr ? pack->pack.content_hash : 0u;
}

uint32_t mr_policy_pack_format_version(const MRPolicyPackHandle* pack) {
    return pack != nullptr ? pack->pack.format_version : 0u;
}

void mr_policy_pack_destroy(MRPolicyPackHandle* pack) { delete pack; }

int mr_write_policy_rollout_pack(const MRPolicyRolloutPackC* pack, const char* output_path) {
    return translate([&]() {
        require(pack != nullptr && output_path != nullptr && output_path[0] != '\0',
                "PolicyRolloutPack and output path are required");
        (void)mr::write_policy_rollout_pack(*pack, output_path);
    });
}

MRPolicyHandle* mr_policy_create(const MRPolicyPackC* policy, uint32_t environments) {
    return construct<MRPolicyHandle>([&]() {
        require(policy != nullptr, "null policy");
        auto h = std::make_unique<MRPolicyHandle>();
        h->policy = std::make_unique<mr::Policy>(mr::policy_pack_from_c(*policy), environments);
        return h.release();
    });
}

void mr_policy_destroy(MRPolicyHandle* policy) { delete policy; }

uint64_t mr_policy_revision(const MRPolicyHandle* policy) {
    return policy != nullptr ? policy->policy->revision() : 0u;
}

uint32_t mr_policy_actor_input_count(const MRPolicyHandle* policy) {
    return policy != nullptr ? policy->policy->actor_input() : 0u;
}

uint32_t mr_policy_critic_input_count(const MRPolicyHandle* policy) {
    return policy != nullptr ? policy->policy->critic_input() : 0u;
}

uint32_t mr_policy_action_count(const MRPolicyHandle* policy) {
    return policy != nullptr ? policy->policy->actions() : 0u;
}

int mr_policy_act(MRPolicyHandle* policy, const float* actor_observations, const float* critic_observations,
                  const MRTaskStateC* task_states, uint32_t control_step, uint64_t seed, float* actions,
                  float* latents, float* log_probabilities, float* values, void* stream) {
    return translate([&]() {
        require(policy != nullptr, "null policy");
        policy->policy->act(actor_observations, critic_observations, task_states, control_step, seed, actions,
                            latents, log_probabilities, values, static_cast<hipStream_t>(stream));
    });
}

int mr_locomotion_task_set_policy(MRLocomotionTaskHandle* task, const MRPolicyPackC* policy) {
    return translate([&]() {
        require(task != nullptr, "null task");
        require(policy != nullptr, "task policy is null");
        task->task->set_policy(mr::policy_pack_from_c(*policy));
    });
}

int mr_locomotion_task_load_policy_pack(MRLocomotionTaskHandle* task, const char* policy_pack_path) {
    return translate([&]() {
        require(task != nullptr, "null task");
        require(policy_pack_path != nullptr && policy_pack_path[0] != '\0', "PolicyPack path is empty");
        task->task->set_policy(mr::read_policy_pack(policy_pack_path));
    });
}

uint64_t mr_locomotion_task_policy_revision(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->policy_revision() : 0u;
}

int mr_locomotion_task_clear_policy(MRLocomotionTaskHandle* task) {
    return translate([&]() {
        require(task != nullptr, "null task");
        task->task->clear_policy();
    });
}

const float* mr_locomotion_task_device_policy_actions(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->device_policy_actions() : nullptr;
}

const float* mr_locomotion_task_device_policy_latents(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->device_policy_latents() : nullptr;
}

const float* mr_locomotion_task_device_policy_log_probabilities(const MRLocomotionTaskHandle* task) {
    return task != nullptr ? task->task->device_policy_log_probabilities() : nullptr;
}

int mr_locomotion_task_read_policy_outputs(MRLocomotionTaskHandle* task, float* actions, size_t action_count,
                                           float* latents, size_t latent_count, float* log_probabilities,
                                           size_t log_probability_count) {
    return translate([&]() {
        require(task != nullptr, "null task");
        task->task->read_policy_outputs(actions, action_count, latents, latent_count, log_probabilities,
                                        log_probability_count);
    });
}

int mr_locomotion_task_write_policy_rollout_pack(const MRLocomotionTaskHandle* task,
                                                 const MRPolicyRolloutBatchC* batch, const char* batch_id,
                                                 const char* output_path) {
    return translate([&]() {
        require(task != nullptr, "null task");
        require(batch != nullptr && batch_id != nullptr && batch_id[0] != '\0' && output_path != nullptr &&
                    output_path[0] != '\0',
                "policy rollout batch, identity, and output path are required");
        task->task->write_policy_rollout_pack(*batch, batch_id, output_path);
    });
}

} // extern "C"
