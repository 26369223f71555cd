"""Thin ctypes binding for the MI355X temporalCone C ABI (include/mr_c_api.h).

Conventions follow the reference binding (python/metalrobo/native.py:16-20,
225-264, 311-313, 589-590): the library is resolved from ``METALROBO_LIBRARY``
or the in-tree build, ``argtypes``/``restype`` are declared once per loaded
library, native failures raise :class:`MetalRoboError` carrying
``mr_last_error()``, and C structs are mirrored by ctypes ``Structure``s with
size checks. PyTorch tensors are accepted only at this edge, as device
inputs already resident in HBM.
"""

from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path
from typing import Final, Optional

import numpy as np

_LIBRARY_ENV: Final = "METALROBO_LIBRARY"

MR_INVALID_INDEX: Final = 0xFFFFFFFF
MR_MANIFOLD_POINT_CAPACITY: Final = 4

MR_SYNTHETIC_CHAIN_BOXES: Final = 1
MR_SYNTHETIC_HUMANOID_TERRAIN: Final = 2
MR_SYNTHETIC_RANDOM_TREES: Final = 3
MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES: Final = 100
MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN: Final = 101
MR_RUN_SOURCE_SYNTHETIC_RANDOM_TREES: Final = 102
MR_LOCOMOTION_SURFACE_GROUND: Final = 0
MR_LOCOMOTION_SURFACE_TERRAIN: Final = 1

MR_ACTUATION_EFFORT: Final = 0
MR_ACTUATION_IMPLICIT_POSITION_DRIVE: Final = 1
MR_SOLVER_ORDERED_SWEEPS: Final = 0
MR_SOLVER_RELAXED_JACOBI: Final = 1

STATUS_NAMES: Final = {
    0: "success", 1: "fixed_budget_complete", 2: "nonfinite_input", 3: "nonfinite_result",
    4: "pair_capacity_overflow", 5: "contact_capacity_overflow", 6: "manifold_capacity_overflow",
    7: "constraint_capacity_overflow", 8: "island_capacity_overflow", 9: "ccd_event_budget_exhausted",
    10: "factorization_failed", 11: "did_not_converge", 12: "unsupported", 13: "ccd_capacity_overflow",
}


class MetalRoboError(RuntimeError):
    """Raised when the native runtime reports an error."""


# ----------------------------------------------------------------- records

class _F4(ct.Structure):
    _fields_ = [("x", ct.c_float), ("y", ct.c_float), ("z", ct.c_float), ("w", ct.c_float)]


class HeightfieldC(ct.Structure):
    _fields_ = [("nx", ct.c_uint32), ("ny", ct.c_uint32), ("present", ct.c_uint32), ("reserved", ct.c_uint32),
                ("origin_spacing", _F4), ("bounds_lower", _F4), ("bounds_upper", _F4)]


class WorldModelDescC(ct.Structure):
    _fields_ = [
        ("model_count", ct.c_uint32), ("environment_count", ct.c_uint32), ("max_nq", ct.c_uint32),
        ("max_nv", ct.c_uint32), ("max_scene_bodies", ct.c_uint32), ("max_shapes", ct.c_uint32),
        ("max_pairs", ct.c_uint32), ("reserved", ct.c_uint32), ("gravity", _F4),
        ("models", ct.c_void_p), ("environment_model", ct.c_void_p), ("bodies", ct.c_void_p),
        ("dofs", ct.c_void_p), ("shapes", ct.c_void_p), ("materials", ct.c_void_p), ("pairs", ct.c_void_p),
        ("body_total", ct.c_uint32), ("dof_total", ct.c_uint32), ("shape_total", ct.c_uint32),
        ("material_total", ct.c_uint32), ("pair_total", ct.c_uint32), ("heightfield_vertex_count", ct.c_uint32),
        ("heightfield", HeightfieldC), ("heightfield_heights", ct.c_void_p),
        ("actuators", ct.c_void_p), ("actuator_total", ct.c_uint32), ("reserved_actuators", ct.c_uint32),
        ("geometries", ct.c_void_p), ("geometry_vertices", ct.c_void_p), ("triangles", ct.c_void_p),
        ("bvh_nodes", ct.c_void_p), ("geometry_total", ct.c_uint32), ("geometry_vertex_total", ct.c_uint32),
        ("triangle_total", ct.c_uint32), ("bvh_node_total", ct.c_uint32),
    ]


class WorldStepConfigC(ct.Structure):
    _fields_ = [
        ("control_timestep", ct.c_float), ("physics_substeps", ct.c_uint32),
        ("velocity_iterations", ct.c_uint32), ("final_velocity_iterations", ct.c_uint32),
        ("actuation_mode", ct.c_uint32), ("warm_start", ct.c_uint32), ("apply_body_damping", ct.c_uint32),
        ("capture_evidence", ct.c_uint32), ("capture_substep", ct.c_uint32), ("pair_capacity", ct.c_uint32),
        ("manifold_capacity", ct.c_uint32), ("constraint_capacity", ct.c_uint32),
        ("manifold_breaking_separation", ct.c_float), ("manifold_breaking_tangential", ct.c_float),
        ("manifold_merge_distance", ct.c_float), ("manifold_normal_cosine", ct.c_float),
        ("contact_slop", ct.c_float), ("max_depenetration_velocity", ct.c_float),
        ("residual_tolerance", ct.c_float), ("max_iterations", ct.c_uint32),
        ("solver", ct.c_uint32), ("maximum_normal_impulse", ct.c_float),
        ("ccd_mode", ct.c_uint32), ("max_ccd_events", ct.c_uint32), ("max_ccd_iterations", ct.c_uint32),
        ("ccd_reserved", ct.c_uint32), ("ccd_distance_tolerance", ct.c_float), ("ccd_time_tolerance", ct.c_float),
        ("ccd_minimum_advance", ct.c_float), ("ccd_reserved1", ct.c_float),
    ]

    def copy(self) -> "WorldStepConfigC":
        out = WorldStepConfigC()
        ct.memmove(ct.byref(out), ct.byref(self), ct.sizeof(self))
        return out


class SyntheticWorldConfigC(ct.Structure):
    _fields_ = [("kind", ct.c_uint32), ("environment_count", ct.c_uint32), ("seed", ct.c_uint64)]


class WorldAdvanceC(ct.Structure):
    _fields_ = [
        ("control_step_count", ct.c_uint32), ("successful_environment_steps", ct.c_uint32),
        ("failed_environment_steps", ct.c_uint32), ("first_failing_environment", ct.c_uint32),
        ("first_failing_control_step", ct.c_uint32), ("first_gpu_status_code", ct.c_uint32),
        ("maximum_active_pairs", ct.c_uint32), ("maximum_constraints", ct.c_uint32),
        ("gpu_milliseconds", ct.c_double),
    ]


class WorldLayoutC(ct.Structure):
    _fields_ = [
        ("environment_count", ct.c_uint32), ("max_nq", ct.c_uint32), ("max_nv", ct.c_uint32),
        ("max_scene_bodies", ct.c_uint32), ("pair_capacity", ct.c_uint32), ("manifold_capacity", ct.c_uint32),
        ("constraint_capacity", ct.c_uint32), ("lds_bytes_per_environment", ct.c_uint32),
        ("retained_device_bytes", ct.c_uint64),
    ]


class TaskRolloutDynamicSphereC(ct.Structure):
    """c_api.h:55-61 (the reference's layout; not supported by this library's sources)."""
    _fields_ = [("position", ct.c_float * 3), ("linear_velocity", ct.c_float * 3), ("radius", ct.c_float),
                ("mass", ct.c_float), ("launch_step", ct.c_uint32)]


class TaskRolloutConfigC(ct.Structure):
    """c_api.h:62-88, byte for byte."""
    _fields_ = [
        ("environment_count", ct.c_uint32), ("physics_substeps", ct.c_uint32),
        ("velocity_iterations", ct.c_uint32), ("final_velocity_iterations", ct.c_uint32),
        ("control_timestep_seconds", ct.c_float), ("seed", ct.c_uint64),
        ("dynamic_spheres", ct.c_void_p), ("dynamic_sphere_count", ct.c_uint32),
        ("disable_task_terminations", ct.c_uint32), ("interaction_reference_mode", ct.c_uint32),
        ("interaction_student_authority", ct.c_float), ("override_interaction_student_authority", ct.c_uint32),
        ("interaction_reset_phase_fraction", ct.c_float), ("override_interaction_reset_phase_fraction", ct.c_uint32),
        ("materialize_articulated_contact_responses", ct.c_uint32), ("override_difficulty_band_range", ct.c_uint32),
        ("minimum_difficulty_band", ct.c_uint32), ("maximum_difficulty_band", ct.c_uint32),
        ("interaction_reset_phase_probability", ct.c_float),
        ("override_interaction_reset_phase_probability", ct.c_uint32),
        ("interaction_reset_maximum_phase", ct.c_float), ("override_interaction_reset_maximum_phase", ct.c_uint32),
    ]


class RunManifestC(ct.Structure):
    """c_api.h:171-190, byte for byte."""
    _fields_ = [("profile", TaskRolloutConfigC), ("source", ct.c_uint32), ("surface", ct.c_uint32),
                ("task", ct.c_uint32)] + [(n, ct.c_char_p) for n in (
                    "urdf_path", "srdf_path", "world_pack_path", "task_pack_path", "robot_actuator_pack_path",
                    "sensor_program_pack_path", "reality_program_pack_path", "teacher_pack_path",
                    "teacher_clip_id")] + [("visual_sensor_program", ct.c_void_p),
                                           ("inspection_visual_program", ct.c_void_p),
                                           ("metallib_path", ct.c_char_p)]


class TaskRolloutLayoutC(ct.Structure):
    """c_api.h:128-158, byte for byte."""
    _fields_ = [(n, ct.c_uint32) for n in (
        "environment_count", "nq", "nv", "action_count", "actor_observation_count", "critic_observation_count",
        "scene_body_count", "motion_feature_count", "maximum_episode_steps")] + [(n, ct.c_uint64) for n in (
            "world_fingerprint", "task_fingerprint", "observation_fingerprint", "action_fingerprint",
            "run_fingerprint", "robot_fingerprint", "sensor_fingerprint", "reality_fingerprint",
            "teacher_fingerprint", "submitted_control_steps", "completed_environment_steps",
            "submission_count")] + [(n, ct.c_size_t) for n in (
                "retained_buffer_bytes", "immutable_private_bytes", "persistent_state_private_bytes",
                "transient_private_bytes", "shared_boundary_bytes", "peak_aliased_bytes")] + [
        ("total_gpu_milliseconds", ct.c_double), ("total_submission_milliseconds", ct.c_double)]


class TaskRolloutStageHighWaterC(ct.Structure):
    """c_api.h:192-220, byte for byte."""
    _fields_ = [(n, ct.c_uint32) for n in (
        "candidate_pairs", "raw_contacts", "manifolds", "constraint_blocks", "constraint_rows", "islands",
        "hard_convex_pairs", "mesh_triangle_candidates", "solver_tiles", "spill_rows", "ccd_candidates",
        "ccd_events", "endpoint_runtime_records", "articulation_point_queries", "rod_candidate_pairs",
        "rod_raw_contacts", "rod_manifolds", "rod_ccd_events", "quality_generalized_velocities", "quality_rows",
        "quality_krylov_vectors", "quality_direct_tiles", "dynamic_nodes", "island_node_references",
        "island_constraint_references", "rod_factor_blocks", "operator_velocity_elements")]


class MlpWeightsC(ct.Structure):
    """Mirror of MRMlpWeightsC (include/mr_mlp.h): device pointers."""
    _fields_ = [("w1", ct.c_void_p), ("b1", ct.c_void_p), ("w2", ct.c_void_p), ("b2", ct.c_void_p),
                ("w3", ct.c_void_p), ("b3", ct.c_void_p), ("in_shift", ct.c_void_p), ("in_scale", ct.c_void_p),
                ("out_scale", ct.c_void_p), ("min_correction", ct.c_float),
                ("max_normal_residual", ct.c_float)]


class MlpTraceC(ct.Structure):
    """Mirror of MRMlpTraceC."""
    _fields_ = [("xq", ct.c_void_p), ("z1", ct.c_void_p), ("h1", ct.c_void_p), ("z2", ct.c_void_p),
                ("h2", ct.c_void_p)]


class MlpGradsC(ct.Structure):
    """Mirror of MRMlpGradsC."""
    _fields_ = [("w1", ct.c_void_p), ("b1", ct.c_void_p), ("w2", ct.c_void_p), ("b2", ct.c_void_p),
                ("w3", ct.c_void_p), ("b3", ct.c_void_p)]


class TaskRolloutAdvanceC(ct.Structure):
    """c_api.h:222-235, byte for byte."""
    _fields_ = [
        ("control_step_count", ct.c_uint32), ("successful_environment_steps", ct.c_uint32),
        ("failed_environment_steps", ct.c_uint32), ("first_failing_environment", ct.c_uint32),
        ("first_failing_control_step", ct.c_uint32), ("first_gpu_status_code", ct.c_uint32),
        ("host_requested_resets", ct.c_uint32), ("maximum_active_contacts", ct.c_uint32),
        ("maximum_manifolds", ct.c_uint32), ("high_water", TaskRolloutStageHighWaterC),
        ("gpu_milliseconds", ct.c_double), ("submission_milliseconds", ct.c_double),
    ]


class LocomotionTaskProgramC(ct.Structure):
    """MRLocomotionTaskProgramC (include/mr_task.h)."""
    _fields_ = [
        ("actions", ct.c_void_p), ("observations", ct.c_void_p), ("rewards", ct.c_void_p),
        ("terminations", ct.c_void_p), ("default_q", ct.c_void_p), ("action_count", ct.c_uint32),
        ("observation_count", ct.c_uint32), ("reward_count", ct.c_uint32), ("termination_count", ct.c_uint32),
        ("history_length", ct.c_uint32), ("delay_slots", ct.c_uint32), ("max_episode_steps", ct.c_uint32),
        ("flags", ct.c_uint32), ("root_q_offset", ct.c_uint32), ("root_v_offset", ct.c_uint32),
        ("actuator_delay", ct.c_uint32), ("reserved", ct.c_uint32), ("root_reference", ct.c_float * 4),
        ("base_height_target", ct.c_float), ("gait_period", ct.c_float), ("command_lower", ct.c_float * 4),
        ("command_upper", ct.c_float * 4), ("command_seconds", ct.c_float * 2), ("seed", ct.c_uint64),
    ]


class PolicyDenseLayerC(ct.Structure):
    """MRPolicyDenseLayerC (include/mr_packs.h, c_api.h:334-342)."""
    _fields_ = [("input_count", ct.c_uint32), ("output_count", ct.c_uint32), ("activation", ct.c_uint32),
                ("weights", ct.c_void_p), ("weight_count", ct.c_size_t), ("bias", ct.c_void_p),
                ("bias_count", ct.c_size_t)]


_POLICY_TABLES = (("observation_mean", "observation_mean_count"),
                  ("observation_inverse_standard_deviation", "observation_inverse_standard_deviation_count"),
                  ("layers", "layer_count"), ("critic_observation_mean", "critic_observation_mean_count"),
                  ("critic_observation_inverse_standard_deviation",
                   "critic_observation_inverse_standard_deviation_count"),
                  ("critic_layers", "critic_layer_count"),
                  ("action_log_standard_deviation", "action_log_standard_deviation_count"),
                  ("action_bias", "action_bias_count"), ("action_scale", "action_scale_count"))


class PolicyPackC(ct.Structure):
    """MRPolicyPackC (include/mr_packs.h, c_api.h:344-372)."""
    _fields_ = ([("id", ct.c_char_p), ("revision", ct.c_uint64), ("contract_version", ct.c_uint64),
                 ("world_fingerprint", ct.c_uint64), ("task_fingerprint", ct.c_uint64),
                 ("observation_fingerprint", ct.c_uint64), ("action_fingerprint", ct.c_uint64)]
                + [f for table, count in _POLICY_TABLES for f in ((table, ct.c_void_p), (count, ct.c_size_t))]
                + [("observation_clip", ct.c_float), ("action_clip", ct.c_float)])


_ROLLOUT_BATCH_TABLES = (("actor_observations", "actor_observation_count"),
                         ("critic_observations", "critic_observation_count"),
                         ("motion_features", "motion_feature_count"), ("teacher_actions", "teacher_action_count"),
                         ("latents", "latent_count"), ("log_probabilities", "log_probability_count"),
                         ("values", "value_count"), ("bootstrap_values", "bootstrap_value_count"),
                         ("transitions", "transition_count"))


class PolicyRolloutBatchC(ct.Structure):
    """MRPolicyRolloutBatchC (include/mr_packs.h, c_api.h:374-394)."""
    _fields_ = [("control_step_count", ct.c_uint32)] + [
        fld for table, count in _ROLLOUT_BATCH_TABLES for fld in ((table, ct.c_void_p), (count, ct.c_size_t))]


class PolicyOutcomeDescriptorC(ct.Structure):
    """MRPolicyOutcomeDescriptorC (include/mr_packs.h)."""
    _fields_ = [("id", ct.c_char_p), ("unit", ct.c_char_p), ("direction", ct.c_uint32)]


class PolicyRolloutPackC(ct.Structure):
    """MRPolicyRolloutPackC (include/mr_packs.h)."""
    _fields_ = [("id", ct.c_char_p), ("task_fingerprint", ct.c_uint64), ("policy_fingerprint", ct.c_uint64),
                ("policy_revision", ct.c_uint64)] + [(n, ct.c_uint32) for n in (
                    "environment_count", "control_step_count", "actor_observation_count", "critic_observation_count",
                    "action_count", "motion_feature_count")] + [
                ("actor_observations", ct.c_void_p), ("actor_observation_count_total", ct.c_size_t),
                ("critic_observations", ct.c_void_p), ("critic_observation_count_total", ct.c_size_t),
                ("motion_features", ct.c_void_p), ("motion_feature_count_total", ct.c_size_t),
                ("teacher_actions", ct.c_void_p), ("teacher_action_count", ct.c_size_t),
                ("latents", ct.c_void_p), ("latent_count", ct.c_size_t),
                ("log_probabilities", ct.c_void_p), ("log_probability_count", ct.c_size_t),
                ("values", ct.c_void_p), ("value_count", ct.c_size_t),
                ("bootstrap_values", ct.c_void_p), ("bootstrap_value_count", ct.c_size_t),
                ("outcomes", ct.c_void_p), ("outcome_count", ct.c_size_t),
                ("outcome_values", ct.c_void_p), ("outcome_value_count", ct.c_size_t),
                ("transitions", ct.c_void_p), ("transition_count", ct.c_size_t)]


assert ct.sizeof(PolicyDenseLayerC) == 48 and ct.sizeof(PolicyPackC) == 208 and ct.sizeof(PolicyRolloutBatchC) == 152


# numpy mirrors of the bulk records (include/mr_world_types.h)
F4 = ("<f4", (4,))
MODEL_HEADER_DTYPE: Final = np.dtype([(n, "<u4") for n in (
    "body_offset", "body_count", "link_count", "root_type", "dof_offset", "nq", "nv", "shape_offset",
    "shape_count", "pair_offset", "pair_count", "level_count", "scene_body_count", "material_offset",
    "material_count", "reserved")])
BODY_DTYPE: Final = np.dtype([
    ("parent", "<u4"), ("joint_type", "<u4"), ("q_index", "<u4"), ("v_index", "<u4"), ("motion_type", "<u4"),
    ("depth", "<u4"), ("flags", "<u4"), ("reserved", "<u4"), ("mass", *F4), ("inertia", "<f4", (3, 4)),
    ("inverse_inertia", "<f4", (3, 4)), ("axis", *F4), ("parent_anchor", *F4), ("child_anchor", *F4),
    ("parent_rotation", *F4), ("child_rotation", *F4), ("damping", *F4), ("damping_scale", *F4)])
DOF_DTYPE: Final = np.dtype([("body", "<u4"), ("q_index", "<u4"), ("flags", "<u4"), ("reserved", "<u4"),
                             ("limits", *F4), ("drive", *F4)])
SHAPE_DTYPE: Final = np.dtype([
    ("body", "<u4"), ("shape_type", "<u4"), ("material", "<u4"), ("flags", "<u4"), ("collision_group", "<u4"),
    ("collision_mask", "<u4"), ("slot_generation", "<u4"), ("geometry", "<u4"), ("local_position", *F4),
    ("local_rotation", *F4), ("dimensions", *F4), ("contact_rest_radius", *F4)])
MATERIAL_DTYPE: Final = np.dtype([("friction", *F4), ("response", *F4)])
GEOMETRY_DTYPE: Final = np.dtype([(n, "<u4") for n in (
    "kind", "flags", "vertex_offset", "vertex_count", "triangle_offset", "triangle_count", "node_offset",
    "node_count")] + [("local_lower", *F4), ("local_upper", *F4)])
TRIANGLE_DTYPE: Final = np.dtype([("vertices", "<u4", (3,)), ("feature", "<u4")])
BVH_NODE_DTYPE: Final = np.dtype([(n, *F4) for n in ("lower_x", "lower_y", "lower_z", "upper_x", "upper_y",
                                                     "upper_z")] + [("child", "<u4", (4,)), ("count", "<u4", (4,))])
assert GEOMETRY_DTYPE.itemsize == 64 and TRIANGLE_DTYPE.itemsize == 16 and BVH_NODE_DTYPE.itemsize == 128
PAIR_DTYPE: Final = np.dtype([("collider_a", "<u4"), ("collider_b", "<u4"), ("pair_class", "<u4"),
                              ("reserved", "<u4")])
BODY_STATE_DTYPE: Final = np.dtype([
    ("position", *F4), ("orientation", *F4), ("linear_velocity_inv_mass", *F4), ("angular_velocity", *F4),
    ("inverse_inertia_world", "<f4", (3, 4)), ("flags_indices", "<u4", (4,))])
MANIFOLD_HEADER_DTYPE: Final = np.dtype([
    ("pair_and_count", "<u4", (4,)), ("generations_and_flags", "<u4", (4,)), ("normal_and_age", *F4),
    ("tangent_and_metric", *F4)])
MANIFOLD_POINT_DTYPE: Final = np.dtype([
    ("local_anchor_a", *F4), ("local_anchor_b", *F4), ("impulses", *F4), ("feature_and_life", "<u4", (4,))])
WORLD_STATUS_DTYPE: Final = np.dtype([
    ("code", "<u4"), ("environment", "<u4"), ("control_step", "<u4"), ("successful_substeps", "<u4"),
    ("aba_code", "<u4"), ("failing_substep", "<u4"), ("failing_index", "<u4"), ("flags", "<u4"),
    ("diagnostics", *F4)])
CONTACT_STATUS_DTYPE: Final = np.dtype([(n, "<u4") for n in (
    "code", "active_pairs", "raw_contacts", "manifolds", "constraints", "islands", "retained_points",
    "new_points", "required_pairs", "required_manifolds", "required_constraints", "solver_iterations",
    "first_failing_pair", "first_failing_constraint", "max_active_pairs", "max_constraints")] +
    [("residuals", *F4), ("diagnostics", *F4)])
PAIR_RECORD_DTYPE: Final = np.dtype([("environment", "<u4"), ("collider_a", "<u4"), ("collider_b", "<u4"),
                                     ("pair_class", "<u4")])
# MRActuatorProfileC (engine_types.h MRActuatorProfileGPU layout)
ACTUATOR_DTYPE: Final = np.dtype([("motor_and_speed", *F4), ("transmission_and_envelope", *F4),
                                  ("identity", "<u4", (4,))])
MR_ACTUATOR_PROFILE_ACTIVE: Final = 1

# task-program records (include/mr_task.h; task_program_types.h:445-623, byte for byte)
U4 = ("<u4", (4,))
TASK_ACTION_DTYPE: Final = np.dtype([("indices", *U4), ("parameters", *F4), ("drive", *F4), ("actuator", *U4)])
TASK_OBSERVATION_DTYPE: Final = np.dtype([("source", *U4), ("transform", *F4), ("auxiliary", *U4)])
TASK_REWARD_DTYPE: Final = np.dtype([("source", *U4), ("parameters", *F4), ("auxiliary", *F4)])
TASK_TERMINATION_DTYPE: Final = np.dtype([("source", *U4), ("parameters", *F4), ("schedule", *U4)])
TASK_STATE_DTYPE: Final = np.dtype([
    ("episode", *U4), ("schedule", *U4), ("status", *U4), ("command_and_phase", *F4), ("command_extension", *F4),
    ("air_return_tracking", *F4), ("recovery", *F4), ("recovery_stats", *U4), ("threat_geometry", *F4),
    ("threat_teacher", *F4), ("threat_metadata", *U4)])
TASK_TRANSITION_DTYPE: Final = np.dtype([
    ("reward_and_state", *F4), ("termination", *U4), ("reward_breakdown0", *F4), ("reward_breakdown1", *F4),
    ("outcome_channels0", *F4), ("outcome_channels1", *F4), ("policy_revision", "<u8"),
    ("timeout_bootstrap_value", "<f4"), ("episode_tracking_score", "<f4"), ("task_progress", *U4)])
# MRLearningTransitionC (include/mr_packs.h; task_program_types.h:628-637)
LEARNING_TRANSITION_DTYPE: Final = np.dtype([("reward_and_bootstrap", *F4), ("termination", *U4), ("context", *U4),
                                             ("policy_revision", "<u8"), ("reserved", "<u8")])
assert LEARNING_TRANSITION_DTYPE.itemsize == 64
MR_ACTUATOR_PROFILE_CALIBRATED: Final = 2

_EXPECTED_SIZES: Final = {
    MODEL_HEADER_DTYPE: 64, BODY_DTYPE: 256, DOF_DTYPE: 48, ACTUATOR_DTYPE: 48, SHAPE_DTYPE: 96, MATERIAL_DTYPE: 32, PAIR_DTYPE: 16,
    BODY_STATE_DTYPE: 128, MANIFOLD_HEADER_DTYPE: 64, MANIFOLD_POINT_DTYPE: 64, WORLD_STATUS_DTYPE: 48,
    CONTACT_STATUS_DTYPE: 96, PAIR_RECORD_DTYPE: 16, TASK_ACTION_DTYPE: 64, TASK_OBSERVATION_DTYPE: 48,
    TASK_REWARD_DTYPE: 48, TASK_TERMINATION_DTYPE: 48, TASK_STATE_DTYPE: 176, TASK_TRANSITION_DTYPE: 128,
}
for _dt, _size in _EXPECTED_SIZES.items():
    assert _dt.itemsize == _size, (_dt, _dt.itemsize, _size)
assert ct.sizeof(WorldStepConfigC) == 120
assert ct.sizeof(WorldModelDescC) == 264
assert ct.sizeof(WorldAdvanceC) == 40
assert ct.sizeof(WorldLayoutC) == 40
assert ct.sizeof(MlpWeightsC) == 80
assert ct.sizeof(LocomotionTaskProgramC) == 160 and LocomotionTaskProgramC.seed.offset == 152
# the reference's task-rollout records (c_api.h), LP64
assert ct.sizeof(TaskRolloutDynamicSphereC) == 36
assert ct.sizeof(TaskRolloutConfigC) == 104
assert ct.sizeof(RunManifestC) == 216 and RunManifestC.source.offset == 104
assert ct.sizeof(TaskRolloutLayoutC) == 200
assert ct.sizeof(TaskRolloutStageHighWaterC) == 108
assert ct.sizeof(TaskRolloutAdvanceC) == 160 and TaskRolloutAdvanceC.gpu_milliseconds.offset == 144


def _ptr(arr: Optional[np.ndarray]) -> Optional[int]:
    if arr is None:
        return None
    if not arr.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays crossing the C ABI must be C-contiguous")
    return arr.ctypes.data


# ------------------------------------------------------------ library load

def _package_directory() -> Path:
    return Path(__file__).resolve().parent


def resolve_library_path(path: str | os.PathLike[str] | None = None) -> Path:
    """Resolve the native library, preferring explicit and environment paths."""
    if path is not None:
        explicit = Path(path).expanduser()
        if not explicit.is_file():
            raise FileNotFoundError(f"MetalRobo library does not exist: {explicit}")
        return explicit.resolve()
    configured = os.environ.get(_LIBRARY_ENV)
    if configured:
        configured_path = Path(configured).expanduser()
        if not configured_path.is_file():
            raise FileNotFoundError(f"{_LIBRARY_ENV} points to a missing file: {configured_path}")
        return configured_path.resolve()
    default = (_package_directory() / ".." / "libmetalrobo_amd.so").resolve()
    if default.is_file():
        return default
    raise FileNotFoundError(
        "Could not find libmetalrobo_amd.so. Build it with `python numi-lab_amd/build.py` or set "
        f"{_LIBRARY_ENV}. Expected the in-tree build at {default}")


_SIGNATURES: Final = {
    "mr_version": ([], ct.c_char_p),
    "mr_runtime_abi_fingerprint": ([], ct.c_uint64),
    "mr_last_error": ([], ct.c_char_p),
    "mr_world_model_create_synthetic": ([ct.POINTER(SyntheticWorldConfigC)], ct.c_void_p),
    "mr_world_model_create": ([ct.POINTER(WorldModelDescC), ct.c_float], ct.c_void_p),
    "mr_world_model_destroy": ([ct.c_void_p], None),
    "mr_world_model_describe": ([ct.c_void_p, ct.POINTER(WorldModelDescC)], ct.c_int),
    "mr_world_model_initial_state": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t,
                                      ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_model_commands": ([ct.c_void_p, ct.c_uint32, ct.c_uint32, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_model_default_config": ([ct.c_void_p, ct.POINTER(WorldStepConfigC)], ct.c_int),
    "mr_world_create": ([ct.c_void_p, ct.POINTER(WorldStepConfigC)], ct.c_void_p),
    "mr_world_destroy": ([ct.c_void_p], None),
    "mr_world_layout": ([ct.c_void_p, ct.POINTER(WorldLayoutC)], ct.c_int),
    "mr_world_initialize_resident_state": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t,
                                            ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_set_reset_state": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t,
                                  ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_write_manifolds": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_world_advance": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t, ct.c_uint32,
                          ct.POINTER(WorldAdvanceC)], ct.c_int),
    "mr_world_enqueue_device": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_locomotion_task_create": ([ct.POINTER(LocomotionTaskProgramC), ct.c_void_p], ct.c_void_p),
    "mr_locomotion_task_destroy": ([ct.c_void_p], None),
    "mr_locomotion_task_reset": ([ct.c_void_p], ct.c_int),
    "mr_locomotion_task_step": ([ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_locomotion_task_read": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t, ct.c_void_p,
                                 ct.c_size_t], ct.c_int),
    "mr_locomotion_task_device_observations": ([ct.c_void_p], ct.c_void_p),
    "mr_locomotion_task_device_transitions": ([ct.c_void_p], ct.c_void_p),
    "mr_locomotion_task_device_targets": ([ct.c_void_p], ct.c_void_p),
    "mr_policy_dense_layer": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_float, ct.c_void_p,
                               ct.c_uint32, ct.c_void_p, ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_uint32,
                               ct.c_uint32, ct.c_void_p], ct.c_int),
    "mr_policy_sample_and_score": ([ct.c_void_p, ct.c_uint32, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_float,
                                    ct.c_uint64, ct.c_uint64, ct.c_void_p, ct.c_uint32, ct.c_uint32, ct.c_uint32,
                                    ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_learning_pack_content_hash": ([ct.c_void_p, ct.c_size_t], ct.c_uint64),
    "mr_write_policy_pack": ([ct.POINTER(PolicyPackC), ct.c_char_p], ct.c_int),
    "mr_read_policy_pack": ([ct.c_char_p], ct.c_void_p),
    "mr_policy_pack_view": ([ct.c_void_p], ct.POINTER(PolicyPackC)),
    "mr_policy_pack_content_hash": ([ct.c_void_p], ct.c_uint64),
    "mr_policy_pack_format_version": ([ct.c_void_p], ct.c_uint32),
    "mr_policy_pack_destroy": ([ct.c_void_p], None),
    "mr_write_policy_rollout_pack": ([ct.POINTER(PolicyRolloutPackC), ct.c_char_p], ct.c_int),
    "mr_policy_create": ([ct.POINTER(PolicyPackC), ct.c_uint32], ct.c_void_p),
    "mr_policy_destroy": ([ct.c_void_p], None),
    "mr_policy_revision": ([ct.c_void_p], ct.c_uint64),
    "mr_policy_actor_input_count": ([ct.c_void_p], ct.c_uint32),
    "mr_policy_critic_input_count": ([ct.c_void_p], ct.c_uint32),
    "mr_policy_action_count": ([ct.c_void_p], ct.c_uint32),
    "mr_policy_act": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint32, ct.c_uint64, ct.c_void_p,
                       ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_locomotion_task_set_policy": ([ct.c_void_p, ct.POINTER(PolicyPackC)], ct.c_int),
    "mr_locomotion_task_load_policy_pack": ([ct.c_void_p, ct.c_char_p], ct.c_int),
    "mr_locomotion_task_policy_revision": ([ct.c_void_p], ct.c_uint64),
    "mr_locomotion_task_clear_policy": ([ct.c_void_p], ct.c_int),
    "mr_locomotion_task_device_policy_actions": ([ct.c_void_p], ct.c_void_p),
    "mr_locomotion_task_device_policy_latents": ([ct.c_void_p], ct.c_void_p),
    "mr_locomotion_task_device_policy_log_probabilities": ([ct.c_void_p], ct.c_void_p),
    "mr_locomotion_task_read_policy_outputs": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t,
                                                ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_locomotion_task_write_policy_rollout_pack": ([ct.c_void_p, ct.POINTER(PolicyRolloutBatchC), ct.c_char_p,
                                                      ct.c_char_p], ct.c_int),
    "mr_world_enqueue_device_horizon": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_world_read_world_step_microseconds": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_enqueue_device_static": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint32, ct.c_void_p,
                                        ct.c_void_p], ct.c_int),
    "mr_world_wait": ([ct.c_void_p, ct.POINTER(WorldAdvanceC)], ct.c_int),
    "mr_world_read_state": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t, ct.c_void_p,
                             ct.c_size_t], ct.c_int),
    "mr_world_read_manifolds": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_world_read_statuses": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_read_contact_statuses": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_read_acceleration": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_read_evidence": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_stream": ([ct.c_void_p], ct.c_void_p),
    "mr_world_last_kernel_milliseconds": ([ct.c_void_p], ct.c_double),
    "mr_world_set_step_timing": ([ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_world_read_step_milliseconds": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_set_stage_profiling": ([ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_world_read_stage_cycles": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_create_task_rollout": ([ct.POINTER(RunManifestC)], ct.c_void_p),
    "mr_task_rollout_destroy": ([ct.c_void_p], None),
    "mr_task_rollout_reset": ([ct.c_void_p, ct.c_uint64], ct.c_int),
    "mr_task_rollout_set_state_readback": ([ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_task_rollout_advance": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t, ct.c_uint32,
                                 ct.c_uint64, ct.c_uint32, ct.POINTER(TaskRolloutAdvanceC)], ct.c_int),
    "mr_task_rollout_layout": ([ct.c_void_p], TaskRolloutLayoutC),
    "mr_task_rollout_status_codes": ([ct.c_void_p], ct.POINTER(ct.c_uint32)),
    "mr_task_rollout_active_contacts": ([ct.c_void_p], ct.POINTER(ct.c_uint32)),
    "mr_task_rollout_final_q": ([ct.c_void_p], ct.POINTER(ct.c_float)),
    "mr_task_rollout_copy_final_scene_states": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_world_set_schedule": ([ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_uint32], ct.c_int),
    "mr_world_set_schedule_reorder": ([ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_world_resident_workgroups": ([ct.c_void_p], ct.c_uint32),
    "mr_world_enqueue_graph": ([ct.c_void_p, ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_world_set_warm_start_mlp": ([ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_world_set_mlp_capture": ([ct.c_void_p, ct.c_uint32], ct.c_int),
    "mr_world_read_mlp_samples": ([ct.c_void_p, ct.c_void_p, ct.c_size_t], ct.c_int),
    "mr_mlp_forward": ([ct.POINTER(MlpWeightsC), ct.c_void_p, ct.c_uint32, ct.c_void_p, ct.POINTER(MlpTraceC),
                        ct.c_void_p], ct.c_int),
    "mr_mlp_loss_scratch_floats": ([ct.c_uint32], ct.c_uint32),
    "mr_mlp_loss": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint32, ct.c_float, ct.c_void_p, ct.c_void_p,
                     ct.c_void_p, ct.c_void_p], ct.c_int),
    "mr_mlp_backward": ([ct.POINTER(MlpWeightsC), ct.POINTER(MlpTraceC), ct.c_void_p, ct.c_uint32, ct.c_void_p,
                         ct.c_void_p, ct.POINTER(MlpGradsC), ct.c_void_p], ct.c_int),
    "mr_mlp_adam": ([ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint32, ct.c_float,
                     ct.c_float, ct.c_float, ct.c_float, ct.c_uint32, ct.c_void_p], ct.c_int),
}


class _Bindings:
    """Configured ctypes function table for one loaded library."""

    def __init__(self, path: str | os.PathLike[str] | None = None) -> None:
        self.path = resolve_library_path(path)
        self.lib = ct.CDLL(str(self.path))
        for name, (argtypes, restype) in _SIGNATURES.items():
            fn = getattr(self.lib, name)
            fn.argtypes = argtypes
            fn.restype = restype

    def last_error(self) -> str:
        raw = self.lib.mr_last_error()
        return (raw.decode() if raw else "") or "unknown native error"

    def check(self, status: int) -> None:
        if status != 0:
            raise MetalRoboError(self.last_error())


_binding_cache: dict[Path, _Bindings] = {}


def load(path: str | os.PathLike[str] | None = None) -> _Bindings:
    resolved = resolve_library_path(path)
    b = _binding_cache.get(resolved)
    if b is None:
        b = _Bindings(resolved)
        _binding_cache[resolved] = b
    return b


def version() -> str:
    return load().lib.mr_version().decode()


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


# ------------------------------------------------------------- high level

class WorldModel:
    """Owned immutable world set (one model block per environment or shared)."""

    def __init__(self, handle: int, bindings: _Bindings) -> None:
        self._b = bindings
        self._h = handle
        self.desc = WorldModelDescC()
        bindings.check(bindings.lib.mr_world_model_describe(handle, ct.byref(self.desc)))

    @classmethod
    def synthetic(cls, kind: int, environment_count: int, seed: int = 0,
                  library: str | os.PathLike[str] | None = None) -> "WorldModel":
        b = load(library)
        cfg = SyntheticWorldConfigC(kind, environment_count, seed)
        h = b.lib.mr_world_model_create_synthetic(ct.byref(cfg))
        if not h:
            raise MetalRoboError(b.last_error())
        return cls(h, b)

    def close(self) -> None:
        if self._h:
            self._b.lib.mr_world_model_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> int:
        return self._h

    @property
    def environment_count(self) -> int:
        return self.desc.environment_count

    @property
    def max_nq(self) -> int:
        return self.desc.max_nq

    @property
    def max_nv(self) -> int:
        return self.desc.max_nv

    @property
    def max_scene_bodies(self) -> int:
        return self.desc.max_scene_bodies

    def _view(self, ptr: int, dtype: np.dtype, count: int) -> np.ndarray:
        if count == 0:
            return np.zeros(0, dtype=dtype)
        buf = (ct.c_char * (dtype.itemsize * count)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype, count=count).copy()

    def models(self) -> np.ndarray:
        return self._view(self.desc.models, MODEL_HEADER_DTYPE, self.desc.model_count)

    def environment_model(self) -> np.ndarray:
        return self._view(self.desc.environment_model, np.dtype("<u4"), self.desc.environment_count)

    def bodies(self) -> np.ndarray:
        return self._view(self.desc.bodies, BODY_DTYPE, self.desc.body_total)

    def dofs(self) -> np.ndarray:
        return self._view(self.desc.dofs, DOF_DTYPE, self.desc.dof_total)

    def actuators(self) -> np.ndarray:
        """Per-DoF actuator profiles (empty when every DoF is inactive)."""
        return self._view(self.desc.actuators, ACTUATOR_DTYPE, self.desc.actuator_total)

    def shapes(self) -> np.ndarray:
        return self._view(self.desc.shapes, SHAPE_DTYPE, self.desc.shape_total)

    def materials(self) -> np.ndarray:
        return self._view(self.desc.materials, MATERIAL_DTYPE, self.desc.material_total)

    def pairs(self) -> np.ndarray:
        return self._view(self.desc.pairs, PAIR_DTYPE, self.desc.pair_total)

    def heights(self) -> np.ndarray:
        hf = self.desc.heightfield
        if not hf.present:
            return np.zeros((0, 0), np.float32)
        return self._view(self.desc.heightfield_heights, np.dtype("<f4"), hf.nx * hf.ny).reshape(hf.ny, hf.nx)

    def initial_state(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        E = self.environment_count
        q = np.zeros((E, self.max_nq), np.float32)
        v = np.zeros((E, self.max_nv), np.float32)
        s = np.zeros((E, self.max_scene_bodies), BODY_STATE_DTYPE)
        self._b.check(self._b.lib.mr_world_model_initial_state(self._h, _ptr(q), q.size, _ptr(v), v.size, _ptr(s),
                                                               s.size))
        return q, v, s

    def commands(self, first_step: int, step_count: int) -> np.ndarray:
        out = np.zeros((step_count, self.environment_count, self.max_nv), np.float32)
        self._b.check(self._b.lib.mr_world_model_commands(self._h, first_step, step_count, _ptr(out), out.size))
        return out

    def default_config(self) -> WorldStepConfigC:
        cfg = WorldStepConfigC()
        self._b.check(self._b.lib.mr_world_model_default_config(self._h, ct.byref(cfg)))
        return cfg


class World:
    """Resident batched world on one GPU (physics-only MetalWorldContext mirror)."""

    def __init__(self, model: WorldModel, config: Optional[WorldStepConfigC] = None) -> None:
        self._b = model._b
        self.model = model
        self.config = (config if config is not None else model.default_config()).copy()
        h = self._b.lib.mr_world_create(model.handle, ct.byref(self.config))
        if not h:
            raise MetalRoboError(self._b.last_error())
        self._h = h
        self.last_steps = 0

    def close(self) -> None:
        if self._h:
            self._b.lib.mr_world_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    @property
    def environment_count(self) -> int:
        return self.model.environment_count

    def layout(self) -> WorldLayoutC:
        out = WorldLayoutC()
        self._b.check(self._b.lib.mr_world_layout(self._h, ct.byref(out)))
        return out

    def set_state(self, q: np.ndarray, v: np.ndarray, scene: np.ndarray) -> None:
        q = np.ascontiguousarray(q, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        scene = np.ascontiguousarray(scene, BODY_STATE_DTYPE)
        self._b.check(self._b.lib.mr_world_initialize_resident_state(self._h, _ptr(q), q.size, _ptr(v), v.size,
                                                                     _ptr(scene), scene.size))

    def set_reset_state(self, q: np.ndarray, v: np.ndarray, scene: np.ndarray) -> None:
        q = np.ascontiguousarray(q, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        scene = np.ascontiguousarray(scene, BODY_STATE_DTYPE)
        self._b.check(self._b.lib.mr_world_set_reset_state(self._h, _ptr(q), q.size, _ptr(v), v.size, _ptr(scene),
                                                           scene.size))

    def write_manifolds(self, headers: np.ndarray, points: np.ndarray, counts: np.ndarray) -> None:
        headers = np.ascontiguousarray(headers, MANIFOLD_HEADER_DTYPE)
        points = np.ascontiguousarray(points, MANIFOLD_POINT_DTYPE)
        counts = np.ascontiguousarray(counts, np.uint32)
        self._b.check(self._b.lib.mr_world_write_manifolds(self._h, _ptr(headers), _ptr(points), _ptr(counts)))

    def state(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        E = self.environment_count
        q = np.zeros((E, self.model.max_nq), np.float32)
        v = np.zeros((E, self.model.max_nv), np.float32)
        s = np.zeros((E, self.model.max_scene_bodies), BODY_STATE_DTYPE)
        self._b.check(self._b.lib.mr_world_read_state(self._h, _ptr(q), q.size, _ptr(v), v.size, _ptr(s), s.size))
        return q, v, s

    def manifolds(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        E, cap = self.environment_count, self.config.manifold_capacity
        h = np.zeros((E, cap), MANIFOLD_HEADER_DTYPE)
        p = np.zeros((E, cap, MR_MANIFOLD_POINT_CAPACITY), MANIFOLD_POINT_DTYPE)
        c = np.zeros(E, np.uint32)
        self._b.check(self._b.lib.mr_world_read_manifolds(self._h, _ptr(h), _ptr(p), _ptr(c)))
        return h, p, c

    def advance(self, commands: np.ndarray, reset_masks: Optional[np.ndarray] = None) -> WorldAdvanceC:
        commands = np.ascontiguousarray(commands, np.float32)
        steps = commands.shape[0]
        masks = None if reset_masks is None else np.ascontiguousarray(reset_masks, np.uint32)
        out = WorldAdvanceC()
        self._b.check(self._b.lib.mr_world_advance(self._h, _ptr(commands), commands.size, _ptr(masks),
                                                   0 if masks is None else masks.size, steps, ct.byref(out)))
        self.last_steps = steps
        return out

    def _device_commands(self, commands) -> int:
        """Checks a device command tensor [steps][E][max_nv] (contiguous float32 on the GPU): the
        kernel reads steps x E x max_nv floats from its address, so nothing smaller may pass."""
        import torch

        steps = int(commands.shape[0]) if commands.dim() > 0 else 0
        expect = (steps, self.environment_count, self.model.max_nv)
        if (steps == 0 or tuple(commands.shape) != expect or commands.dtype != torch.float32
                or not commands.is_contiguous() or not commands.is_cuda):
            raise ValueError(f"device commands must be a contiguous float32 GPU tensor of shape {expect}")
        return steps

    def _device_masks(self, reset_masks, steps: int) -> Optional[int]:
        import torch

        if reset_masks is None:
            return None
        if (tuple(reset_masks.shape) != (steps, self.environment_count) or not reset_masks.is_cuda
                or not reset_masks.is_contiguous() or reset_masks.dtype not in (torch.int32, torch.uint32)):
            raise ValueError("device reset masks must be a contiguous int32/uint32 GPU tensor "
                             "[steps][environments]")
        return reset_masks.data_ptr()

    def enqueue_device(self, commands, reset_masks=None) -> None:
        """Enqueue with inputs already in device memory (torch tensors on the GPU)."""
        steps = self._device_commands(commands)
        mask_ptr = self._device_masks(reset_masks, steps)
        self._b.check(self._b.lib.mr_world_enqueue_device(self._h, commands.data_ptr(), mask_ptr, steps))
        self.last_steps = steps

    def enqueue_device_horizon(self, commands, reset_masks=None) -> None:
        """mr_world_enqueue_device_horizon: with a work-queue schedule, every scheduled world runs all
        control steps of `commands` ([steps][E][max_nv] on the GPU) inside one persistent launch."""
        steps = self._device_commands(commands)
        mask_ptr = self._device_masks(reset_masks, steps)
        self._b.check(self._b.lib.mr_world_enqueue_device_horizon(self._h, commands.data_ptr(), mask_ptr, steps))
        self.last_steps = steps

    def world_step_microseconds(self) -> np.ndarray:
        """[steps][E] wall-clock microseconds of each control step of each world (step timing on)."""
        out = np.zeros((self.last_steps, self.environment_count), np.float32)
        self._b.check(self._b.lib.mr_world_read_world_step_microseconds(self._h, _ptr(out), out.size))
        return out

    def enqueue_device_static(self, commands, status, contact_status, reset_masks=None) -> None:
        """mr_world_enqueue_device_static: enqueue_device that allocates no device memory. The
        per-step scratch comes from the library's static device array; `status` and
        `contact_status` are the caller's contiguous uint8 GPU tensors of steps x E x 48 and
        steps x E x 96 bytes that receive the statuses."""
        import torch

        steps = self._device_commands(commands)
        mask_ptr = self._device_masks(reset_masks, steps)
        E = self.environment_count
        for t, rec in ((status, WORLD_STATUS_DTYPE), (contact_status, CONTACT_STATUS_DTYPE)):
            if (t.dtype != torch.uint8 or not t.is_cuda or not t.is_contiguous()
                    or t.numel() != steps * E * rec.itemsize):
                raise ValueError(f"status buffers must be contiguous uint8 GPU tensors of "
                                 f"{steps} x {E} x {rec.itemsize} bytes")
        self._b.check(self._b.lib.mr_world_enqueue_device_static(self._h, commands.data_ptr(), mask_ptr, steps,
                                                                 status.data_ptr(), contact_status.data_ptr()))
        self.last_steps = steps

    def wait(self) -> WorldAdvanceC:
        out = WorldAdvanceC()
        self._b.check(self._b.lib.mr_world_wait(self._h, ct.byref(out)))
        return out

    def statuses(self) -> np.ndarray:
        out = np.zeros((self.last_steps, self.environment_count), WORLD_STATUS_DTYPE)
        self._b.check(self._b.lib.mr_world_read_statuses(self._h, _ptr(out), out.size))
        return out

    def contact_statuses(self) -> np.ndarray:
        out = np.zeros((self.last_steps, self.environment_count), CONTACT_STATUS_DTYPE)
        self._b.check(self._b.lib.mr_world_read_contact_statuses(self._h, _ptr(out), out.size))
        return out

    def acceleration(self) -> np.ndarray:
        out = np.zeros((self.environment_count, self.model.max_nv), np.float32)
        self._b.check(self._b.lib.mr_world_read_acceleration(self._h, _ptr(out), out.size))
        return out

    def evidence(self) -> tuple[np.ndarray, np.ndarray]:
        E, cap = self.environment_count, self.config.pair_capacity
        pairs = np.zeros((E, cap), PAIR_RECORD_DTYPE)
        counts = np.zeros(E, np.uint32)
        self._b.check(self._b.lib.mr_world_read_evidence(self._h, _ptr(pairs), pairs.size, _ptr(counts),
                                                         counts.size))
        return pairs, counts

    def stream(self) -> int:
        return self._b.lib.mr_world_stream(self._h) or 0

    def last_kernel_milliseconds(self) -> float:
        return float(self._b.lib.mr_world_last_kernel_milliseconds(self._h))

    def set_step_timing(self, enabled: bool) -> None:
        self._b.check(self._b.lib.mr_world_set_step_timing(self._h, 1 if enabled else 0))

    def set_schedule(self, order=None, grid: int = 0) -> None:
        """grid = 0: one workgroup per environment. grid > 0: the persistent work queue with
        `grid` workgroups over `order` (environment indices, a subset allowed)."""
        if grid == 0:
            self._b.check(self._b.lib.mr_world_set_schedule(self._h, None, 0, 0))
            return
        o = np.ascontiguousarray(order, dtype=np.uint32)
        self._b.check(self._b.lib.mr_world_set_schedule(self._h, _ptr(o), o.size, int(grid)))

    def set_schedule_reorder(self, enabled: bool = True) -> None:
        """Re-sort the work queue after every control step by that step's cone sweeps, heaviest
        first (needs set_schedule over every environment)."""
        self._b.check(self._b.lib.mr_world_set_schedule_reorder(self._h, 1 if enabled else 0))

    def resident_workgroups(self) -> int:
        n = int(self._b.lib.mr_world_resident_workgroups(self._h))
        if n == 0:
            raise MetalRoboError(self._b.last_error())
        return n

    def enqueue_graph(self, commands) -> None:
        """Graph-captured horizon over device commands [steps][E][max_nv] (a torch tensor kept at
        one address between calls); wait() afterwards."""
        steps = self._device_commands(commands)
        self._b.check(self._b.lib.mr_world_enqueue_graph(self._h, commands.data_ptr(), steps))
        self.last_steps = steps

    def set_warm_start_mlp(self, mlp) -> None:
        """Warm-start later advances from `mlp` (a metalrobo.warmstart.WarmStartMLP, kept alive
        here while set), or restore the plain warm start with None."""
        self._mlp = mlp
        if mlp is None:
            self._b.check(self._b.lib.mr_world_set_warm_start_mlp(self._h, None))
        else:
            self._mlp_weights = mlp.weights()
            self._b.check(self._b.lib.mr_world_set_warm_start_mlp(self._h, ct.byref(self._mlp_weights)))

    def set_mlp_capture(self, enabled: bool) -> None:
        self._b.check(self._b.lib.mr_world_set_mlp_capture(self._h, 1 if enabled else 0))

    def mlp_samples(self) -> np.ndarray:
        """[E][constraint_capacity][20]: 8 features, converged impulse (3), valid flag, the warm start the
        solve began from (3), 1.0 where the network replaced the plain warm start, the network's scaled
        correction (3) and 1.0 where its features were in bounds (last captured step)."""
        cap = self.config.constraint_capacity
        out = np.zeros((self.environment_count, cap, 20), np.float32)
        self._b.check(self._b.lib.mr_world_read_mlp_samples(self._h, _ptr(out), out.size))
        return out

    STAGE_NAMES = ("kinematics", "aba", "projection", "broadphase", "narrowphase", "emission", "responses",
                   "evaluation", "warm_start", "sweeps", "residuals", "integration")

    def set_stage_profiling(self, enabled: bool) -> None:
        self._b.check(self._b.lib.mr_world_set_stage_profiling(self._h, 1 if enabled else 0))

    def stage_cycles(self) -> np.ndarray:
        out = np.zeros((self.environment_count, 16), np.uint64)
        self._b.check(self._b.lib.mr_world_read_stage_cycles(self._h, _ptr(out), out.size))
        return out

    def step_milliseconds(self) -> np.ndarray:
        out = np.zeros(self.last_steps, np.float64)
        self._b.check(self._b.lib.mr_world_read_step_milliseconds(self._h, _ptr(out), out.size))
        return out


class TaskRollout:
    """Stepping surface named after the reference's task-rollout ABI (c_api.h:583-746)."""

    def __init__(self, source: int, environment_count: int, seed: int = 0, physics_substeps: int = 0,
                 velocity_iterations: int = 0, final_velocity_iterations: int = 0,
                 control_timestep_seconds: float = 0.0, library: str | os.PathLike[str] | None = None) -> None:
        self._b = load(library)
        manifest = RunManifestC()
        manifest.profile.environment_count = environment_count
        manifest.profile.physics_substeps = physics_substeps
        manifest.profile.velocity_iterations = velocity_iterations
        manifest.profile.final_velocity_iterations = final_velocity_iterations
        manifest.profile.control_timestep_seconds = control_timestep_seconds
        manifest.profile.seed = seed
        manifest.source = source
        manifest.surface = MR_LOCOMOTION_SURFACE_TERRAIN if source == MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN \
            else MR_LOCOMOTION_SURFACE_GROUND
        h = self._b.lib.mr_create_task_rollout(ct.byref(manifest))
        if not h:
            raise MetalRoboError(self._b.last_error())
        self._h = h

    def close(self) -> None:
        if self._h:
            self._b.lib.mr_task_rollout_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def layout(self) -> TaskRolloutLayoutC:
        return self._b.lib.mr_task_rollout_layout(self._h)

    def reset(self, seed: int) -> None:
        self._b.check(self._b.lib.mr_task_rollout_reset(self._h, seed))

    def set_state_readback(self, enabled: bool) -> None:
        self._b.check(self._b.lib.mr_task_rollout_set_state_readback(self._h, 1 if enabled else 0))

    def advance(self, normalized_actions: np.ndarray, reset_masks: Optional[np.ndarray] = None,
                policy_revision: int = 0) -> TaskRolloutAdvanceC:
        actions = np.ascontiguousarray(normalized_actions, np.float32)
        steps = actions.shape[0]
        masks = None if reset_masks is None else np.ascontiguousarray(reset_masks, np.uint32)
        out = TaskRolloutAdvanceC()
        self._b.check(self._b.lib.mr_task_rollout_advance(self._h, _ptr(actions), actions.size, _ptr(masks),
                                                          0 if masks is None else masks.size, steps,
                                                          policy_revision, 0, ct.byref(out)))
        self._last_steps = steps
        return out

    def status_codes(self) -> np.ndarray:
        lay = self.layout()
        ptr = self._b.lib.mr_task_rollout_status_codes(self._h)
        n = self._last_steps * lay.environment_count
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self._last_steps, lay.environment_count).copy()

    def active_contacts(self) -> np.ndarray:
        lay = self.layout()
        ptr = self._b.lib.mr_task_rollout_active_contacts(self._h)
        n = self._last_steps * lay.environment_count
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self._last_steps, lay.environment_count).copy()

    def final_q(self) -> np.ndarray:
        lay = self.layout()
        ptr = self._b.lib.mr_task_rollout_final_q(self._h)
        if not ptr:
            raise MetalRoboError("no final q (enable state readback and advance first)")
        return np.ctypeslib.as_array(ptr, shape=(lay.environment_count * lay.nq,)).reshape(
            lay.environment_count, lay.nq).copy()

    def final_scene_states(self) -> np.ndarray:
        lay = self.layout()
        out = np.zeros((lay.environment_count, lay.scene_body_count, 13), np.float32)
        self._b.check(self._b.lib.mr_task_rollout_copy_final_scene_states(self._h, _ptr(out), out.size))
        return out


class LocomotionTask:
    """mr_locomotion_task_*: the task layer either side of a World's control step. `program` is a
    dict with the record arrays (TASK_*_DTYPE) and the MRLocomotionTaskProgramC scalars; the world's
    reset state becomes the program's nominal pose at rest."""

    def __init__(self, world: World, program: dict) -> None:
        self._b = load()
        self.world = world
        self.program = program
        self._keep = {k: np.ascontiguousarray(program[k], dt) for k, dt in (
            ("actions", TASK_ACTION_DTYPE), ("observations", TASK_OBSERVATION_DTYPE),
            ("rewards", TASK_REWARD_DTYPE), ("terminations", TASK_TERMINATION_DTYPE))}
        self._keep["default_q"] = np.ascontiguousarray(program["default_q"], np.float32)
        self.c_program = task_program_struct(program, self._keep)
        self._h = self._b.lib.mr_locomotion_task_create(ct.byref(self.c_program), world._h)
        if not self._h:
            raise MetalRoboError(self._b.last_error())
        self.environment_count = world.environment_count
        self.history_length = int(program["history_length"])
        self.observation_count = len(self._keep["observations"])
        self.action_count = len(self._keep["actions"])

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._b.lib.mr_locomotion_task_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> None:
        self._b.check(self._b.lib.mr_locomotion_task_reset(self._h))

    def step(self, actions=None) -> None:
        """One control step with device actions (contiguous float32 GPU tensor [E][action_count]),
        or None when a policy is installed."""
        import torch

        if actions is None:
            self._b.check(self._b.lib.mr_locomotion_task_step(self._h, None))
            return
        if (tuple(actions.shape) != (self.environment_count, self.action_count) or actions.dtype != torch.float32
                or not actions.is_cuda or not actions.is_contiguous()):
            raise ValueError(f"actions must be a contiguous float32 GPU tensor of shape "
                             f"{(self.environment_count, self.action_count)}")
        self._b.check(self._b.lib.mr_locomotion_task_step(self._h, actions.data_ptr()))

    def set_policy(self, pack) -> None:
        """mr_locomotion_task_set_policy with a metalrobo.packs.PolicyPack."""
        c, keep = pack.to_c()
        self._b.check(self._b.lib.mr_locomotion_task_set_policy(self._h, ct.byref(c)))
        del keep

    def load_policy_pack(self, path: str) -> None:
        self._b.check(self._b.lib.mr_locomotion_task_load_policy_pack(self._h, str(path).encode()))

    def clear_policy(self) -> None:
        self._b.check(self._b.lib.mr_locomotion_task_clear_policy(self._h))

    def policy_revision(self) -> int:
        return int(self._b.lib.mr_locomotion_task_policy_revision(self._h))

    def read_policy_outputs(self):
        """(actions [E][A], latents [E][A], log probabilities [E]) of the last policy step, after waiting."""
        E, A = self.environment_count, self.action_count
        actions = np.zeros((E, A), np.float32)
        latents = np.zeros((E, A), np.float32)
        logp = np.zeros(E, np.float32)
        self._b.check(self._b.lib.mr_locomotion_task_read_policy_outputs(
            self._h, _ptr(actions), actions.size, _ptr(latents), latents.size, _ptr(logp), logp.size))
        return actions, latents, logp

    def write_policy_rollout_pack(self, path: str, batch_id: str, control_step_count: int, actor_observations,
                                  critic_observations, latents, log_probabilities, values, bootstrap_values,
                                  transitions, teacher_actions=None) -> None:
        """mr_locomotion_task_write_policy_rollout_pack from host numpy tables (step-major)."""
        keep = []
        c = PolicyRolloutBatchC()
        c.control_step_count = int(control_step_count)
        for (name, count), arr in zip(_ROLLOUT_BATCH_TABLES, (
                actor_observations, critic_observations, None, teacher_actions, latents, log_probabilities, values,
                bootstrap_values, transitions)):
            if arr is None:
                setattr(c, name, None)
                setattr(c, count, 0)
                continue
            a = (np.ascontiguousarray(arr, TASK_TRANSITION_DTYPE) if name == "transitions"
                 else np.ascontiguousarray(np.asarray(arr, np.float32).reshape(-1)))
            keep.append(a)
            setattr(c, name, a.ctypes.data if a.size else None)
            setattr(c, count, a.size)
        self._b.check(self._b.lib.mr_locomotion_task_write_policy_rollout_pack(
            self._h, ct.byref(c), batch_id.encode(), str(path).encode()))

    def read(self):
        """(observations [E][history][O], transitions [E], task states [E]) after waiting."""
        E = self.environment_count
        obs = np.zeros((E, self.history_length, self.observation_count), np.float32)
        tr = np.zeros(E, TASK_TRANSITION_DTYPE)
        st = np.zeros(E, TASK_STATE_DTYPE)
        self._b.check(self._b.lib.mr_locomotion_task_read(self._h, _ptr(obs), obs.size, _ptr(tr), tr.size, _ptr(st),
                                                          st.size))
        return obs, tr, st

    def device_observations(self) -> int:
        return self._b.lib.mr_locomotion_task_device_observations(self._h)

    def device_targets(self) -> int:
        return self._b.lib.mr_locomotion_task_device_targets(self._h)


def task_program_struct(program: dict, arrays: dict) -> LocomotionTaskProgramC:
    """The MRLocomotionTaskProgramC of `program`, pointing into `arrays` (kept alive by the caller)."""
    c = LocomotionTaskProgramC()
    c.actions = arrays["actions"].ctypes.data
    c.observations = arrays["observations"].ctypes.data
    c.rewards = arrays["rewards"].ctypes.data if len(arrays["rewards"]) else None
    c.terminations = arrays["terminations"].ctypes.data if len(arrays["terminations"]) else None
    c.default_q = arrays["default_q"].ctypes.data
    c.action_count = len(arrays["actions"])
    c.observation_count = len(arrays["observations"])
    c.reward_count = len(arrays["rewards"])
    c.termination_count = len(arrays["terminations"])
    for name in ("history_length", "delay_slots", "max_episode_steps", "flags", "root_q_offset", "root_v_offset",
                 "actuator_delay", "seed"):
        setattr(c, name, int(program.get(name, 0)))
    for name in ("base_height_target", "gait_period"):
        setattr(c, name, float(program[name]))
    for name, n in (("root_reference", 4), ("command_lower", 4), ("command_upper", 4), ("command_seconds", 2)):
        vals = list(program.get(name, [0.0] * n)) + [0.0] * n
        getattr(c, name)[:] = [float(x) for x in vals[:n]]
    return c


def policy_dense_layer(weights, bias, inp, out, activation: int, mean=None, inverse_std=None,
                       observation_clip: float = 0.0) -> None:
    """mr_policy_dense_layer on torch GPU tensors: weights [out][in], bias [out], inp [E][>=in],
    out [E][>=out] (row strides from the tensors), optional mean / inverse_std [in]."""
    import torch

    b = load()
    E, in_w = int(inp.shape[0]), int(weights.shape[1])
    out_w = int(weights.shape[0])
    for t in (weights, bias, inp, out) + tuple(x for x in (mean, inverse_std) if x is not None):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("policy tensors must be contiguous float32 GPU tensors")
    stream = torch.cuda.current_stream().cuda_stream
    b.check(b.lib.mr_policy_dense_layer(weights.data_ptr(), bias.data_ptr(),
                                        None if mean is None else mean.data_ptr(),
                                        None if inverse_std is None else inverse_std.data_ptr(),
                                        float(observation_clip), inp.data_ptr(), int(inp.stride(0)), out.data_ptr(),
                                        int(out.stride(0)), E, in_w, out_w, int(activation), stream))


def policy_sample_and_score(mean, action_bias, action_scale, log_std, action_clip: float, seed: int, revision: int,
                            task_states, control_step: int, actions_out, latents_out, log_prob_out) -> None:
    """mr_policy_sample_and_score on torch GPU tensors; task_states is a uint8 GPU tensor holding
    [E] MRTaskStateC records (176 bytes each); log_std may be None (deterministic)."""
    import torch

    b = load()
    E, A = int(mean.shape[0]), int(action_bias.shape[0])
    stream = torch.cuda.current_stream().cuda_stream
    b.check(b.lib.mr_policy_sample_and_score(mean.data_ptr(), int(mean.stride(0)), action_bias.data_ptr(),
                                             action_scale.data_ptr(), None if log_std is None else log_std.data_ptr(),
                                             float(action_clip), int(seed), int(revision), task_states.data_ptr(), E,
                                             A, int(control_step), actions_out.data_ptr(), latents_out.data_ptr(),
                                             log_prob_out.data_ptr(), stream))
