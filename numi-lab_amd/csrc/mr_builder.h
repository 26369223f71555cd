// Synthetic world construction for the benchmark configurations and a
// generic owned model container shared by the executor and the C ABI.
#pragma once

#include "../../include/mr_world_types.h"

#include <cstdint>
#include <string>
#include <vector>

namespace mr {

// Counter-based SplitMix64 keyed on (seed, environment, step, channel): the
// published SplitMix64 finalizer (Steele, Lea, Flood 2014) applied to a
// chained key, so every input is reproducible from its coordinates.
uint64_t splitmix64(uint64_t z);
uint64_t random_key(uint64_t seed, uint64_t env, uint64_t step, uint64_t channel);
double random_unit(uint64_t seed, uint64_t env, uint64_t step, uint64_t channel);
double random_normal(uint64_t seed, uint64_t env, uint64_t step, uint64_t channel);

enum SyntheticKind : uint32_t {
    kChainBoxes = 1u,       // 7-DoF chain + 20 boxes on a plane (effort mode)
    kHumanoidTerrain = 2u,  // floating 29-DoF humanoid on a heightfield (PD mode)
    kRandomTrees = 3u       // per-world random 6-40 DoF revolute tree + free boxes/spheres (effort mode)
};

// Owned model + initial state. All per-environment arrays use the strides
// max_nq / max_nv / max_scene_bodies.
struct ModelData {
    std::vector<MRModelHeaderC> models;
    std::vector<uint32_t> environment_model;
    std::vector<MRBodyC> bodies;
    std::vector<MRDofC> dofs;
    std::vector<MRActuatorProfileC> actuators;  // parallel to dofs, or empty: every DoF inactive
    std::vector<MRShapeC> shapes;
    std::vector<MRMaterialC> materials;
    std::vector<MRCompiledPairC> pairs;
    std::vector<float> heights;
    MRHeightfieldC heightfield{};
    // convex and mesh geometry, shared by every model block
    std::vector<MRGeometryC> geometries;
    std::vector<mr_float4> geometry_vertices;
    std::vector<MRTriangleC> triangles;   // each mesh's triangles in BVH leaf order
    std::vector<MRBvhNodeC> bvh_nodes;    // built by compile_geometry
    mr_float4 gravity{0.0f, 0.0f, -9.81f, 0.0f};
    uint32_t max_nq = 0, max_nv = 0, max_scene_bodies = 0, max_shapes = 0, max_pairs = 0;
    std::vector<float> initial_q, initial_v;
    std::vector<MRBodyStateC> initial_scene;
    // Synthetic command generator state.
    uint32_t kind = 0;
    uint64_t seed = 0;
    std::vector<float> standing_pose; // implicit-drive centre (max_nv per model 0)
    MRWorldStepConfigC default_config{};

    uint32_t environment_count() const { return (uint32_t)environment_model.size(); }
    MRWorldModelDescC describe() const;
    // Fills the per-step command stream [step][environment][max_nv] for
    // control steps [first_step, first_step + step_count).
    void commands(uint32_t first_step, uint32_t step_count, float* out) const;
};

// Validates the convex and mesh geometries, fills their bounds and builds
// each mesh's BVH4 (reordering its triangles into leaf order; the triangle
// feature ids travel with the triangles). Throws std::runtime_error.
void compile_geometry(ModelData& model);
// Validates topology/capacity rules the fused kernel relies on and fills
// derived fields (levels, damping scales for timestep h). Throws
// std::runtime_error with a message on violation.
void finalize_model(ModelData& model, float microstep);
// Copies a caller-described model, validating it (finalize_model) before
// the default initial state is sized and filled.
ModelData copy_model(const MRWorldModelDescC& desc, float microstep);
ModelData build_synthetic(uint32_t kind, uint32_t environment_count, uint64_t seed);

// Canonical eligible pairs (MetalWorld.mm:14546-14615): every unordered
// collider pair A<B except disabled shapes, same-body pairs, pairs failing
// the group/mask filter, static/kinematic-only pairs, links within
// `exclusion_distance` tree edges of each other, and unsupported classes.
void compile_pairs(const MRModelHeaderC& header, const MRBodyC* bodies, const MRShapeC* shapes,
                   uint32_t exclusion_distance, std::vector<MRCompiledPairC>& out);
uint32_t pair_class(uint32_t type_a, uint32_t type_b);

} // namespace mr
