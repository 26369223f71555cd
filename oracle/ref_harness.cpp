// TEST INFRASTRUCTURE — a C ABI over the reference's own FP64 CPU oracle,
// compiled from the reference sources where they lie (oracle/build_ref.py)
// into oracle/_ref/libmr_ref.so. Only tests/ load it, to check this
// repository's FP32 oracle against the reference's code; nothing in the
// product links or calls it.
//
// Wrapped reference entry points:
//   * computeArticulatedForwardDynamics  (ArticulatedDynamics.hpp:173-182,
//     ArticulatedDynamics.cpp:2241) — FP64 CRBA + RNEA forward dynamics;
//   * collideCpuReference                 (Collision.hpp, Collision.cpp:2959-3228)
//     — FP64 SAP broadphase, analytic narrowphase, persistent manifolds;
//   * solveReferenceConicProblem          (ReferenceConicSolver.hpp,
//     ReferenceConicSolver.cpp:182) — exact circular-cone contact solve;
//   * solveContactConstraints             (ConstraintSolver.hpp:101-105,
//     ConstraintSolver.cpp:1047) — projected Gauss-Seidel over free bodies.
//
// Arguments are plain pointers and sizes. Records that cross this boundary
// are declared here with the byte layout of the matching reference records,
// so the caller packs them from its own tables without including either
// project's headers.

#include "metalrobo/ArticulatedDynamics.hpp"
#include "metalrobo/Collision.hpp"
#include "metalrobo/ConstraintSolver.hpp"
#include "metalrobo/EngineModel.hpp"
#include "metalrobo/ReferenceConicSolver.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <span>
#include <string>
#include <vector>

namespace {

using namespace metalrobo;

void set_error(char* err, size_t len, const std::string& msg) {
    if (err != nullptr && len > 0) {
        std::snprintf(err, len, "%s", msg.c_str());
    }
}

mr_float4 f4(const float* p, float w = 0.0f) { return mr_float4{p[0], p[1], p[2], w}; }

mr_float4 q4(const float* p) { return mr_float4{p[0], p[1], p[2], p[3]}; }

}  // namespace

extern "C" {

// One link of a tree articulation, in the caller's order (link 0 is the root,
// every parent precedes its children). Inertia is the body-frame tensor about
// the COM; anchors are relative to each body's COM; rotations are (x, y, z, w)
// joint frames; joint_type uses MRJointType values (revolute, continuous,
// fixed) and q_index / v_index are local coordinates of a one-DoF joint.
typedef struct MRRefLink {
    uint32_t parent;
    uint32_t joint_type;
    uint32_t q_index;
    uint32_t v_index;
    float mass;
    float inertia[9];
    float axis[3];
    float parent_anchor[3];
    float child_anchor[3];
    float parent_rotation[4];
    float child_rotation[4];
    float damping[2];  // linear, angular
} MRRefLink;

// Generalized accelerations of one tree (fixed or floating root) at (q, v)
// under generalized force tau. armature[nv] is the generalized rotor inertia
// per velocity coordinate. Returns 0, or the reference's
// ArticulatedDynamicsStatus (> 0), or -1 when the model is rejected (err).
int mr_ref_forward_dynamics(uint32_t root_type, uint32_t link_count, const MRRefLink* links, uint32_t nq,
                            uint32_t nv, const float* armature, const float* gravity, double timestep,
                            uint32_t apply_body_damping, const double* q, const double* v, const double* tau,
                            double* qdd, char* err, size_t errlen) {
    try {
        EngineModel m;
        const bool floating = root_type == MR_ROOT_FLOATING;
        m.world.abiVersion = MR_ENGINE_ABI_VERSION;
        m.world.bodyCount = link_count;
        m.world.articulationCount = 1u;
        m.world.jointCount = link_count - 1u;
        m.world.shapeCount = 0u;
        m.world.materialCount = 0u;
        m.world.nq = nq;
        m.world.nv = nv;
        m.world.pairCapacity = 1u;
        m.world.contactCapacity = 1u;
        m.world.constraintCapacity = 1u;
        m.world.islandCapacity = 1u;
        m.world.gravityAndTimestep = mr_float4{gravity[0], gravity[1], gravity[2], (float)timestep};
        m.world.solverScales = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};

        MRArticulationGPU a{};
        a.rootBody = 0u;
        a.rootType = root_type;
        a.firstBody = 0u;
        a.bodyCount = link_count;
        a.firstJoint = 0u;
        a.jointCount = link_count - 1u;
        a.qOffset = 0u;
        a.nq = nq;
        a.vOffset = 0u;
        a.nv = nv;
        m.articulations.push_back(a);

        m.dofs.resize(nv);
        if (floating) {
            // the reference's root coordinates carry no drive or armature (EngineModel.cpp:194-209)
            for (uint32_t d = 0; d < 6u; ++d) {
                if (armature[d] != 0.0f) {
                    set_error(err, errlen, "floating-root coordinates cannot carry armature");
                    return -1;
                }
                MRDofPropertiesGPU& p = m.dofs[d];
                std::memset(&p, 0, sizeof(p));
                p.articulationIndex = 0u;
                p.jointIndex = MR_INVALID_INDEX;
                p.qIndex = d < 3u ? d : MR_INVALID_INDEX;
                p.vIndex = d;
                p.localDof = d;
                p.flags = MR_DOF_FLAG_ROOT;
            }
        }
        for (uint32_t i = 0; i < link_count; ++i) {
            const MRRefLink& l = links[i];
            MRBodyPropertiesGPU b{};
            b.articulationIndex = 0u;
            b.parentBody = i == 0u ? MR_INVALID_INDEX : l.parent;
            b.inboundJoint = i == 0u ? MR_INVALID_INDEX : i - 1u;
            b.motionType = MR_MOTION_DYNAMIC;
            b.massAndInverseMass = mr_float4{l.mass, l.mass > 0.0f ? 1.0f / l.mass : 0.0f, 0.0f, 0.0f};
            b.centerOfMass = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
            b.inertiaRow0 = f4(l.inertia + 0);
            b.inertiaRow1 = f4(l.inertia + 3);
            b.inertiaRow2 = f4(l.inertia + 6);
            // the reference reads the inverse from its own rows; give it the exact inverse of
            // the (diagonal or general) inertia in double, rounded once
            const double I[9] = {l.inertia[0], l.inertia[1], l.inertia[2], l.inertia[3], l.inertia[4],
                                 l.inertia[5], l.inertia[6], l.inertia[7], l.inertia[8]};
            const double det = I[0] * (I[4] * I[8] - I[5] * I[7]) - I[1] * (I[3] * I[8] - I[5] * I[6]) +
                               I[2] * (I[3] * I[7] - I[4] * I[6]);
            double inv[9] = {0};
            if (det != 0.0) {
                inv[0] = (I[4] * I[8] - I[5] * I[7]) / det;
                inv[1] = (I[2] * I[7] - I[1] * I[8]) / det;
                inv[2] = (I[1] * I[5] - I[2] * I[4]) / det;
                inv[3] = (I[5] * I[6] - I[3] * I[8]) / det;
                inv[4] = (I[0] * I[8] - I[2] * I[6]) / det;
                inv[5] = (I[2] * I[3] - I[0] * I[5]) / det;
                inv[6] = (I[3] * I[7] - I[4] * I[6]) / det;
                inv[7] = (I[1] * I[6] - I[0] * I[7]) / det;
                inv[8] = (I[0] * I[4] - I[1] * I[3]) / det;
            }
            b.inverseInertiaRow0 = mr_float4{(float)inv[0], (float)inv[1], (float)inv[2], 0.0f};
            b.inverseInertiaRow1 = mr_float4{(float)inv[3], (float)inv[4], (float)inv[5], 0.0f};
            b.inverseInertiaRow2 = mr_float4{(float)inv[6], (float)inv[7], (float)inv[8], 0.0f};
            b.dampingAndSpeedLimits = mr_float4{l.damping[0], l.damping[1], 0.0f, 0.0f};
            m.bodies.push_back(b);
            if (i == 0u) continue;
            MRJointDescriptorGPU j{};
            j.parentBody = l.parent;
            j.childBody = i;
            j.jointType = l.joint_type;
            const bool one_dof = l.joint_type == MR_JOINT_REVOLUTE || l.joint_type == MR_JOINT_CONTINUOUS;
            j.qOffset = one_dof ? l.q_index : 0u;
            j.nq = one_dof ? 1u : 0u;
            j.vOffset = one_dof ? l.v_index : 0u;
            j.nv = one_dof ? 1u : 0u;
            j.axis0 = f4(l.axis);
            j.parentAnchor = f4(l.parent_anchor);
            j.childAnchor = f4(l.child_anchor);
            j.parentRotation = q4(l.parent_rotation);
            j.childRotation = q4(l.child_rotation);
            m.joints.push_back(j);
            if (one_dof) {
                MRDofPropertiesGPU& p = m.dofs[l.v_index];
                std::memset(&p, 0, sizeof(p));
                p.articulationIndex = 0u;
                p.jointIndex = i - 1u;
                p.qIndex = l.q_index;
                p.vIndex = l.v_index;
                p.localDof = 0u;
                p.drive = mr_float4{0.0f, 0.0f, armature[l.v_index], 0.0f};
            }
        }
        m.defaultQ.assign(nq, 0.0f);
        m.defaultV.assign(nv, 0.0f);
        if (floating) m.defaultQ[6] = 1.0f;
        std::string reason;
        if (!m.valid(&reason)) {
            set_error(err, errlen, "reference model rejected: " + reason);
            return -1;
        }
        ArticulatedDynamicsConfig cfg;
        cfg.gravity = {gravity[0], gravity[1], gravity[2]};
        cfg.timestep = timestep;
        cfg.applyBodyDamping = apply_body_damping != 0u;
        const ArticulatedDynamicsDiagnostics d = computeArticulatedForwardDynamics(
            m, 0u, std::span<const double>(q, nq), std::span<const double>(v, nv), std::span<const double>(tau, nv),
            {}, std::span<double>(qdd, nv), cfg);
        if (!d.succeeded()) {
            set_error(err, errlen, "forward dynamics status " + std::to_string((uint32_t)d.status));
            return (int)d.status;
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Shape record with the byte layout of the caller's 96-byte collider table
// entry: body, type, material, flags, group, mask, generation, reserved, then
// local position, local rotation (x, y, z, w), dimensions and
// (contact offset, rest offset, bounding radius, -).
typedef struct MRRefShape {
    uint32_t body;
    uint32_t shape_type;
    uint32_t material;
    uint32_t flags;
    uint32_t collision_group;
    uint32_t collision_mask;
    uint32_t slot_generation;
    uint32_t reserved;
    float local_position[4];
    float local_rotation[4];
    float dimensions[4];
    float contact_rest_radius[4];
} MRRefShape;

// One environment through the reference's FP64 collision oracle from an empty
// manifold cache. bodies: 128-byte body-state records in the layout of
// MRBodyStateGPU (engine_types.h:418-434). exclusions: [count][2] collider
// pairs removed before the broadphase.
// Outputs: pairs [pair_capacity][2] (canonical A < B order), raw contacts
// [raw_capacity][10] (normal A->B, separation, point A, point B) with
// raw_u32 [raw_capacity][3] (feature A, feature B, index into pairs), and the
// counts. Returns 0, the reference's MRStepStatusCode (> 0), or -1 (err).
int mr_ref_collide(uint32_t shape_count, const MRRefShape* shapes, uint32_t body_count, const void* bodies,
                   uint32_t exclusion_count, const uint32_t* exclusions, uint32_t pair_capacity,
                   uint32_t raw_capacity, uint32_t* pairs, uint32_t* pair_count, float* raw, uint32_t* raw_u32,
                   uint32_t* raw_count, char* err, size_t errlen) {
    try {
        static_assert(sizeof(MRBodyStateGPU) == 128, "body state layout");
        static_assert(sizeof(MRRefShape) == 96, "shape layout");
        std::vector<MRShapeGPU> sh(shape_count);
        for (uint32_t s = 0; s < shape_count; ++s) {
            const MRRefShape& in = shapes[s];
            MRShapeGPU& o = sh[s];
            std::memset(&o, 0, sizeof(o));
            o.bodyIndex = in.body;
            o.shapeType = in.shape_type;
            o.materialIndex = in.material;
            o.flags = in.flags;
            o.collisionGroup = in.collision_group;
            o.collisionMask = in.collision_mask;
            o.slotGeneration = in.slot_generation;
            o.localPosition = q4(in.local_position);
            o.localRotation = q4(in.local_rotation);
            o.dimensions = q4(in.dimensions);
            o.contactRestAndBoundingRadius = q4(in.contact_rest_radius);
        }
        std::vector<MRBodyStateGPU> bs(body_count);
        std::memcpy(bs.data(), bodies, (size_t)body_count * sizeof(MRBodyStateGPU));
        std::vector<CollisionPairExclusion> ex(exclusion_count);
        for (uint32_t i = 0; i < exclusion_count; ++i) ex[i] = CollisionPairExclusion{exclusions[2 * i], exclusions[2 * i + 1]};
        CollisionConfig cfg;
        cfg.environment = 0u;
        cfg.capacities.pairCapacity = pair_capacity;
        cfg.capacities.rawContactCapacity = raw_capacity;
        cfg.capacities.manifoldCapacity = pair_capacity;
        PersistentManifoldCache cache;
        const CollisionFrame f = collideCpuReference(sh, bs, cfg, cache, ex);
        if (!f.succeeded()) {
            set_error(err, errlen, "collision status " + std::to_string((uint32_t)f.diagnostics.code));
            return (int)f.diagnostics.code;
        }
        *pair_count = (uint32_t)f.pairs.size();
        for (size_t p = 0; p < f.pairs.size(); ++p) {
            pairs[2 * p] = f.pairs[p].colliderA;
            pairs[2 * p + 1] = f.pairs[p].colliderB;
        }
        *raw_count = (uint32_t)f.rawContacts.size();
        for (size_t c = 0; c < f.rawContacts.size(); ++c) {
            const MRRawContactGPU& r = f.rawContacts[c];
            float* o = raw + c * 10u;
            o[0] = r.normalAndSeparation.x;
            o[1] = r.normalAndSeparation.y;
            o[2] = r.normalAndSeparation.z;
            o[3] = r.normalAndSeparation.w;
            o[4] = r.pointAWorld.x;
            o[5] = r.pointAWorld.y;
            o[6] = r.pointAWorld.z;
            o[7] = r.pointBWorld.x;
            o[8] = r.pointBWorld.y;
            o[9] = r.pointBWorld.z;
            raw_u32[3 * c] = r.featureAndFlags[0];
            raw_u32[3 * c + 1] = r.featureAndFlags[1];
            raw_u32[3 * c + 2] = f.rawContactPairIndices[c];
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Exact circular-cone solve of one dense problem: inverse_mass [nv][nv]
// row-major, free_velocity [nv], jacobian [contacts][3][nv] (normal, tangent u,
// tangent v rows), target / regularization / warm [contacts][3], friction
// [contacts]. Outputs velocity [nv], impulses [contacts][3] and
// info [4] = (iterations, objective, scaled optimality residual, max cone
// violation). Returns 0, the status code (> 0), or -1 (err).
int mr_ref_conic_solve(uint32_t nv, const double* inverse_mass, const double* free_velocity, uint32_t contacts,
                       const double* jacobian, const double* target, const double* regularization,
                       const double* warm, const double* friction, uint32_t max_iterations, double tolerance,
                       double* velocity, double* impulses, double* info, char* err, size_t errlen) {
    try {
        DenseConicProblem p;
        p.nv = nv;
        p.inverseMass.assign(inverse_mass, inverse_mass + (size_t)nv * nv);
        p.freeVelocity.assign(free_velocity, free_velocity + nv);
        for (uint32_t c = 0; c < contacts; ++c) {
            DenseContactBlock b;
            const double* J = jacobian + (size_t)c * 3u * nv;
            b.normalJacobian.assign(J, J + nv);
            b.tangentUJacobian.assign(J + nv, J + 2u * nv);
            b.tangentVJacobian.assign(J + 2u * nv, J + 3u * nv);
            for (int k = 0; k < 3; ++k) {
                b.targetVelocity[k] = target[3 * c + k];
                b.regularization[k] = regularization[3 * c + k];
                b.warmImpulse[k] = warm != nullptr ? warm[3 * c + k] : 0.0;
            }
            b.friction = friction[c];
            p.contacts.push_back(std::move(b));
        }
        ReferenceConicSolverConfig cfg;
        cfg.maximumIterations = max_iterations;
        cfg.optimalityTolerance = tolerance;
        cfg.enableWarmStart = warm != nullptr;
        const ReferenceConicSolution s = solveReferenceConicProblem(p, cfg);
        info[0] = (double)s.iterations;
        info[1] = s.objective;
        info[2] = s.scaledOptimalityResidual;
        info[3] = s.maximumConeViolation;
        if (!s.converged()) {
            set_error(err, errlen, "conic solve status " + std::to_string((uint32_t)s.code));
            return (int)s.code;
        }
        for (uint32_t i = 0; i < nv; ++i) velocity[i] = s.velocity[i];
        for (size_t i = 0; i < s.impulses.size(); ++i) impulses[i] = s.impulses[i];
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

static_assert(sizeof(MRBodyStateGPU) == 128, "MRBodyStateGPU layout");
static_assert(sizeof(MRContactConstraintGPU) == 144, "MRContactConstraintGPU layout");

// Projected Gauss-Seidel contact solve of free bodies (solveContactConstraints,
// ConstraintSolver.cpp:1047-1233), in place: bodies [body_count] and contacts
// [contact_count] use the reference's MRBodyStateGPU / MRContactConstraintGPU
// byte layouts; velocities and impulses are updated. Runs exactly `iterations`
// sweeps (deterministic mode, no early exit). info [4] = (iterations, last
// sweep's maximum impulse change, maximum normal residual, maximum cone
// violation). Returns the reference's status code (0 or a failure code).
int mr_ref_pgs_solve(uint32_t body_count, void* bodies, uint32_t contact_count, void* contacts, double timestep,
                     double error_reduction, double penetration_slop, double max_depenetration_velocity,
                     uint32_t iterations, double* info, char* err, size_t errlen) {
    try {
        ContactSolverConfig cfg;
        cfg.timestep = timestep;
        cfg.errorReduction = error_reduction;
        cfg.penetrationSlop = penetration_slop;
        cfg.maxDepenetrationVelocity = max_depenetration_velocity;
        cfg.velocityIterations = iterations;
        cfg.enableWarmStart = false;
        cfg.deterministic = true;
        std::span<MRBodyStateGPU> b(static_cast<MRBodyStateGPU*>(bodies), body_count);
        std::span<MRContactConstraintGPU> c(static_cast<MRContactConstraintGPU*>(contacts), contact_count);
        const ContactSolverDiagnostics d = solveContactConstraints(b, c, cfg);
        info[0] = (double)d.iterations;
        info[1] = d.maximumImpulseDelta;
        info[2] = d.maximumNormalResidual;
        info[3] = d.maximumConeViolation;
        if (!d.succeeded()) {
            set_error(err, errlen, "contact solve status " + std::to_string((uint32_t)d.code));
            return (int)d.code;
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

}  // extern "C"
