// Launch record of the fused control-step kernel, shared by the host
// executor and the device code.
#pragma once

#include "../../include/mr_mlp.h"
#include "../../include/mr_world_types.h"

#include <stdint.h>

namespace mr {

// Stage indices of the optional per-stage cycle counters.
enum StageIndex : uint32_t {
    kStageKinematics = 0,
    kStageABA,
    kStageProjection,
    kStageBroadphase,
    kStageNarrowphase,
    kStageEmission,
    kStageResponses,
    kStageEvaluation,
    kStageWarmStart,
    kStageSweeps,
    kStageResiduals,
    kStageIntegration,
    kStageCount = 16
};

// Endpoint kinds of a constraint, decoded once at emission.
enum HotKind : uint32_t {
    kHotArticulated = 1u, // an endpoint is an articulation link
    kHotDynamicA = 2u,    // body_a is a dynamic scene body
    kHotDynamicB = 4u,    // body_b is a dynamic scene body
};

// Per-constraint data of the cone solve (global scratch, read-only during
// the sweeps); the impulses themselves live in LDS. The first 64 bytes are
// what every sweep reads (the tangent rows' targets and regularisation are
// always zero); the geometry behind them is read only for constraints with a
// dynamic scene endpoint.
struct HotConstraint {
    float winv[9];       // 3x3 block inverse (J M^-1 J^T while it is evaluated)
    float target0, reg0; // normal-row target velocity and regularisation
    float mu_u, mu_v;
    float stage[3];      // relative velocity rows between the evaluation passes
    uint32_t body_a, body_b;
    uint32_t flags, kind; // kind: HotKind bits
    float dir[9];        // normal, tangent u, tangent v
    float point[3];      // contact midpoint
};

// The leading 64 bytes of a HotConstraint.
struct SolveBlock {
    float winv[9];
    float target0, reg0, mu_u, mu_v;
    float stage[3];
};

// The trailing 64 bytes of a HotConstraint: what the sweeps need for a
// contact with a dynamic scene endpoint.
struct GeoBlock {
    uint32_t body_a, body_b;
    uint32_t flags, kind;
    float dir[9];
    float point[3];
};

// Per-constraint data produced by emission and consumed by the Jacobian and
// IR evaluation passes (global scratch).
struct ColdConstraint {
    float pa[3], pb[3];
    float separation;
    uint32_t manifold, point, flags;
    float friction[4];
    float restitution, restitution_threshold, compliance, dissipation;
    float lambda[3];
    float pad;
};

// Tree topology of one environment's model, cached in LDS once per launch so
// that tree walks never chase parent indices through global memory. Link and
// DoF indices fit in a byte (0xff = none); masks are bit sets over links.
struct TopoCache {
    unsigned long long anc[MR_MAX_BODIES];      // links on the path link -> root, root excluded
    unsigned long long children[MR_MAX_BODIES]; // direct child links
    uint8_t parent[MR_MAX_BODIES];
    uint8_t vidx[MR_MAX_BODIES];
    uint8_t qidx[MR_MAX_BODIES];
    uint8_t jtype[MR_MAX_BODIES];
    uint8_t depth[MR_MAX_BODIES];
    uint8_t motion[MR_MAX_BODIES];              // motion type of every body
    uint8_t dof_body[MR_MAX_BODIES];            // owning link of every DoF
    uint8_t slot[MR_MAX_BODIES];                // outward-pass slot of each link (0xff: leaf)
};

#if defined(__HIPCC__)
#define MR_HOST_DEVICE __host__ __device__
#else
#define MR_HOST_DEVICE
#endif

// Slot plan of the response outward pass. Links are visited in index order
// (parents first); a link's spatial acceleration must stay available until
// its last child has read it, so it takes the lowest free slot and frees it
// at that child. Leaves take none. `parent(i)` gives the parent link of
// i >= 1. Returns the number of slots used at the peak.
template <class ParentFn>
MR_HOST_DEVICE inline uint32_t plan_response_slots(ParentFn parent, uint32_t n, uint8_t* slot) {
    uint8_t last_child[MR_MAX_BODIES];
    for (uint32_t i = 0; i < n; ++i) last_child[i] = 0xffu;
    for (uint32_t i = 1; i < n; ++i) last_child[parent(i)] = (uint8_t)i;
    unsigned long long used = 0ull;
    uint32_t peak = 0;
    auto take = [&]() -> uint8_t {
        uint32_t s = 0;
        while ((used >> s) & 1ull) ++s;
        used |= 1ull << s;
        peak = s + 1u > peak ? s + 1u : peak;
        return (uint8_t)s;
    };
    slot[0] = last_child[0] != 0xffu ? take() : (uint8_t)0xffu;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t p = parent(i);
        if (last_child[p] == i) used &= ~(1ull << slot[p]);
        slot[i] = last_child[i] != 0xffu ? take() : (uint8_t)0xffu;
    }
    return peak;
}

// Joint constants of one link for forward kinematics, staged in LDS once per
// launch: parent-side joint rotation, conjugated child-side rotation,
// normalized joint axis and the two anchors.
struct LinkConst {
    float prot[4];
    float crot_conj[4];
    float axis[3];
    float panchor[3];
    float canchor[3];
};

// Byte offsets of the per-workgroup dynamic LDS arrays.
// One heightfield query result (32 bytes): unit normal, signed surface
// distance (NaN when the point is off the grid), surface point, triangle
// feature key.
struct HfHit {
    float n[3];
    float d;
    float p[3];
    uint32_t feature;
};
static_assert(sizeof(HfHit) == 32, "HfHit layout");
// Heightfield query points staged per control-step microstep: a humanoid on
// terrain needs at most 64 (18 spheres, 10 capsules x 3, 2 boxes x 8).
constexpr uint32_t kHfSlots = 128u;

struct LdsLayout {
    uint32_t topo, lconst, arm;
    uint32_t q, v, vcand, effort, accel, cmd;
    uint32_t pos, rot, w, vl, sa, sl, ca, cl, r;
    uint32_t I, pt, pb, Ut, Ub, D, u, alpha, acc, L;
    uint32_t scene, col, active, npts, moff, okey;
    uint32_t hfitem, hfq;  // heightfield query table (kHfSlots entries each), collision phase only
    uint32_t lam;    // [constraint_capacity][3] floats: cone impulses
    uint32_t cpair;  // [constraint_capacity]: body_a | body_b << 8 | kind << 16
    // response scratch, overlaying the collision arrays (dead after emission)
    uint32_t rslot;  // [slots][6][64] floats: outward spatial accelerations
    uint32_t rstack; // [stack][64] floats: joint terms of the inward path
    uint32_t isl;    // [MR_MAX_BODIES + 1] u8 island parents + used flags (after responses)
    uint32_t resid;  // [2][constraint_capacity] floats: per-constraint residual terms (after responses)
    uint32_t jrelax; // [constraint_capacity] floats: relaxed-Jacobi damping per contact (after responses)
    uint32_t jdelta; // [constraint_capacity][3] floats: relaxed-Jacobi impulse deltas (after responses)
    // MLP instance of the kernel only (past the plain layout's end):
    uint32_t mlp;        // warm-start MLP tile scratch (kMlpTileLdsBytes)
    uint32_t mlp_lam;    // [constraint_capacity][3] floats: the plain warm start, kept for the fallback
    uint32_t mlp_scene;  // [max_scene_bodies] MRBodyStateC: scene bodies before the warm start
    uint32_t total;
};

struct KernelArgs {
    // immutable model
    const MRModelHeaderC* models;
    const uint32_t* environment_model;
    const uint8_t* response_slots;  // [model_count][MR_MAX_BODIES]: plan_response_slots per model
    const MRBodyC* bodies;
    const MRDofC* dofs;
    const MRActuatorProfileC* actuators;  // parallel to dofs, or null: every DoF inactive
    const MRShapeC* shapes;
    const MRMaterialC* materials;
    const MRCompiledPairC* pairs;
    const float* heights;
    MRHeightfieldC heightfield;
    // convex and mesh geometry, shared by every model block (null without any)
    const MRGeometryC* geometries;
    const mr_float4* geometry_vertices;
    const MRTriangleC* triangles;
    const MRBvhNodeC* bvh_nodes;
    uint32_t geometry;  // 1: some shape is a convex hull or a mesh (the kGeo instances run)
    float gravity[4];
    MRWorldStepConfigC cfg;
    float h;
    uint32_t environment_count;
    uint32_t max_nq, max_nv, max_scene;
    uint32_t row_stride;  // floats per Jacobian / response row: max(1, max_nv)
    uint32_t control_step;
    uint32_t pair_capacity, manifold_capacity, constraint_capacity;
    // persistent state
    float* q;
    float* v;
    MRBodyStateC* scene;
    MRManifoldHeaderC* mh;   // [3][E][manifold_capacity]
    MRManifoldPointC* mp;    // [3][E][manifold_capacity][4]
    uint32_t* mcount;        // [3][E]
    uint32_t* mcurrent;      // [E]
    // inputs of this control step
    const float* command;       // [E][max_nv] ([horizon][E][max_nv] for the horizon instance)
    const uint32_t* reset_mask; // [E] or null ([horizon][E] for the horizon instance)
    // work-queue instance only: environments in processing order, and the
    // head every workgroup takes its next entry from (zeroed before a launch)
    const uint32_t* queue;
    uint32_t queue_count;
    uint32_t* queue_head;
    const float* reset_q;
    const float* reset_v;
    const MRBodyStateC* reset_scene;
    // outputs of this control step
    MRWorldStatusC* status;     // [E] ([horizon][E] for the horizon instance)
    MRContactStatusC* cstatus;  // [E] ([horizon][E] for the horizon instance)
    float* step_us;             // [horizon][E] wall-clock microseconds per control step of each world, or null
    uint32_t horizon;           // control steps per launch of the horizon instance
    float us_per_tick;          // microseconds per tick of the constant-rate clock (s_memrealtime)
    float* acceleration;        // [E][max_nv] or null
    MRPairRecordC* evidence_pairs; // [E][pair_capacity] or null
    uint32_t* evidence_counts;     // [E] or null
    // global scratch
    MRManifoldHeaderC* stage_h;   // [E][pair_capacity]
    MRManifoldPointC* stage_p;    // [E][pair_capacity][4]
    float* jrow;                  // [E][constraint_capacity][3][row_stride]
    float* resp;                  // [E][constraint_capacity][3][row_stride]
    ColdConstraint* cold;         // [E][constraint_capacity]
    HotConstraint* hot;           // [E][constraint_capacity]
    unsigned long long* stage_cycles; // [E][kStageCount] shader-clock cycles per stage, or null
    LdsLayout lds;
    // warm-start MLP (read by the MLP instance of the kernel only)
    MRMlpWeightsC mlp;     // device weights, used when mlp_enabled
    uint32_t mlp_enabled;
    float* mlp_samples;    // [E][constraint_capacity][kMlpSampleFloats] or null: capture at cfg.capture_substep
};

// LDS bytes of one warm-start MLP tile: 16 x 8 FP32 inputs and two
// 16 x 264 BF16 hidden layers.
constexpr uint32_t kMlpTileLdsBytes = 16u * 8u * 4u + 2u * 16u * 264u * 2u;
// One captured sample: 8 features, the converged impulse (3), a valid flag,
// the warm start the cone solve began from (3), 1.0 where the network
// replaced the plain warm start (0.0 otherwise), and with inference on the
// network's scaled correction (3) and 1.0 where every normalized feature was
// within bounds (0.0 and a zero correction otherwise).
constexpr uint32_t kMlpSampleFloats = 20u;

// Device-side summary of one advance (all steps x environments), reduced on
// the GPU so that waiting for an advance reads back 24 bytes instead of
// every status record.
struct AdvanceSummaryDev {
    unsigned long long first_failure; // (step << 40) | (environment << 8) | code; ~0 if none
    uint32_t failed;
    uint32_t max_active_pairs;
    uint32_t max_constraints;
    uint32_t pad;
};

} // namespace mr
