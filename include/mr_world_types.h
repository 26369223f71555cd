/*
 * Pointer-free records shared by the C ABI, the HIP kernels and the CPU
 * oracle of the batched temporalCone microstep.
 *
 * Layouts follow the reference's shared CPU/GPU records where a record
 * crosses the boundary unchanged:
 *   MRBodyStateC        <- MRBodyStateGPU        include/metalrobo/engine_types.h:418-434
 *   MRShapeC            <- MRShapeGPU            include/metalrobo/engine_types.h:1430-1455
 *   MRManifoldHeaderC   <- MRManifoldHeaderGPU   include/metalrobo/engine_types.h:1503-1512
 *   MRManifoldPointC    <- MRManifoldPointGPU    include/metalrobo/engine_types.h:1514-1521
 *   MRWorldStatusC      <- MRMetalWorldStatusGPU include/metalrobo/engine_types.h:726-741
 *   MRStepStatusCode    <- MRStepStatusCode      include/metalrobo/engine_types.h:171-186
 * The model records (MRBodyC, MRDofC, MRMaterialC, MRCompiledPairC) fold the
 * reference's articulation/joint/dof/body-property records into one body
 * record per link, because this engine keeps exactly one articulation per
 * world (the temporalCone single-articulation graph).
 *
 * Every record is a multiple of 16 bytes; sizes are asserted below.
 */
#ifndef MR_WORLD_TYPES_H
#define MR_WORLD_TYPES_H

#include <stddef.h>
#include <stdint.h>

#define MR_WORLD_ABI_VERSION 1u
#define MR_INVALID_INDEX 0xffffffffu

/* Per-world topology capacities of the fused kernel (one wavefront per world). */
#define MR_MAX_BODIES 64u          /* articulation links + scene bodies */
#define MR_MAX_LINKS 48u           /* articulation links incl. root */
#define MR_MAX_DOFS 48u            /* generalized velocities */
#define MR_MAX_Q 49u               /* generalized positions */
#define MR_MAX_SHAPES 64u
#define MR_MAX_LEVELS 48u

#define MR_MANIFOLD_POINT_CAPACITY 4u
#define MR_RAW_CONTACTS_PER_PAIR 8u
#define MR_MAX_PHYSICS_SUBSTEPS 64u

/* Conservative relative AABB / query pad, 2^-17 (engine_types.h:84-85). */
#define MR_COLLISION_AABB_RELATIVE_PAD 0.00000762939453125f
#define MR_COLLISION_QUERY_RELATIVE_PAD 0.00000762939453125f

enum MRShapeType {
    MR_SHAPE_SPHERE = 0u,
    MR_SHAPE_CAPSULE = 1u,
    MR_SHAPE_BOX = 2u,
    MR_SHAPE_PLANE = 3u,
    MR_SHAPE_CONVEX = 5u,         /* convex hull of a geometry's vertices (gpu_types.h:41-50) */
    MR_SHAPE_TRIANGLE_MESH = 6u,  /* static triangle mesh with a BVH4 */
    MR_SHAPE_HEIGHTFIELD = 7u
};

/* MR_SHAPE_FLAG_ENABLE_CCD opts a collider into the hybrid mode's continuous
 * collision query (MRShapeGPU flags, engine_types.h; Collision.metal:7205). */
enum MRShapeFlags {
    MR_SHAPE_FLAG_SIMULATION_DISABLED = 1u,
    MR_SHAPE_FLAG_ENABLE_CCD = 2u,
    MR_SHAPE_FLAG_MESH_TWO_SIDED = 4u /* triangle meshes are one-sided unless set */
};

/* MetalWorldCCDMode (MetalWorld.hpp:420): speculative contacts only, or
 * speculative contacts plus microsteps split at the earliest time of impact
 * of CCD-enabled pairs. */
enum MRCcdMode { MR_CCD_SPECULATIVE = 0u, MR_CCD_HYBRID = 1u };
#define MR_CCD_MAX_EVENTS 16u

enum MRMotionType {
    MR_MOTION_STATIC = 0u,
    MR_MOTION_KINEMATIC = 1u,
    MR_MOTION_DYNAMIC = 2u
};

enum MRRootType { MR_ROOT_FIXED = 0u, MR_ROOT_FLOATING = 1u };

enum MRJointType {
    MR_JOINT_REVOLUTE = 0u,
    MR_JOINT_CONTINUOUS = 2u,
    MR_JOINT_FIXED = 5u,
    MR_JOINT_ROOT = 15u /* body 0 of an articulation */
};

/* Narrowphase classes; numbering follows MRCollisionPairClass
 * (engine_types.h:139-153); heightfield pairs use the surface class. */
enum MRPairClass {
    MR_PAIR_UNSUPPORTED = 0u,
    MR_PAIR_SPHERE_SPHERE = 1u,
    MR_PAIR_SPHERE_PLANE = 2u,
    MR_PAIR_CAPSULE_PLANE = 3u,
    MR_PAIR_BOX_PLANE = 4u,
    MR_PAIR_SPHERE_CAPSULE = 6u,
    MR_PAIR_CAPSULE_CAPSULE = 7u,
    MR_PAIR_SPHERE_BOX = 8u,
    MR_PAIR_CAPSULE_BOX = 9u,
    MR_PAIR_BOX_BOX = 10u,
    MR_PAIR_CONVEX = 11u,      /* a convex hull with a sphere, capsule, box, plane or hull */
    MR_PAIR_HEIGHTFIELD = 12u,
    MR_PAIR_MESH = 13u         /* a static triangle mesh with a sphere, capsule, box or hull */
};

/* engine_types.h:171-186 */
enum MRStepStatusCode {
    MR_STEP_SUCCESS = 0u,
    MR_STEP_FIXED_BUDGET_COMPLETE = 1u,
    MR_STEP_NONFINITE_INPUT = 2u,
    MR_STEP_NONFINITE_RESULT = 3u,
    MR_STEP_PAIR_CAPACITY_OVERFLOW = 4u,
    MR_STEP_CONTACT_CAPACITY_OVERFLOW = 5u,
    MR_STEP_MANIFOLD_CAPACITY_OVERFLOW = 6u,
    MR_STEP_CONSTRAINT_CAPACITY_OVERFLOW = 7u,
    MR_STEP_ISLAND_CAPACITY_OVERFLOW = 8u,
    MR_STEP_CCD_EVENT_BUDGET_EXHAUSTED = 9u,
    MR_STEP_FACTORIZATION_FAILED = 10u,
    MR_STEP_DID_NOT_CONVERGE = 11u,
    MR_STEP_UNSUPPORTED = 12u,
    MR_STEP_CCD_CAPACITY_OVERFLOW = 13u,
    MR_STEP_STATUS_COUNT = 14u
};

/* engine_types.h:268-273 */
enum MRDofFlags {
    MR_DOF_FLAG_ROOT = 1u << 0u,
    MR_DOF_FLAG_ACTUATED = 1u << 1u,
    MR_DOF_FLAG_POSITION_LIMIT = 1u << 2u,
    MR_DOF_FLAG_VELOCITY_LIMIT = 1u << 3u,
    MR_DOF_FLAG_EFFORT_LIMIT = 1u << 4u,
    MR_DOF_FLAG_DRIVE = 1u << 5u
};

enum MRActuationMode {
    MR_ACTUATION_EFFORT = 0u,
    MR_ACTUATION_IMPLICIT_POSITION_DRIVE = 1u
};

/* Constraint flags (engine_types.h MRConstraintFlags). */
enum MRConstraintFlags {
    MR_CONSTRAINT_FLAG_NEW_IMPACT = 1u << 0u,
    MR_CONSTRAINT_FLAG_WARM_STARTED = 1u << 1u
};

typedef struct mr_float4 {
    float x, y, z, w;
} mr_float4;

/* ---------------------------------------------------------------- model -- */

/* One body of a world. Bodies [0, link_count) form the single articulation
 * (body 0 is its root, parent < child); bodies [link_count, body_count) are
 * scene bodies (free dynamic, or static). The joint fields describe the
 * inbound joint of a link. Anchors are relative to each body's COM-centred
 * origin, matching the reference's joint convention (NUMERICS.md "Joint
 * anchors are expressed from each body's COM"). */
typedef struct MRBodyC {
    uint32_t parent;       /* link index of the parent, MR_INVALID_INDEX for root/scene */
    uint32_t joint_type;   /* MRJointType */
    uint32_t q_index;      /* local q coordinate of the 1-DoF joint, else invalid */
    uint32_t v_index;      /* local v coordinate of the 1-DoF joint, else invalid */
    uint32_t motion_type;  /* MRMotionType */
    uint32_t depth;        /* tree depth (root 0); scene bodies 0 */
    uint32_t flags;
    uint32_t reserved;
    mr_float4 mass;        /* x mass, y inverse mass */
    mr_float4 inertia_row0;          /* body-frame inertia about the COM */
    mr_float4 inertia_row1;
    mr_float4 inertia_row2;
    mr_float4 inverse_inertia_row0;
    mr_float4 inverse_inertia_row1;
    mr_float4 inverse_inertia_row2;
    mr_float4 axis;            /* joint axis in the joint frame */
    mr_float4 parent_anchor;   /* joint point in the parent's COM frame */
    mr_float4 child_anchor;    /* joint point in this body's COM frame */
    mr_float4 parent_rotation; /* joint frame in the parent (x,y,z,w) */
    mr_float4 child_rotation;  /* joint frame in this body (x,y,z,w) */
    /* linear damping, angular damping, max linear speed, max angular speed */
    mr_float4 damping;
    /* precomputed per-microstep exponential damping scales exp(-h*d) for
     * scene-body prediction: x linear, y angular (written at world creation) */
    mr_float4 damping_scale;
} MRBodyC;

/* One record per generalized velocity coordinate (engine_types.h:357-384). */
typedef struct MRDofC {
    uint32_t body;     /* link owning the coordinate */
    uint32_t q_index;  /* scalar configuration coordinate or MR_INVALID_INDEX */
    uint32_t flags;    /* MRDofFlags */
    uint32_t reserved;
    mr_float4 limits;  /* lower, upper, maximum velocity, maximum effort */
    mr_float4 drive;   /* stiffness, damping, armature, dry friction */
} MRDofC;

/* Optional motor / transmission record per generalized velocity
 * (MRActuatorProfileGPU, engine_types.h:386-395, 48 bytes). An inactive
 * record (flags without MR_ACTUATOR_PROFILE_ACTIVE) must be all zero; it
 * executes as no-load speed FLT_MAX, efficiency 1 and stall torque equal to
 * the DoF's effort limit (FLT_MAX without one), as MetalWorld.mm:1376-1410
 * expands it. Every command is clamped to the torque-speed envelope
 *   min(limits.w, stall * efficiency * (1 - clamp(|v| / no_load_speed, 0, 1)))
 * (MetalWorld.metal:252-269, 345-365), so a DoF without an effort limit
 * receives no effort. */
enum MRActuatorProfileFlags {
    MR_ACTUATOR_PROFILE_ACTIVE = 1u << 0u,
    MR_ACTUATOR_PROFILE_CALIBRATED = 1u << 1u
};

typedef struct MRActuatorProfileC {
    /* joint-side torque constant N m/A, current limit A, no-load speed rad/s, efficiency [0, 1] */
    mr_float4 motor_and_speed;
    /* backlash play, command delay s, stall torque N m (= torque constant * current limit), reserved */
    mr_float4 transmission_and_envelope;
    uint32_t identity[4]; /* model-local v index, MRActuatorProfileFlags, reserved, reserved */
} MRActuatorProfileC;

/* engine_types.h:1430-1455 */
typedef struct MRShapeC {
    uint32_t body;
    uint32_t shape_type;
    uint32_t material;
    uint32_t flags;
    uint32_t collision_group;
    uint32_t collision_mask;
    uint32_t slot_generation;
    uint32_t geometry;         /* convex and mesh shapes: index into the model's geometry table */
    mr_float4 local_position;
    mr_float4 local_rotation;
    /* Sphere: x=radius. Capsule: x=radius, y=half length (local +Y).
     * Box: xyz=half extents. Convex and mesh: xyz=scale of the geometry's
     * vertices. Heightfield: unused (the world header holds it). */
    mr_float4 dimensions;
    /* contact offset, rest offset, bounding radius, reserved */
    mr_float4 contact_rest_radius;
} MRShapeC;

/* Cooked geometry of convex and mesh shapes (MRGeometryHeaderGPU,
 * engine_types.h:1360-1383, restated in a compact layout). Vertices are
 * local, before the shape's scale. The model compiler fills the bounds and a
 * mesh's BVH4 (node_offset, node_count); callers leave them zero. */
#define MR_MAX_CONVEX_VERTICES 64u
#define MR_MESH_LEAF_TRIANGLES 4u
typedef struct MRGeometryC {
    uint32_t kind;            /* MR_SHAPE_CONVEX or MR_SHAPE_TRIANGLE_MESH */
    uint32_t flags;
    uint32_t vertex_offset;   /* into the vertex table */
    uint32_t vertex_count;    /* convex: 4 .. MR_MAX_CONVEX_VERTICES */
    uint32_t triangle_offset; /* meshes: into the triangle table */
    uint32_t triangle_count;
    uint32_t node_offset;     /* meshes: into the BVH node table (model compiler) */
    uint32_t node_count;
    mr_float4 local_lower;    /* vertex bounds (model compiler) */
    mr_float4 local_upper;
} MRGeometryC;

/* One mesh triangle: vertex indices local to its geometry, and the stable
 * feature id its contacts carry (MRMeshTriangleGPU, engine_types.h:1411-1418). */
typedef struct MRTriangleC {
    uint32_t vertices[3];
    uint32_t feature;
} MRTriangleC;

/* BVH4 node: the four children's bounds in structure-of-arrays form. A child
 * is an inner node (count 0: child = node index within the geometry), a leaf
 * (count 1 .. MR_MESH_LEAF_TRIANGLES triangles from child, local to the
 * geometry) or empty (child = MR_INVALID_INDEX). */
typedef struct MRBvhNodeC {
    mr_float4 lower_x, lower_y, lower_z;
    mr_float4 upper_x, upper_y, upper_z;
    uint32_t child[4];
    uint32_t count[4];
} MRBvhNodeC;

/* engine_types.h:1349-1356, with per-axis coefficients for elliptic cones. */
typedef struct MRMaterialC {
    mr_float4 friction; /* static u, static v, dynamic u, dynamic v */
    mr_float4 response; /* restitution, restitution threshold, compliance, dissipation */
} MRMaterialC;

/* engine_types.h:761-766 */
typedef struct MRCompiledPairC {
    uint32_t collider_a;
    uint32_t collider_b;
    uint32_t pair_class;
    uint32_t reserved;
} MRCompiledPairC;

/* One model block. A world references one model per environment; worlds with
 * identical robots share one block, randomized worlds own one each. All
 * offsets index the flat arrays of MRWorldModelDescC. */
typedef struct MRModelHeaderC {
    uint32_t body_offset;
    uint32_t body_count;
    uint32_t link_count;
    uint32_t root_type;     /* MRRootType */
    uint32_t dof_offset;
    uint32_t nq;
    uint32_t nv;
    uint32_t shape_offset;
    uint32_t shape_count;
    uint32_t pair_offset;
    uint32_t pair_count;
    uint32_t level_count;   /* 1 + maximum link depth */
    uint32_t scene_body_count;
    uint32_t material_offset;
    uint32_t material_count;
    uint32_t reserved;
} MRModelHeaderC;

/* Shared static heightfield (one per world set): axis-aligned, z up, vertices
 * on a regular grid, cell (ix,iy) split into triangles (v00,v10,v11) and
 * (v00,v11,v01). */
typedef struct MRHeightfieldC {
    uint32_t nx;
    uint32_t ny;
    uint32_t present;
    uint32_t reserved;
    mr_float4 origin_spacing; /* x,y origin of vertex (0,0); z spacing; w reserved */
    mr_float4 bounds_lower;   /* xyz lower corner of the vertex AABB */
    mr_float4 bounds_upper;
} MRHeightfieldC;

typedef struct MRWorldModelDescC {
    uint32_t model_count;
    uint32_t environment_count;
    uint32_t max_nq;
    uint32_t max_nv;
    uint32_t max_scene_bodies;
    uint32_t max_shapes;
    uint32_t max_pairs;
    uint32_t reserved;
    mr_float4 gravity;                 /* xyz */
    const MRModelHeaderC* models;      /* [model_count] */
    const uint32_t* environment_model; /* [environment_count] */
    const MRBodyC* bodies;
    const MRDofC* dofs;
    const MRShapeC* shapes;
    const MRMaterialC* materials;
    const MRCompiledPairC* pairs;
    uint32_t body_total;
    uint32_t dof_total;
    uint32_t shape_total;
    uint32_t material_total;
    uint32_t pair_total;
    uint32_t heightfield_vertex_count;
    MRHeightfieldC heightfield;
    const float* heightfield_heights;  /* [ny][nx] */
    /* [dof_total], parallel to dofs, or null: every DoF inactive */
    const MRActuatorProfileC* actuators;
    uint32_t actuator_total;           /* dof_total, or 0 with actuators null */
    uint32_t reserved_actuators;
    /* geometry of convex and mesh shapes, shared by every model block */
    const MRGeometryC* geometries;     /* [geometry_total] */
    const mr_float4* geometry_vertices; /* [geometry_vertex_total], xyz */
    const MRTriangleC* triangles;      /* [triangle_total] */
    /* the meshes' BVH4 nodes: filled by mr_world_model_describe (with the
     * triangles in leaf order), ignored by mr_world_model_create, which builds them */
    const MRBvhNodeC* bvh_nodes;
    uint32_t geometry_total;
    uint32_t geometry_vertex_total;
    uint32_t triangle_total;
    uint32_t bvh_node_total;
} MRWorldModelDescC;

/* ---------------------------------------------------------------- state -- */

/* engine_types.h:418-434 (128 bytes) */
typedef struct MRBodyStateC {
    mr_float4 position;                 /* xyz COM, w = 1 */
    mr_float4 orientation;              /* body-to-world quaternion (x,y,z,w) */
    mr_float4 linear_velocity_inv_mass; /* xyz COM velocity, w inverse mass */
    mr_float4 angular_velocity;         /* xyz world angular velocity */
    mr_float4 inverse_inertia_world_row0;
    mr_float4 inverse_inertia_world_row1;
    mr_float4 inverse_inertia_world_row2;
    uint32_t flags_indices[4];          /* motion type, articulation, body, flags */
} MRBodyStateC;

/* engine_types.h:1503-1512 */
typedef struct MRManifoldHeaderC {
    uint32_t pair_and_count[4];        /* environment, collider A, collider B, points */
    uint32_t generations_and_flags[4]; /* generation A, generation B, patch, flags */
    mr_float4 normal_and_age;          /* body-A-local normal, frames since rebuild */
    mr_float4 tangent_and_metric;      /* body-A-local tangent, breaking metric */
} MRManifoldHeaderC;

/* engine_types.h:1514-1521 */
typedef struct MRManifoldPointC {
    mr_float4 local_anchor_a;
    mr_float4 local_anchor_b;
    /* normal, tangent u, tangent v; w: 0, or in the hybrid CCD mode the duration (s)
     * of the segment whose solve produced them (a warm start scales them by the
     * next segment's duration over it) */
    mr_float4 impulses;
    uint32_t feature_and_life[4]; /* feature A, feature B, lifetime, flags */
} MRManifoldPointC;

/* ------------------------------------------------------------- statuses -- */

/* engine_types.h:726-741 (48 bytes) */
typedef struct MRWorldStatusC {
    uint32_t code;
    uint32_t environment;
    uint32_t control_step;
    uint32_t successful_substeps;
    uint32_t aba_code;
    uint32_t failing_substep;
    uint32_t failing_index;
    uint32_t flags;
    /* minimum ABA pivot, maximum ABA pivot, max |acceleration|, root quat norm error */
    mr_float4 diagnostics;
} MRWorldStatusC;

/* Per-environment contact evidence of the last executed microstep of a
 * control step (subset of MRMetalWorldContactStatusGPU, engine_types.h:1215). */
typedef struct MRContactStatusC {
    uint32_t code;
    uint32_t active_pairs;
    uint32_t raw_contacts;
    uint32_t manifolds;
    uint32_t constraints;
    uint32_t islands;
    uint32_t retained_points;
    uint32_t new_points;
    uint32_t required_pairs;
    uint32_t required_manifolds;
    uint32_t required_constraints;
    uint32_t solver_iterations;
    uint32_t first_failing_pair;
    uint32_t first_failing_constraint;
    uint32_t max_active_pairs;      /* high water over the control step */
    uint32_t max_constraints;       /* high water over the control step */
    /* final-sweep max |delta impulse|, normal residual, cone violation,
     * cone sweeps run over the control step (all microsteps, fallbacks included) */
    mr_float4 residuals;
    /* max penetration, retained fraction, microsteps whose learned warm start
     * fell back to the plain one, and w: those fallbacks the residual guard made
     * or, in the hybrid CCD mode (which excludes the learned warm start), the
     * segments that ended at a time of impact over the control step */
    mr_float4 diagnostics;
} MRContactStatusC;

/* Teacher-forced evidence record: one active broadphase pair of the captured
 * microstep, in canonical (environment, collider A, collider B) order. */
typedef struct MRPairRecordC {
    uint32_t environment;
    uint32_t collider_a;
    uint32_t collider_b;
    uint32_t pair_class;
} MRPairRecordC;

/* --------------------------------------------------------------- config -- */

/* Mirrors the MetalWorldStepConfig fields of the temporalCone path
 * (include/metalrobo/MetalWorld.hpp:380-452). */
/* Iteration of the cone solve (MRWorldStepConfigC.solver). Both start from
 * the projected warm start and use the same per-contact 3x3 block inverse.
 *   ORDERED_SWEEPS: ordered block Gauss-Seidel, each contact update applied
 *     before the next (mr_world_solve_contact_islands,
 *     MetalWorldContact.metal:7856-8316);
 *   RELAXED_JACOBI: the Wave32 production iteration (mrWorldWave32SolvePacket,
 *     MetalWorldContact.metal:6142-7213): every contact proposes its update
 *     from the velocity at the start of the iteration, damped by
 *     diag / sum_j |W_nn(c, j)| clamped to [1/32, 1], and the velocity then
 *     takes all deltas at once; residuals.x is the largest |delta| over all
 *     iterations of the microstep. */
enum MRSolverMode { MR_SOLVER_ORDERED_SWEEPS = 0u, MR_SOLVER_RELAXED_JACOBI = 1u };

typedef struct MRWorldStepConfigC {
    float control_timestep;           /* seconds per control step */
    uint32_t physics_substeps;        /* microsteps per control step */
    uint32_t velocity_iterations;     /* cone sweeps per microstep */
    uint32_t final_velocity_iterations; /* extra sweeps on the final microstep */
    uint32_t actuation_mode;          /* MRActuationMode */
    uint32_t warm_start;
    uint32_t apply_body_damping;
    uint32_t capture_evidence;        /* record the pair list of capture_substep */
    uint32_t capture_substep;
    uint32_t pair_capacity;           /* per environment */
    uint32_t manifold_capacity;       /* per environment */
    uint32_t constraint_capacity;     /* per environment */
    float manifold_breaking_separation; /* 0.02 m */
    float manifold_breaking_tangential; /* 0.02 m */
    float manifold_merge_distance;      /* 0.002 m */
    float manifold_normal_cosine;       /* 0.95 */
    float contact_slop;               /* added to penetration before stabilization */
    float max_depenetration_velocity; /* clamp of the stabilization velocity */
    float residual_tolerance;         /* >0: stop sweeps once final |dlambda| <= tol */
    uint32_t max_iterations;          /* sweep cap when residual_tolerance > 0 */
    uint32_t solver;                  /* MRSolverMode */
    float maximum_normal_impulse;     /* > 0: cap of every contact's normal impulse (the cone's
                                         maximumNormalImpulse, MetalWorldContact.metal:488-493); 0: none */
    /* continuous collision (MetalWorld.hpp:420-426, MetalWorld.mm:3157-3215) */
    uint32_t ccd_mode;                /* MRCcdMode */
    uint32_t max_ccd_events;          /* impact segments per microstep, <= MR_CCD_MAX_EVENTS (default 4) */
    uint32_t max_ccd_iterations;      /* conservative-advancement iterations per pair (default 16) */
    uint32_t ccd_reserved;
    float ccd_distance_tolerance;     /* m: a pair this close counts as touching (default 1e-3) */
    float ccd_time_tolerance;         /* s: shortest impact segment (default 1e-5) */
    float ccd_minimum_advance;        /* s: smallest conservative-advancement step (default 1e-6) */
    float ccd_reserved1;
} MRWorldStepConfigC;

#ifdef __cplusplus
#define MR_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define MR_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
MR_STATIC_ASSERT(sizeof(MRBodyC) == 256, "MRBodyC size");
MR_STATIC_ASSERT(sizeof(MRDofC) == 48, "MRDofC size");
MR_STATIC_ASSERT(sizeof(MRActuatorProfileC) == 48, "MRActuatorProfileC size");
MR_STATIC_ASSERT(sizeof(MRShapeC) == 96, "MRShapeC size");
MR_STATIC_ASSERT(sizeof(MRMaterialC) == 32, "MRMaterialC size");
MR_STATIC_ASSERT(sizeof(MRGeometryC) == 64, "MRGeometryC size");
MR_STATIC_ASSERT(sizeof(MRTriangleC) == 16, "MRTriangleC size");
MR_STATIC_ASSERT(sizeof(MRBvhNodeC) == 128, "MRBvhNodeC size");
MR_STATIC_ASSERT(sizeof(MRCompiledPairC) == 16, "MRCompiledPairC size");
MR_STATIC_ASSERT(sizeof(MRModelHeaderC) == 64, "MRModelHeaderC size");
MR_STATIC_ASSERT(sizeof(MRBodyStateC) == 128, "MRBodyStateC size");
MR_STATIC_ASSERT(sizeof(MRManifoldHeaderC) == 64, "MRManifoldHeaderC size");
MR_STATIC_ASSERT(sizeof(MRManifoldPointC) == 64, "MRManifoldPointC size");
MR_STATIC_ASSERT(sizeof(MRWorldStatusC) == 48, "MRWorldStatusC size");
MR_STATIC_ASSERT(sizeof(MRContactStatusC) == 96, "MRContactStatusC size");
MR_STATIC_ASSERT(sizeof(MRPairRecordC) == 16, "MRPairRecordC size");
MR_STATIC_ASSERT(sizeof(MRWorldStepConfigC) == 120, "MRWorldStepConfigC size");

#endif /* MR_WORLD_TYPES_H */
