// TEST INFRASTRUCTURE — the CPU oracle of the batched temporalCone microstep.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load or run anything under oracle/; the product never does.
//
// An FP32 restatement of the reference's production contact graph
// (docs/METAL_WORLD.md:181-201), one environment at a time, written with a
// fixed evaluation order and compiled with -ffp-contract=off. Each stage cites
// the reference source it restates:
//   drive refresh          MetalWorld.metal:280-366
//   forward kinematics     ArticulatedABA.metal:742-900
//   ABA                    ArticulatedABA.metal:352-1440 (+ effectiveArmature :301-315)
//   scene prediction       MetalWorldContact.metal:2509-2567
//   collider projection    Collision.metal:321-600 (inflateFiniteBounds :110-120)
//   broadphase flags       Collision.metal:796-858, 7905-7962
//   narrowphase            Collision.metal:1003-1720, 4379-4830
//   heightfield            first-principles support-point / triangle-plane test
//                          (the reference's CPU path has no heightfield oracle)
//   manifold refresh       Collision.metal:5108-5530 (buildPersistentWorldManifold)
//   constraint emission    Collision.metal:9180-9470
//   IR evaluation          MetalWorldContact.metal:3800-3995
//   inverse-ABA responses  ArticulatedInverseMass.metal:1652-1700 (streamed responses)
//   block cone solve       MetalWorldContact.metal:482-720, 7856-8316
//   joint limits           MetalWorldContact.metal:2878-2960
//   integration            MetalWorldContact.metal:8876-9043
//   transactional commit   MetalWorld.metal:372-519, MetalWorldContact.metal:9051-9082
//
// Reference parity is pinned by tests/ (FP64 CRBA/RNEA dynamics, closed-form
// collision cases, the free-body and exact-cone known answers, brute-force
// pair lists); see the function-level notes below for what each check covers.

#include "oracle_math.h"

#include "../include/mr_world_types.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

using namespace oracle;

namespace {

constexpr float kTiny = 1.0e-14f;
constexpr float kFloatEps = 1.1920928955078125e-7f;
constexpr float kAbsolutePivotFloor = 1.0e-12f;
constexpr float kMatrixFloor = 1.0e-10f;
constexpr float kConeEpsilon = 1.0e-7f;
constexpr float kContactRegularization = 1.0e-2f;
constexpr uint32_t kMaxCandidates = 12u;

V3 f3(const mr_float4& v) { return V3{v.x, v.y, v.z}; }
Q4 q4(const mr_float4& v) { return Q4{v.x, v.y, v.z, v.w}; }

uint32_t feature_key(uint32_t shape_type, uint32_t local) {
    return ((shape_type & 0x0fu) << 28u) | (local & 0x0fffffffu);
}

// ----------------------------------------------------------- spatial algebra

// Articulated inertia [[A, B], [B^T, C]] in a world-aligned frame at a body's
// COM; A and C symmetric (xx, xy, xz, yy, yz, zz), B row-major.
struct SpI {
    float A[6];
    float B[9];
    float C[6];
};

V3 sym_mul(const float s[6], V3 w) {
    return V3{s[0] * w.x + s[1] * w.y + s[2] * w.z,
              s[1] * w.x + s[3] * w.y + s[4] * w.z,
              s[2] * w.x + s[4] * w.y + s[5] * w.z};
}
V3 mat_mul(const float b[9], V3 v) {
    return V3{b[0] * v.x + b[1] * v.y + b[2] * v.z,
              b[3] * v.x + b[4] * v.y + b[5] * v.z,
              b[6] * v.x + b[7] * v.y + b[8] * v.z};
}
V3 matT_mul(const float b[9], V3 w) {
    return V3{b[0] * w.x + b[3] * w.y + b[6] * w.z,
              b[1] * w.x + b[4] * w.y + b[7] * w.z,
              b[2] * w.x + b[5] * w.y + b[8] * w.z};
}
inline int sym_index(int i, int j) {
    static const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return idx[i][j];
}

// I' = X^T I X for a pure translation r (child COM minus parent COM):
// A' = A - (Q + Q^T) - r~ C r~,  B' = B + r~ C,  C' = C,  Q = B r~.
void shift_inertia(const SpI& in, V3 r, SpI& out) {
    float Cm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Cm[i][j] = in.C[sym_index(i, j)];
    float T[3][3];
    for (int j = 0; j < 3; ++j) {
        T[0][j] = r.y * Cm[2][j] - r.z * Cm[1][j];
        T[1][j] = r.z * Cm[0][j] - r.x * Cm[2][j];
        T[2][j] = r.x * Cm[1][j] - r.y * Cm[0][j];
    }
    float Q[3][3];
    for (int i = 0; i < 3; ++i) {
        const float b0 = in.B[3 * i + 0], b1 = in.B[3 * i + 1], b2 = in.B[3 * i + 2];
        Q[i][0] = b1 * r.z - b2 * r.y;
        Q[i][1] = b2 * r.x - b0 * r.z;
        Q[i][2] = b0 * r.y - b1 * r.x;
    }
    float P[3][3];
    for (int i = 0; i < 3; ++i) {
        P[i][0] = T[i][1] * r.z - T[i][2] * r.y;
        P[i][1] = T[i][2] * r.x - T[i][0] * r.z;
        P[i][2] = T[i][0] * r.y - T[i][1] * r.x;
    }
    static const int si[6] = {0, 0, 0, 1, 1, 2};
    static const int sj[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
        const int i = si[k], j = sj[k];
        out.A[k] = in.A[k] - (Q[i][j] + Q[j][i]) - P[i][j];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out.B[3 * i + j] = in.B[3 * i + j] + T[i][j];
    for (int k = 0; k < 6; ++k) out.C[k] = in.C[k];
}

float spi_max_abs(const SpI& I) {
    float m = 0.0f;
    for (int k = 0; k < 6; ++k) m = fmaxd(m, std::fabs(I.A[k]));
    for (int k = 0; k < 9; ++k) m = fmaxd(m, std::fabs(I.B[k]));
    for (int k = 0; k < 6; ++k) m = fmaxd(m, std::fabs(I.C[k]));
    return m;
}

float spi6(const SpI& I, int r, int c) {
    if (r < 3 && c < 3) return I.A[sym_index(r, c)];
    if (r < 3) return I.B[3 * r + (c - 3)];
    if (c < 3) return I.B[3 * c + (r - 3)];
    return I.C[sym_index(r - 3, c - 3)];
}

// ------------------------------------------------------------ world views

struct EnvModel {
    const MRModelHeaderC* header;
    const MRBodyC* bodies;
    const MRDofC* dofs;
    const MRActuatorProfileC* actuators;  // [nv] or null
    const MRShapeC* shapes;
    const MRMaterialC* materials;
    const MRCompiledPairC* pairs;
    const MRHeightfieldC* heightfield;
    const float* heights;
    const MRGeometryC* geometries;
    const mr_float4* geometry_vertices;
    const MRTriangleC* triangles;
    const MRBvhNodeC* bvh_nodes;
    V3 gravity;
};

EnvModel env_model(const MRWorldModelDescC* desc, uint32_t env) {
    EnvModel m;
    m.header = &desc->models[desc->environment_model[env]];
    m.bodies = desc->bodies + m.header->body_offset;
    m.dofs = desc->dofs + m.header->dof_offset;
    m.actuators = desc->actuators != nullptr ? desc->actuators + m.header->dof_offset : nullptr;
    m.shapes = desc->shapes + m.header->shape_offset;
    m.materials = desc->materials + m.header->material_offset;
    m.pairs = desc->pairs + m.header->pair_offset;
    m.heightfield = &desc->heightfield;
    m.heights = desc->heightfield_heights;
    m.geometries = desc->geometries;
    m.geometry_vertices = desc->geometry_vertices;
    m.triangles = desc->triangles;
    m.bvh_nodes = desc->bvh_nodes;
    m.gravity = V3{desc->gravity.x, desc->gravity.y, desc->gravity.z};
    return m;
}

// Projected collider (MRProjectedColliderGPU + the derived WorldShape fields).
struct Collider {
    uint32_t type;
    uint32_t body;
    uint32_t disabled;
    V3 center;
    Q4 rotation;
    V3 lower, upper;
    float radius;
    float half_length;
    float contact_offset;
    V3 half_extents;
    V3 p0, p1;        // capsule endpoints
    V3 plane_normal;  // plane
};

struct RawContact {
    V3 normal;
    float separation;
    V3 point_a, point_b;
    uint32_t feature_a, feature_b;
};

struct RawBatch {
    RawContact c[MR_RAW_CONTACTS_PER_PAIR];
    uint32_t count;
};

struct Candidate {
    MRManifoldPointC point;
    V3 world_point;
    float separation;
    float drift;
};

// Solver-side constraint (MRContactConstraintGPU + evaluated IR rows/cone).
struct Constraint {
    uint32_t body_a, body_b;    // model body indices
    uint32_t flags;
    uint32_t manifold, point;
    V3 point_world, pa, pb;     // midpoint, anchor world points
    float separation;
    V3 dir[3];                  // normal, tangent u, tangent v
    float friction[4];          // static u, static v, dynamic u, dynamic v
    float restitution, restitution_threshold, compliance, dissipation;
    float target[3];
    float reg[3];
    float mu_u, mu_v;
    float lambda[3];
    float winv[3][3];
};

struct Work {
    // articulation
    V3 pos[MR_MAX_LINKS];
    Q4 rot[MR_MAX_LINKS];
    V3 w[MR_MAX_LINKS], vl[MR_MAX_LINKS];
    V3 sa[MR_MAX_LINKS], sl[MR_MAX_LINKS];
    V3 ca[MR_MAX_LINKS], cl[MR_MAX_LINKS];
    V3 r[MR_MAX_LINKS];
    SpI I[MR_MAX_LINKS];
    V3 pt[MR_MAX_LINKS], pb[MR_MAX_LINKS];
    V3 Ut[MR_MAX_LINKS], Ub[MR_MAX_LINKS];
    float D[MR_MAX_LINKS], u[MR_MAX_LINKS];
    V3 alpha[MR_MAX_LINKS], acc[MR_MAX_LINKS];
    float L[36];
    float effort[MR_MAX_DOFS];
    float accel[MR_MAX_DOFS];
    float vcand[MR_MAX_DOFS];
    float minPivot, maxPivot;
    // scene (indexed by model body index; articulation links hold poses only)
    MRBodyStateC bodies[MR_MAX_BODIES];
    Collider col[MR_MAX_SHAPES];
};

struct StepContext {
    const MRWorldModelDescC* desc;
    const MRWorldStepConfigC* cfg;
    EnvModel m;
    uint32_t env;
    float h;
    uint32_t manifold_capacity;
    uint32_t pair_capacity;
    uint32_t constraint_capacity;
};

// Optional capture of one environment's final-microstep contact problem, for
// the FP64 exact-cone comparison (serial control-step entry point only).
struct ProblemCapture {
    uint32_t env = 0xffffffffu;
    float* out = nullptr;
    uint32_t capacity = 0;  // floats
    uint32_t used = 0;
};
thread_local ProblemCapture g_capture;

struct Failure {
    uint32_t code = MR_STEP_SUCCESS;
    uint32_t index = MR_INVALID_INDEX;
};

// ------------------------------------------------------------------ drives

// MetalWorld.metal:252-269 / 345-365: the actuator torque-speed envelope
//   min(limits.w, stall * efficiency * (1 - clamp(|v| / no_load, 0, 1))),
// with an inactive profile executing as no-load FLT_MAX, efficiency 1 and
// stall = the effort limit (FLT_MAX without one), MetalWorld.mm:1376-1410.
float actuator_envelope(const MRDofC& dof, const MRActuatorProfileC* act, float vel) {
    const float big = std::numeric_limits<float>::max();
    float no_load = big, efficiency = 1.0f;
    float stall = ((dof.flags & MR_DOF_FLAG_EFFORT_LIMIT) != 0u && dof.limits.w > 0.0f) ? dof.limits.w : big;
    if (act != nullptr && (act->identity[1] & MR_ACTUATOR_PROFILE_ACTIVE) != 0u) {
        no_load = act->motor_and_speed.z;
        efficiency = act->motor_and_speed.w;
        stall = act->transmission_and_envelope.z;
    }
    const float speed_fraction = clampd(std::fabs(vel) / fmaxd(no_load, std::numeric_limits<float>::min()), 0.0f, 1.0f);
    return fmind(dof.limits.w, stall * efficiency * (1.0f - speed_fraction));
}

// MetalWorld.metal:184-270 (prepare: effort commands, clamped at the control
// step's start velocity and held for its microsteps) and 280-366 (implicit
// position drive refreshed from the current q/v every microstep); every
// command is clamped to the actuator envelope.
void refresh_drives(const StepContext& ctx, const float* q, const float* v,
                    const float* command, Work& wk) {
    const uint32_t nv = ctx.m.header->nv;
    for (uint32_t j = 0; j < nv; ++j) {
        const MRDofC& dof = ctx.m.dofs[j];
        const float envelope = actuator_envelope(dof, ctx.m.actuators != nullptr ? &ctx.m.actuators[j] : nullptr, v[j]);
        if (ctx.cfg->actuation_mode == MR_ACTUATION_EFFORT) {
            wk.effort[j] = clampd(command[j], -envelope, envelope);
            continue;
        }
        float cmd = 0.0f;
        if ((dof.flags & MR_DOF_FLAG_DRIVE) != 0u && dof.q_index != MR_INVALID_INDEX) {
            float target = command[j];
            if ((dof.flags & MR_DOF_FLAG_POSITION_LIMIT) != 0u)
                target = clampd(target, dof.limits.x, dof.limits.y);
            const float vel = v[j];
            cmd = dof.drive.x * (target - q[dof.q_index] - ctx.h * vel) - dof.drive.y * vel;
            const float dry = dof.drive.w;
            if (dry > 0.0f) {
                if (std::fabs(vel) > 1.0e-4f)
                    cmd -= vel > 0.0f ? dry : -dry;
                else
                    cmd -= clampd(cmd, -dry, dry);
            }
            if ((dof.flags & MR_DOF_FLAG_EFFORT_LIMIT) != 0u && dof.limits.w > 0.0f)
                cmd = clampd(cmd, -dof.limits.w, dof.limits.w);
        }
        wk.effort[j] = clampd(cmd, -envelope, envelope);
    }
}

// ArticulatedABA.metal:301-315
float effective_armature(const StepContext& ctx, uint32_t j) {
    const MRDofC& dof = ctx.m.dofs[j];
    float value = dof.drive.z;
    if (ctx.cfg->actuation_mode == MR_ACTUATION_IMPLICIT_POSITION_DRIVE &&
        (dof.flags & MR_DOF_FLAG_DRIVE) != 0u) {
        value += ctx.h * dof.drive.y + ctx.h * ctx.h * dof.drive.x;
    }
    return value;
}

// --------------------------------------------------------- kinematics + ABA

// Forward kinematics, motion subspaces and velocity-product accelerations of
// every link (ArticulatedABA.metal:680-900). Returns false on a bad root
// quaternion / non-finite pose.
bool forward_kinematics(const StepContext& ctx, const float* q, const float* v, Work& wk,
                        Failure& f) {
    const MRModelHeaderC& h = *ctx.m.header;
    if (h.root_type == MR_ROOT_FLOATING) {
        Q4 rq;
        if (!checked_quat(Q4{q[3], q[4], q[5], q[6]}, rq)) {
            f.code = MR_STEP_NONFINITE_INPUT;
            f.index = 3u;
            return false;
        }
        wk.pos[0] = V3{q[0], q[1], q[2]};
        wk.rot[0] = rq;
        wk.vl[0] = V3{v[0], v[1], v[2]};
        wk.w[0] = V3{v[3], v[4], v[5]};
    } else {
        wk.pos[0] = V3{0.0f, 0.0f, 0.0f};
        wk.rot[0] = Q4{0.0f, 0.0f, 0.0f, 1.0f};
        wk.vl[0] = V3{0.0f, 0.0f, 0.0f};
        wk.w[0] = V3{0.0f, 0.0f, 0.0f};
    }
    const V3 zero{0.0f, 0.0f, 0.0f};
    wk.r[0] = zero;
    wk.sa[0] = zero;
    wk.sl[0] = zero;
    wk.ca[0] = zero;
    wk.cl[0] = zero;
    for (uint32_t i = 1; i < h.link_count; ++i) {
        const MRBodyC& b = ctx.m.bodies[i];
        const uint32_t p = b.parent;
        const Q4 pjr = qmul(wk.rot[p], q4(b.parent_rotation));
        const bool moving = b.joint_type == MR_JOINT_REVOLUTE || b.joint_type == MR_JOINT_CONTINUOUS;
        const V3 axis_n = normalize(f3(b.axis));
        float qi = 0.0f, qd = 0.0f;
        Q4 motion{0.0f, 0.0f, 0.0f, 1.0f};
        if (moving) {
            qi = q[b.q_index];
            qd = v[b.v_index];
            motion = axis_angle(axis_n, qi);
        }
        Q4 child;
        if (!normalized_quat(qmul(qmul(pjr, motion), qconj(q4(b.child_rotation))), child)) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = i;
            return false;
        }
        wk.rot[i] = child;
        const V3 axis_w = qrot(pjr, axis_n);
        const V3 joint_pos = add(wk.pos[p], qrot(wk.rot[p], f3(b.parent_anchor)));
        const V3 child_anchor = qrot(child, f3(b.child_anchor));
        wk.pos[i] = sub(joint_pos, child_anchor);
        wk.r[i] = sub(wk.pos[i], wk.pos[p]);
        const V3 p2j = sub(joint_pos, wk.pos[p]);
        const V3 wp = wk.w[p];
        wk.w[i] = moving ? add(wp, mul(qd, axis_w)) : wp;
        wk.vl[i] = sub(add(wk.vl[p], cross(wp, p2j)), cross(wk.w[i], child_anchor));
        wk.sa[i] = moving ? axis_w : zero;
        wk.sl[i] = moving ? neg(cross(axis_w, child_anchor)) : zero;
        wk.ca[i] = moving ? mul(qd, cross(wp, axis_w)) : zero;
        wk.cl[i] = sub(sub(cross(wp, cross(wp, p2j)), cross(wk.ca[i], child_anchor)),
                       cross(wk.w[i], cross(wk.w[i], child_anchor)));
        if (!finite3(wk.pos[i]) || !finite4(wk.rot[i]) || !finite3(wk.w[i]) ||
            !finite3(wk.vl[i]) || !finite3(wk.cl[i]) || !finite3(wk.ca[i])) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = i;
            return false;
        }
    }
    return true;
}

// World-frame inertia R * I_body * R^T as (xx, xy, xz, yy, yz, zz).
void world_inertia(const MRBodyC& b, Q4 rot, float out[6]) {
    const V3 c0 = qrot(rot, V3{1.0f, 0.0f, 0.0f});
    const V3 c1 = qrot(rot, V3{0.0f, 1.0f, 0.0f});
    const V3 c2 = qrot(rot, V3{0.0f, 0.0f, 1.0f});
    const float R[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
    const float I[3][3] = {{b.inertia_row0.x, b.inertia_row0.y, b.inertia_row0.z},
                           {b.inertia_row1.x, b.inertia_row1.y, b.inertia_row1.z},
                           {b.inertia_row2.x, b.inertia_row2.y, b.inertia_row2.z}};
    static const int si[6] = {0, 0, 0, 1, 1, 2};
    static const int sj[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
        const int r = si[k], c = sj[k];
        float acc = 0.0f;
        for (int a = 0; a < 3; ++a)
            for (int b2 = 0; b2 < 3; ++b2) acc += R[r][a] * I[a][b2] * R[c][b2];
        out[k] = acc;
    }
}

// Articulated-body algorithm: accelerations for the current state, effort and
// gravity (ArticulatedABA.metal:900-1440). Stores U, D, u, the articulated
// inertias and the root factor for the inverse-ABA response pass.
bool aba(const StepContext& ctx, const float* v, Work& wk, Failure& f) {
    const MRModelHeaderC& h = *ctx.m.header;
    const uint32_t n = h.link_count;
    const bool damping = ctx.cfg->apply_body_damping != 0u;
    for (uint32_t i = 0; i < n; ++i) {
        const MRBodyC& b = ctx.m.bodies[i];
        SpI& I = wk.I[i];
        world_inertia(b, wk.rot[i], I.A);
        for (int k = 0; k < 9; ++k) I.B[k] = 0.0f;
        const float m = b.mass.x;
        I.C[0] = m; I.C[1] = 0.0f; I.C[2] = 0.0f;
        I.C[3] = m; I.C[4] = 0.0f; I.C[5] = m;
        const V3 wi = wk.w[i];
        V3 bt = cross(wi, sym_mul(I.A, wi));
        V3 bf = neg(mul(m, ctx.m.gravity));
        if (damping) {
            bt = add(bt, mul(b.damping.y, wi));
            bf = add(bf, mul(b.damping.x, wk.vl[i]));
        }
        wk.pt[i] = bt;
        wk.pb[i] = bf;
    }
    if (h.root_type == MR_ROOT_FLOATING) {
        SpI& I = wk.I[0];
        I.C[0] += effective_armature(ctx, 0);
        I.C[3] += effective_armature(ctx, 1);
        I.C[5] += effective_armature(ctx, 2);
        I.A[0] += effective_armature(ctx, 3);
        I.A[3] += effective_armature(ctx, 4);
        I.A[5] += effective_armature(ctx, 5);
        wk.pt[0] = sub(wk.pt[0], V3{wk.effort[3], wk.effort[4], wk.effort[5]});
        wk.pb[0] = sub(wk.pb[0], V3{wk.effort[0], wk.effort[1], wk.effort[2]});
    }
    wk.minPivot = INFINITY;
    wk.maxPivot = 0.0f;
    for (uint32_t i = n - 1; i >= 1; --i) {
        const MRBodyC& b = ctx.m.bodies[i];
        SpI& I = wk.I[i];
        const V3 Ica_t = add(sym_mul(I.A, wk.ca[i]), mat_mul(I.B, wk.cl[i]));
        const V3 Ica_b = add(matT_mul(I.B, wk.ca[i]), sym_mul(I.C, wk.cl[i]));
        V3 pt = add(wk.pt[i], Ica_t);
        V3 pb = add(wk.pb[i], Ica_b);
        SpI Ip = I;
        if (b.v_index != MR_INVALID_INDEX) {
            const V3 sa = wk.sa[i], sl = wk.sl[i];
            const V3 Ut = add(sym_mul(I.A, sa), mat_mul(I.B, sl));
            const V3 Ub = add(matT_mul(I.B, sa), sym_mul(I.C, sl));
            const float D = (dot(sa, Ut) + dot(sl, Ub)) + effective_armature(ctx, b.v_index);
            const float floor = fmaxd(kAbsolutePivotFloor, spi_max_abs(I) * 6.0f * kFloatEps);
            if (!(D > floor) || !std::isfinite(D)) {
                f.code = MR_STEP_FACTORIZATION_FAILED;
                f.index = b.v_index;
                return false;
            }
            const float u = (wk.effort[b.v_index] - dot(sa, pt)) - dot(sl, pb);
            const float uD = u / D;
            pt = add(pt, mul(uD, Ut));
            pb = add(pb, mul(uD, Ub));
            const float ut[3] = {Ut.x, Ut.y, Ut.z};
            const float ub[3] = {Ub.x, Ub.y, Ub.z};
            static const int si[6] = {0, 0, 0, 1, 1, 2};
            static const int sj[6] = {0, 1, 2, 1, 2, 2};
            for (int k = 0; k < 6; ++k) Ip.A[k] = I.A[k] - (ut[si[k]] * ut[sj[k]]) / D;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) Ip.B[3 * a + c] = I.B[3 * a + c] - (ut[a] * ub[c]) / D;
            for (int k = 0; k < 6; ++k) Ip.C[k] = I.C[k] - (ub[si[k]] * ub[sj[k]]) / D;
            wk.Ut[i] = Ut;
            wk.Ub[i] = Ub;
            wk.D[i] = D;
            wk.u[i] = u;
            const float pivot = std::sqrt(D);
            wk.minPivot = fmind(wk.minPivot, pivot);
            wk.maxPivot = fmaxd(wk.maxPivot, pivot);
        }
        SpI shifted;
        shift_inertia(Ip, wk.r[i], shifted);
        SpI& P = wk.I[b.parent];
        for (int k = 0; k < 6; ++k) P.A[k] += shifted.A[k];
        for (int k = 0; k < 9; ++k) P.B[k] += shifted.B[k];
        for (int k = 0; k < 6; ++k) P.C[k] += shifted.C[k];
        wk.pt[b.parent] = add(wk.pt[b.parent], add(pt, cross(wk.r[i], pb)));
        wk.pb[b.parent] = add(wk.pb[b.parent], pb);
    }
    if (h.root_type == MR_ROOT_FLOATING) {
        const SpI& I = wk.I[0];
        const float floor = fmaxd(kAbsolutePivotFloor, spi_max_abs(I) * 6.0f * kFloatEps);
        for (int k = 0; k < 36; ++k) wk.L[k] = 0.0f;
        for (int row = 0; row < 6; ++row) {
            for (int col = 0; col <= row; ++col) {
                float val = spi6(I, row, col);
                for (int k = 0; k < col; ++k) val -= wk.L[row * 6 + k] * wk.L[col * 6 + k];
                if (row == col) {
                    if (!(val > floor) || !std::isfinite(val)) {
                        f.code = MR_STEP_FACTORIZATION_FAILED;
                        f.index = (uint32_t)row;
                        return false;
                    }
                    const float piv = std::sqrt(val);
                    wk.L[row * 6 + row] = piv;
                    wk.minPivot = fmind(wk.minPivot, piv);
                    wk.maxPivot = fmaxd(wk.maxPivot, piv);
                } else {
                    wk.L[row * 6 + col] = val / wk.L[col * 6 + col];
                }
            }
        }
        const float rhs[6] = {-wk.pt[0].x, -wk.pt[0].y, -wk.pt[0].z,
                              -wk.pb[0].x, -wk.pb[0].y, -wk.pb[0].z};
        float y[6], x[6];
        for (int row = 0; row < 6; ++row) {
            float val = rhs[row];
            for (int col = 0; col < row; ++col) val -= wk.L[row * 6 + col] * y[col];
            y[row] = val / wk.L[row * 6 + row];
        }
        for (int row = 5; row >= 0; --row) {
            float val = y[row];
            for (int col = row + 1; col < 6; ++col) val -= wk.L[col * 6 + row] * x[col];
            x[row] = val / wk.L[row * 6 + row];
        }
        wk.alpha[0] = V3{x[0], x[1], x[2]};
        wk.acc[0] = V3{x[3], x[4], x[5]};
        wk.accel[0] = x[3]; wk.accel[1] = x[4]; wk.accel[2] = x[5];
        wk.accel[3] = x[0]; wk.accel[4] = x[1]; wk.accel[5] = x[2];
    } else {
        wk.alpha[0] = V3{0.0f, 0.0f, 0.0f};
        wk.acc[0] = V3{0.0f, 0.0f, 0.0f};
    }
    for (uint32_t i = 1; i < n; ++i) {
        const MRBodyC& b = ctx.m.bodies[i];
        const V3 ap = wk.alpha[b.parent];
        const V3 lp = add(wk.acc[b.parent], cross(ap, wk.r[i]));
        V3 ai = add(ap, wk.ca[i]);
        V3 li = add(lp, wk.cl[i]);
        if (b.v_index != MR_INVALID_INDEX) {
            const float proj = dot(wk.Ut[i], ap) + dot(wk.Ub[i], lp);
            const float qdd = (wk.u[i] - proj) / wk.D[i];
            wk.accel[b.v_index] = qdd;
            ai = add(ai, mul(qdd, wk.sa[i]));
            li = add(li, mul(qdd, wk.sl[i]));
        }
        wk.alpha[i] = ai;
        wk.acc[i] = li;
    }
    const uint32_t nv = h.nv;
    for (uint32_t j = 0; j < nv; ++j) {
        const float a = wk.accel[j];
        const float nvv = v[j] + ctx.h * a;
        if (!std::isfinite(a) || !std::isfinite(nvv)) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = j;
            return false;
        }
        wk.vcand[j] = nvv;
    }
    return true;
}

// Inverse ABA: out = M_eff^{-1} J^T f for point forces on links, sharing the
// articulated inertias, U, D and root factor of the current microstep
// (ArticulatedInverseMass.metal:1652-1700, streamed response columns).
struct PointForce {
    uint32_t link;
    V3 point;
    V3 force;
};

void inverse_aba(const StepContext& ctx, const Work& wk, const PointForce* forces,
                 uint32_t force_count, float* out) {
    const MRModelHeaderC& h = *ctx.m.header;
    const uint32_t n = h.link_count;
    V3 pt[MR_MAX_LINKS], pb[MR_MAX_LINKS], al[MR_MAX_LINKS], li[MR_MAX_LINKS];
    float uu[MR_MAX_LINKS];
    for (uint32_t i = 0; i < n; ++i) {
        pt[i] = V3{0.0f, 0.0f, 0.0f};
        pb[i] = V3{0.0f, 0.0f, 0.0f};
    }
    for (uint32_t k = 0; k < force_count; ++k) {
        const PointForce& pf = forces[k];
        pt[pf.link] = sub(pt[pf.link], cross(sub(pf.point, wk.pos[pf.link]), pf.force));
        pb[pf.link] = sub(pb[pf.link], pf.force);
    }
    for (uint32_t i = n - 1; i >= 1; --i) {
        const MRBodyC& b = ctx.m.bodies[i];
        V3 t = pt[i], fb = pb[i];
        if (b.v_index != MR_INVALID_INDEX) {
            const float u = (0.0f - dot(wk.sa[i], t)) - dot(wk.sl[i], fb);
            const float uD = u / wk.D[i];
            t = add(t, mul(uD, wk.Ut[i]));
            fb = add(fb, mul(uD, wk.Ub[i]));
            uu[i] = u;
        }
        pt[b.parent] = add(pt[b.parent], add(t, cross(wk.r[i], fb)));
        pb[b.parent] = add(pb[b.parent], fb);
    }
    if (h.root_type == MR_ROOT_FLOATING) {
        const float rhs[6] = {-pt[0].x, -pt[0].y, -pt[0].z, -pb[0].x, -pb[0].y, -pb[0].z};
        float y[6], x[6];
        for (int row = 0; row < 6; ++row) {
            float val = rhs[row];
            for (int col = 0; col < row; ++col) val -= wk.L[row * 6 + col] * y[col];
            y[row] = val / wk.L[row * 6 + row];
        }
        for (int row = 5; row >= 0; --row) {
            float val = y[row];
            for (int col = row + 1; col < 6; ++col) val -= wk.L[col * 6 + row] * x[col];
            x[row] = val / wk.L[row * 6 + row];
        }
        al[0] = V3{x[0], x[1], x[2]};
        li[0] = V3{x[3], x[4], x[5]};
        out[0] = x[3]; out[1] = x[4]; out[2] = x[5];
        out[3] = x[0]; out[4] = x[1]; out[5] = x[2];
    } else {
        al[0] = V3{0.0f, 0.0f, 0.0f};
        li[0] = V3{0.0f, 0.0f, 0.0f};
    }
    for (uint32_t i = 1; i < n; ++i) {
        const MRBodyC& b = ctx.m.bodies[i];
        const V3 ap = al[b.parent];
        const V3 lp = add(li[b.parent], cross(ap, wk.r[i]));
        if (b.v_index != MR_INVALID_INDEX) {
            const float proj = dot(wk.Ut[i], ap) + dot(wk.Ub[i], lp);
            const float qdd = (uu[i] - proj) / wk.D[i];
            out[b.v_index] = qdd;
            al[i] = add(ap, mul(qdd, wk.sa[i]));
            li[i] = add(lp, mul(qdd, wk.sl[i]));
        } else {
            al[i] = ap;
            li[i] = lp;
        }
    }
}

// Velocity of a world point x carried by link `link` per unit generalized
// velocity j (one column of the point Jacobian), or zero.
V3 point_jacobian_column(const StepContext& ctx, const Work& wk, uint32_t link, V3 x,
                         uint32_t j) {
    const MRModelHeaderC& h = *ctx.m.header;
    const V3 zero{0.0f, 0.0f, 0.0f};
    if (h.root_type == MR_ROOT_FLOATING && j < 6u) {
        if (j < 3u) {
            V3 e = zero;
            setcomp(e, (int)j, 1.0f);
            return e;
        }
        V3 e = zero;
        setcomp(e, (int)j - 3, 1.0f);
        return cross(e, sub(x, wk.pos[0]));
    }
    const uint32_t owner = ctx.m.dofs[j].body;
    for (uint32_t b = link; b != 0u && b != MR_INVALID_INDEX; b = ctx.m.bodies[b].parent) {
        if (b == owner) return add(wk.sl[b], cross(wk.sa[b], sub(x, wk.pos[b])));
    }
    return zero;
}

// ------------------------------------------------------- collider projection

void inflate_bounds(Collider& c) {
    const V3 scale = add(vmax(vabs(c.lower), vabs(c.upper)), V3{1.0f, 1.0f, 1.0f});
    const V3 pad = mul(MR_COLLISION_AABB_RELATIVE_PAD, scale);
    c.lower = sub(c.lower, pad);
    c.upper = add(c.upper, pad);
}

// Collision.metal:321-600 (makeWorldShape) for the primitive types, plus the
// static axis-aligned heightfield.
// Collider of shape s with its body at (bpos, borient).
bool project_collider_pose(const StepContext& ctx, uint32_t s, V3 bpos, Q4 borient, Collider& c) {
    const MRShapeC& sh = ctx.m.shapes[s];
    Q4 br, lr;
    if (!checked_quat(borient, br) || !checked_quat(q4(sh.local_rotation), lr))
        return false;
    c.type = sh.shape_type;
    c.body = sh.body;
    c.disabled = (sh.flags & MR_SHAPE_FLAG_SIMULATION_DISABLED) != 0u ? 1u : 0u;
    c.rotation = qmul(br, lr);
    c.center = add(bpos, qrot(br, f3(sh.local_position)));
    c.contact_offset = sh.contact_rest_radius.x;
    c.radius = 0.0f;
    c.half_length = 0.0f;
    c.half_extents = V3{0.0f, 0.0f, 0.0f};
    c.p0 = c.p1 = c.plane_normal = V3{0.0f, 0.0f, 0.0f};
    c.lower = c.upper = V3{0.0f, 0.0f, 0.0f};
    if (!finite3(c.center) || !finite4(c.rotation)) return false;
    if (c.disabled) return true;
    const float co = c.contact_offset;
    switch (sh.shape_type) {
    case MR_SHAPE_SPHERE: {
        c.radius = sh.dimensions.x;
        const float e = c.radius + co;
        c.lower = sub(c.center, V3{e, e, e});
        c.upper = add(c.center, V3{e, e, e});
        inflate_bounds(c);
        break;
    }
    case MR_SHAPE_CAPSULE: {
        c.radius = sh.dimensions.x;
        c.half_length = sh.dimensions.y;
        const V3 axis = qrot(c.rotation, V3{0.0f, c.half_length, 0.0f});
        c.p0 = sub(c.center, axis);
        c.p1 = add(c.center, axis);
        const float e = c.radius + co;
        c.lower = sub(vmin(c.p0, c.p1), V3{e, e, e});
        c.upper = add(vmax(c.p0, c.p1), V3{e, e, e});
        inflate_bounds(c);
        break;
    }
    case MR_SHAPE_BOX: {
        c.half_extents = f3(sh.dimensions);
        const V3 bx = qrot(c.rotation, V3{1.0f, 0.0f, 0.0f});
        const V3 by = qrot(c.rotation, V3{0.0f, 1.0f, 0.0f});
        const V3 bz = qrot(c.rotation, V3{0.0f, 0.0f, 1.0f});
        const V3 ext = add(add(add(mul(c.half_extents.x, vabs(bx)), mul(c.half_extents.y, vabs(by))),
                               mul(c.half_extents.z, vabs(bz))),
                           V3{co, co, co});
        c.lower = sub(c.center, ext);
        c.upper = add(c.center, ext);
        inflate_bounds(c);
        break;
    }
    case MR_SHAPE_PLANE: {
        c.plane_normal = normalize(qrot(c.rotation, V3{0.0f, 1.0f, 0.0f}));
        break;
    }
    case MR_SHAPE_HEIGHTFIELD: {
        const MRHeightfieldC& hf = *ctx.m.heightfield;
        c.lower = sub(f3(hf.bounds_lower), V3{co, co, co});
        c.upper = add(f3(hf.bounds_upper), V3{co, co, co});
        inflate_bounds(c);
        break;
    }
    case MR_SHAPE_CONVEX:
    case MR_SHAPE_TRIANGLE_MESH: {
        // half_extents carries the vertex scale; bounds of the transformed
        // vertices (convex) or of the transformed local bounds (mesh)
        c.half_extents = f3(sh.dimensions);
        const MRGeometryC& g = ctx.m.geometries[sh.geometry];
        const V3 sc = c.half_extents;
        V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
        const uint32_t count = sh.shape_type == MR_SHAPE_CONVEX ? g.vertex_count : 8u;
        for (uint32_t i = 0; i < count; ++i) {
            V3 l;
            if (sh.shape_type == MR_SHAPE_CONVEX) {
                const mr_float4& v = ctx.m.geometry_vertices[g.vertex_offset + i];
                l = V3{v.x, v.y, v.z};
            } else {
                l = V3{(i & 1u) ? g.local_upper.x : g.local_lower.x, (i & 2u) ? g.local_upper.y : g.local_lower.y,
                       (i & 4u) ? g.local_upper.z : g.local_lower.z};
            }
            const V3 w = add(c.center, qrot(c.rotation, V3{sc.x * l.x, sc.y * l.y, sc.z * l.z}));
            lo = vmin(lo, w);
            hi = vmax(hi, w);
        }
        c.lower = sub(lo, V3{co, co, co});
        c.upper = add(hi, V3{co, co, co});
        inflate_bounds(c);
        break;
    }
    default:
        return false;
    }
    return finite3(c.lower) && finite3(c.upper);
}

bool project_collider(const StepContext& ctx, const Work& wk, uint32_t s, Collider& c) {
    const MRBodyStateC& body = wk.bodies[ctx.m.shapes[s].body];
    return project_collider_pose(ctx, s, f3(body.position), q4(body.orientation), c);
}

float collider_scale(const Collider& c) {
    return fmaxd(fmaxd(maxabs3(c.center), maxabs3(c.half_extents)),
                 fmaxd(fmaxd(maxabs3(c.p0), maxabs3(c.p1)),
                       fmaxd(fmaxd(std::fabs(c.radius), std::fabs(c.half_length)),
                             fmaxd(std::fabs(c.contact_offset),
                                   fmaxd(maxabs3(c.lower), maxabs3(c.upper))))));
}

float pair_query_padding(const Collider& a, const Collider& b) {
    return (fmaxd(collider_scale(a), collider_scale(b)) + 1.0f) * MR_COLLISION_QUERY_RELATIVE_PAD;
}

// Collision.metal:796-858 (projectedPairMayOverlap).
bool pair_may_overlap(const Collider& a, const Collider& b) {
    if (a.disabled || b.disabled) return false;
    if (a.type != MR_SHAPE_PLANE && b.type != MR_SHAPE_PLANE) {
        return a.lower.x <= b.upper.x && a.lower.y <= b.upper.y && a.lower.z <= b.upper.z &&
            a.upper.x >= b.lower.x && a.upper.y >= b.lower.y && a.upper.z >= b.lower.z;
    }
    const Collider& plane = a.type == MR_SHAPE_PLANE ? a : b;
    const Collider& fin = a.type == MR_SHAPE_PLANE ? b : a;
    const V3 center = mul(0.5f, add(fin.lower, fin.upper));
    const V3 half = mul(0.5f, sub(fin.upper, fin.lower));
    const float scale = fmaxd(fmaxd(maxabs3(plane.center), maxabs3(fin.lower)),
                              fmaxd(maxabs3(fin.upper),
                                    fmaxd(std::fabs(plane.contact_offset),
                                          std::fabs(fin.contact_offset))));
    const float pad = (scale + 1.0f) * MR_COLLISION_QUERY_RELATIVE_PAD;
    const float dist = dot(plane.plane_normal, sub(center, plane.center)) -
        dot(vabs(plane.plane_normal), half);
    return dist <= plane.contact_offset + pad;
}

// ------------------------------------------------------------- narrowphase

RawContact make_contact(V3 n, float sep, V3 pa, V3 pb, uint32_t fa, uint32_t fb) {
    RawContact c;
    c.normal = n;
    c.separation = sep;
    c.point_a = pa;
    c.point_b = pb;
    c.feature_a = fa;
    c.feature_b = fb;
    return c;
}

RawContact swapped(const RawContact& in) {
    RawContact c = in;
    c.normal = neg(in.normal);
    c.point_a = in.point_b;
    c.point_b = in.point_a;
    c.feature_a = in.feature_b;
    c.feature_b = in.feature_a;
    return c;
}

V3 coincident_normal(uint32_t a, uint32_t b) {
    const uint32_t hash = (a * 73856093u) ^ (b * 19349663u);
    V3 r{0.0f, 0.0f, 0.0f};
    setcomp(r, (int)(hash % 3u), (hash & 4u) == 0u ? 1.0f : -1.0f);
    return r;
}

V3 stable_segment_normal(V3 da, V3 db, uint32_t ca, uint32_t cb) {
    const V3 cr = cross(da, db);
    if (dot(cr, cr) > kTiny) return normalize(cr);
    V3 d;
    if (dot(da, da) > kTiny) d = normalize(da);
    else if (dot(db, db) > kTiny) d = normalize(db);
    else return coincident_normal(ca, cb);
    const V3 ad = vabs(d);
    V3 ref;
    if (ad.x <= ad.y && ad.x <= ad.z) ref = V3{1.0f, 0.0f, 0.0f};
    else if (ad.y <= ad.z) ref = V3{0.0f, 1.0f, 0.0f};
    else ref = V3{0.0f, 0.0f, 1.0f};
    V3 res = normalize(cross(d, ref));
    if ((((ca * 73856093u) ^ (cb * 19349663u)) & 4u) != 0u) res = neg(res);
    return res;
}

struct SegPoint {
    V3 point;
    float t;
};

SegPoint closest_on_segment(V3 p, V3 e0, V3 e1) {
    const V3 seg = sub(e1, e0);
    const float l2 = dot(seg, seg);
    if (!(l2 > kTiny)) return SegPoint{e0, 0.0f};
    const float t = clampd(dot(sub(p, e0), seg) / l2, 0.0f, 1.0f);
    return SegPoint{add(e0, mul(t, seg)), t};
}

struct SegPair {
    V3 pa, pb;
    float ta, tb;
};

// Collision.metal:1036-1125 (closestPointsOnSegments).
SegPair closest_segments(V3 a0, V3 a1, V3 b0, V3 b1) {
    const V3 da = sub(a1, a0), db = sub(b1, b0), off = sub(a0, b0);
    const float sa = dot(da, da), sb = dot(db, db);
    const float pa = dot(da, off), pb = dot(db, off);
    float ta = 0.0f, tb = 0.0f;
    if (!(sa > kTiny) && !(sb > kTiny)) return SegPair{a0, b0, 0.0f, 0.0f};
    if (!(sa > kTiny)) {
        tb = clampd(pb / sb, 0.0f, 1.0f);
    } else {
        const float cp = dot(da, db);
        if (!(sb > kTiny)) {
            ta = clampd(-pa / sa, 0.0f, 1.0f);
        } else {
            const float den = sa * sb - cp * cp;
            if (den > kTiny * sa * sb) ta = clampd((cp * pb - pa * sb) / den, 0.0f, 1.0f);
            tb = (cp * ta + pb) / sb;
            if (tb < 0.0f) {
                tb = 0.0f;
                ta = clampd(-pa / sa, 0.0f, 1.0f);
            } else if (tb > 1.0f) {
                tb = 1.0f;
                ta = clampd((cp - pa) / sa, 0.0f, 1.0f);
            }
        }
    }
    return SegPair{add(a0, mul(ta, da)), add(b0, mul(tb, db)), ta, tb};
}

uint32_t capsule_feature(float t) {
    if (t <= 0.0f) return 0u;
    return t >= 1.0f ? 1u : 2u;
}

uint32_t box_region(float c, float h) {
    if (c < -h) return 0u;
    return c > h ? 2u : 1u;
}

// Collision.metal:1152-1239 (sphereBoxWitness).
void sphere_box(const Collider& sphere, const Collider& box, V3& normal, V3& box_point,
                float& sep, uint32_t& feature) {
    const V3 lc = qrot_inv(box.rotation, sub(sphere.center, box.center));
    const V3 he = box.half_extents;
    V3 closest{clampd(lc.x, -he.x, he.x), clampd(lc.y, -he.y, he.y), clampd(lc.z, -he.z, he.z)};
    const V3 ld = sub(closest, lc);
    const float dist = length(ld);
    if (dist > kTiny) {
        normal = qrot(box.rotation, mul(1.0f / dist, ld));
        sep = dist - sphere.radius;
        feature = box_region(lc.x, he.x) + 3u * box_region(lc.y, he.y) + 9u * box_region(lc.z, he.z);
    } else {
        const V3 d = sub(he, vabs(lc));
        int axis = 0;
        if (d.y < comp(d, axis)) axis = 1;
        if (d.z < comp(d, axis)) axis = 2;
        const float sign = comp(lc, axis) >= 0.0f ? 1.0f : -1.0f;
        V3 outward{0.0f, 0.0f, 0.0f};
        setcomp(outward, axis, sign);
        setcomp(closest, axis, sign * comp(he, axis));
        normal = qrot(box.rotation, neg(outward));
        sep = -comp(d, axis) - sphere.radius;
        feature = 27u + 2u * (uint32_t)axis + (sign > 0.0f ? 1u : 0u);
    }
    box_point = add(box.center, qrot(box.rotation, closest));
}

void consider_capsule_box(V3 seg_point, V3 box_point, float t, uint32_t feat, float& best_sq,
                          V3& best_seg, V3& best_box, float& best_t, uint32_t& best_feat) {
    const V3 d = sub(box_point, seg_point);
    const float sq = dot(d, d);
    const float tol = 1.0e-12f * (1.0f + fmaxd(best_sq, sq));
    if (!std::isfinite(best_sq) || sq < best_sq - tol ||
        (std::fabs(sq - best_sq) <= tol && feat < best_feat)) {
        best_sq = sq;
        best_seg = seg_point;
        best_box = box_point;
        best_t = t;
        best_feat = feat;
    }
}

// Collision.metal:1241-1405 (capsuleBoxWitness).
void capsule_box(const Collider& cap, const Collider& box, uint32_t cap_index,
                 uint32_t box_index, V3& normal, V3& cap_point, V3& box_point, float& sep,
                 float& t_out, uint32_t& feature) {
    const V3 s0 = qrot_inv(box.rotation, sub(cap.p0, box.center));
    const V3 s1 = qrot_inv(box.rotation, sub(cap.p1, box.center));
    const V3 dir = sub(s1, s0);
    const V3 he = box.half_extents;
    float enter = 0.0f, exit = 1.0f;
    bool inter = true;
    for (int axis = 0; axis < 3; ++axis) {
        const float da = comp(dir, axis), sa = comp(s0, axis), ha = comp(he, axis);
        if (std::fabs(da) <= kTiny) {
            if (sa < -ha || sa > ha) inter = false;
            continue;
        }
        float first = (-ha - sa) / da;
        float second = (ha - sa) / da;
        if (first > second) {
            const float tmp = first;
            first = second;
            second = tmp;
        }
        enter = fmaxd(enter, first);
        exit = fmind(exit, second);
        if (enter > exit) inter = false;
    }
    if (inter) {
        t_out = clampd(0.5f * (enter + exit), 0.0f, 1.0f);
        const V3 core = add(s0, mul(t_out, dir));
        const V3 d = sub(he, vabs(core));
        int axis = 0;
        if (d.y < d.x) axis = 1;
        if (d.z < comp(d, axis)) axis = 2;
        const float sign = comp(core, axis) >= 0.0f ? 1.0f : -1.0f;
        V3 outward{0.0f, 0.0f, 0.0f};
        setcomp(outward, axis, sign);
        V3 bp = core;
        setcomp(bp, axis, sign * comp(he, axis));
        const V3 ln = neg(outward);
        normal = qrot(box.rotation, ln);
        cap_point = add(box.center, qrot(box.rotation, add(core, mul(cap.radius, ln))));
        box_point = add(box.center, qrot(box.rotation, bp));
        sep = -comp(d, axis) - cap.radius;
        feature = 27u + 2u * (uint32_t)axis + (sign > 0.0f ? 1u : 0u);
        return;
    }
    float best_sq = INFINITY;
    V3 best_seg{0.0f, 0.0f, 0.0f}, best_box{0.0f, 0.0f, 0.0f};
    float best_t = 0.0f;
    uint32_t best_feat = 0xffffffffu;
    for (uint32_t e = 0; e < 2u; ++e) {
        const V3 sp = e == 0u ? s0 : s1;
        const V3 bp{clampd(sp.x, -he.x, he.x), clampd(sp.y, -he.y, he.y), clampd(sp.z, -he.z, he.z)};
        consider_capsule_box(sp, bp, (float)e, e, best_sq, best_seg, best_box, best_t, best_feat);
    }
    uint32_t edge_feature = 2u;
    for (int va = 0; va < 3; ++va) {
        const int f0 = (va + 1) % 3, f1 = (va + 2) % 3;
        for (uint32_t signs = 0; signs < 4u; ++signs) {
            V3 e0{0.0f, 0.0f, 0.0f}, e1{0.0f, 0.0f, 0.0f};
            setcomp(e0, va, -comp(he, va));
            setcomp(e1, va, comp(he, va));
            const float sg0 = (signs & 1u) == 0u ? -1.0f : 1.0f;
            const float sg1 = (signs & 2u) == 0u ? -1.0f : 1.0f;
            setcomp(e0, f0, sg0 * comp(he, f0));
            setcomp(e1, f0, comp(e0, f0));
            setcomp(e0, f1, sg1 * comp(he, f1));
            setcomp(e1, f1, comp(e0, f1));
            const SegPair cl = closest_segments(s0, s1, e0, e1);
            consider_capsule_box(cl.pa, cl.pb, cl.ta, edge_feature, best_sq, best_seg, best_box,
                                 best_t, best_feat);
            ++edge_feature;
        }
    }
    const float dist = std::sqrt(fmaxd(best_sq, 0.0f));
    const V3 ln = dist > kTiny ? mul(1.0f / dist, sub(best_box, best_seg))
                               : qrot_inv(box.rotation, coincident_normal(cap_index, box_index));
    normal = qrot(box.rotation, ln);
    cap_point = add(box.center, qrot(box.rotation, add(best_seg, mul(cap.radius, ln))));
    box_point = add(box.center, qrot(box.rotation, best_box));
    sep = dist - cap.radius;
    t_out = best_t;
    feature = 64u + best_feat;
}

V3 box_axis(const Collider& box, int axis) {
    V3 l{0.0f, 0.0f, 0.0f};
    setcomp(l, axis, 1.0f);
    return qrot(box.rotation, l);
}

float box_projection_radius(const Collider& box, V3 axis) {
    return std::fabs(dot(axis, box_axis(box, 0))) * box.half_extents.x +
        std::fabs(dot(axis, box_axis(box, 1))) * box.half_extents.y +
        std::fabs(dot(axis, box_axis(box, 2))) * box.half_extents.z;
}

V3 box_vertex(const Collider& box, uint32_t v) {
    const V3 he = box.half_extents;
    const V3 l{(v & 1u) == 0u ? -he.x : he.x, (v & 2u) == 0u ? -he.y : he.y,
               (v & 4u) == 0u ? -he.z : he.z};
    return add(box.center, qrot(box.rotation, l));
}

bool point_in_inflated_box(V3 p, const Collider& box, float infl) {
    const V3 l = vabs(qrot_inv(box.rotation, sub(p, box.center)));
    return l.x <= box.half_extents.x + infl && l.y <= box.half_extents.y + infl &&
        l.z <= box.half_extents.z + infl;
}

V3 box_support(const Collider& box, V3 d) {
    V3 res = box.center;
    for (int axis = 0; axis < 3; ++axis) {
        const V3 basis = box_axis(box, axis);
        const float he = comp(box.half_extents, axis);
        res = add(res, mul(dot(d, basis) >= 0.0f ? he : -he, basis));
    }
    return res;
}

void append(RawBatch& out, const RawContact& c) {
    if (out.count < MR_RAW_CONTACTS_PER_PAIR) out.c[out.count++] = c;
}

// Collision.metal:1577-1720 (boxBoxContacts: SAT + inflated vertex tests).
void box_box(uint32_t ia, uint32_t ib, const Collider& A, const Collider& B, float accepted,
             RawBatch& out) {
    V3 axes[15];
    uint32_t nax = 0;
    for (int a = 0; a < 3; ++a) axes[nax++] = box_axis(A, a);
    for (int a = 0; a < 3; ++a) axes[nax++] = box_axis(B, a);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const V3 cr = cross(box_axis(A, a), box_axis(B, b));
            const float sq = dot(cr, cr);
            if (sq > 1.0e-12f) axes[nax++] = mul(1.0f / std::sqrt(sq), cr);
        }
    const V3 delta = sub(B.center, A.center);
    float best_sep = -INFINITY;
    V3 best_n = coincident_normal(ia, ib);
    uint32_t best_axis = 0;
    for (uint32_t k = 0; k < nax; ++k) {
        V3 axis = axes[k];
        const float proj = dot(delta, axis);
        if (proj < 0.0f || (proj == 0.0f && dot(axis, best_n) < 0.0f)) axis = neg(axis);
        const float sep = std::fabs(proj) - box_projection_radius(A, axis) -
            box_projection_radius(B, axis);
        if (sep > accepted) return;
        const float tie = 1.0e-7f * (1.0f + fmaxd(std::fabs(best_sep), std::fabs(sep)));
        if (!std::isfinite(best_sep) || sep > best_sep + tie ||
            (std::fabs(sep - best_sep) <= tie && k < best_axis)) {
            best_sep = sep;
            best_n = axis;
            best_axis = k;
        }
    }
    const float ra = box_projection_radius(A, best_n);
    const float rb = box_projection_radius(B, best_n);
    const float near_b = dot(B.center, best_n) - rb;
    const float far_a = dot(A.center, best_n) + ra;
    for (uint32_t v = 0; v < 8u; ++v) {
        const V3 pa = box_vertex(A, v);
        if (!point_in_inflated_box(pa, B, accepted)) continue;
        const float sep = near_b - dot(pa, best_n);
        if (sep <= accepted)
            append(out, make_contact(best_n, sep, pa, add(pa, mul(sep, best_n)),
                                     feature_key(MR_SHAPE_BOX, v),
                                     feature_key(MR_SHAPE_BOX, 128u + best_axis)));
    }
    for (uint32_t v = 0; v < 8u; ++v) {
        const V3 pb = box_vertex(B, v);
        if (!point_in_inflated_box(pb, A, accepted)) continue;
        const float sep = dot(pb, best_n) - far_a;
        if (sep <= accepted)
            append(out, make_contact(best_n, sep, sub(pb, mul(sep, best_n)), pb,
                                     feature_key(MR_SHAPE_BOX, 128u + best_axis),
                                     feature_key(MR_SHAPE_BOX, v)));
    }
    if (out.count == 0u) {
        const V3 pa = box_support(A, best_n);
        const V3 pb = box_support(B, neg(best_n));
        append(out, make_contact(best_n, dot(sub(pb, pa), best_n), pa, pb,
                                 feature_key(MR_SHAPE_BOX, 256u + best_axis),
                                 feature_key(MR_SHAPE_BOX, 256u + best_axis)));
    }
}

// Collision.metal:1415-1440 (appendFinitePlaneContact).
void append_surface_contact(RawBatch& out, bool surface_is_a, V3 surface_normal, V3 finite_point,
                            float sep, uint32_t finite_feature, uint32_t surface_feature) {
    const V3 sp = sub(finite_point, mul(sep, surface_normal));
    RawContact c = make_contact(neg(surface_normal), sep, finite_point, sp, finite_feature,
                                surface_feature);
    if (surface_is_a) c = swapped(c);
    append(out, c);
}

// Heightfield support-point test (first principles; see file header). A point
// x of the finite shape with radius rho is tested against the plane of the
// grid triangle below it.
void heightfield_point(const StepContext& ctx, RawBatch& out, bool hf_is_a, V3 x, float rho,
                       float accepted, uint32_t finite_feature) {
    const MRHeightfieldC& hf = *ctx.m.heightfield;
    const float s = hf.origin_spacing.z;
    const float fx = (x.x - hf.origin_spacing.x) / s;
    const float fy = (x.y - hf.origin_spacing.y) / s;
    if (!std::isfinite(fx) || !std::isfinite(fy)) return;
    int ix = (int)std::floor(clampd(fx, 0.0f, (float)(hf.nx - 2u)));
    int iy = (int)std::floor(clampd(fy, 0.0f, (float)(hf.ny - 2u)));
    const float tx = fx - (float)ix;
    const float ty = fy - (float)iy;
    const uint32_t nx = hf.nx;
    const float h00 = ctx.m.heights[(uint32_t)iy * nx + (uint32_t)ix];
    const float h10 = ctx.m.heights[(uint32_t)iy * nx + (uint32_t)ix + 1u];
    const float h01 = ctx.m.heights[((uint32_t)iy + 1u) * nx + (uint32_t)ix];
    const float h11 = ctx.m.heights[((uint32_t)iy + 1u) * nx + (uint32_t)ix + 1u];
    const V3 v00{hf.origin_spacing.x + (float)ix * s, hf.origin_spacing.y + (float)iy * s, h00};
    const uint32_t tri = tx >= ty ? 0u : 1u;
    V3 e1, e2;
    if (tri == 0u) {
        e1 = V3{s, 0.0f, h10 - h00};
        e2 = V3{s, s, h11 - h00};
    } else {
        e1 = V3{s, s, h11 - h00};
        e2 = V3{0.0f, s, h01 - h00};
    }
    const V3 n = normalize(cross(e1, e2));
    const float d = dot(n, sub(x, v00)) - rho;
    if (d <= accepted) {
        const uint32_t cell = (uint32_t)iy * (nx - 1u) + (uint32_t)ix;
        append_surface_contact(out, hf_is_a, n, sub(x, mul(rho, n)), d,
                               finite_feature, feature_key(MR_SHAPE_HEIGHTFIELD, 2u * cell + tri));
    }
}

// Collision.metal:4379-4830 (generateContacts) for the supported classes.
// ---- convex and mesh narrowphase (GJK, MPR, support-mapped patches, BVH4)
//
// Role of Collision.metal:2298-3480 (GJK, EPA/MPR, support patches) and
// :4097-4400 (BVH4 mesh traversal), restated independently. A finite
// collider is reduced to a core: a point set searched by scan for support
// points plus a radius (sphere: centre, radius r; capsule: its endpoints,
// radius r; box: its 8 vertices; convex: its scaled, transformed vertices;
// mesh triangle: its 3 vertices). Separated cores: GJK gives the closest
// points; overlapping cores: Minkowski portal refinement gives the normal and
// depth along the centre-to-centre ray. The normal n points from A to B.
// Contacts are a support-mapped patch: the points of each core within a slab
// of its support plane along +-n (its feature); feature points inside the
// other feature's polygon in the plane normal to n, or a feature edge
// clipped against the other polygon; a single pair of witness points when
// either feature is a point, or both are edges.

struct Core {
    uint32_t type;      // MR_SHAPE_* of the collider (MR_SHAPE_TRIANGLE_MESH for a triangle)
    uint32_t n;         // points
    float radius;
    V3 center;
    Q4 rotation;
    V3 scale;           // convex: vertex scale; box: half extents
    V3 p0, p1, p2;      // capsule endpoints / triangle vertices
    const mr_float4* verts;  // convex: the geometry's vertices
};

V3 core_point(const Core& c, uint32_t i) {
    if (c.type == MR_SHAPE_SPHERE) return c.center;
    if (c.type == MR_SHAPE_CAPSULE) return i == 0u ? c.p0 : c.p1;
    if (c.type == MR_SHAPE_TRIANGLE_MESH) return i == 0u ? c.p0 : (i == 1u ? c.p1 : c.p2);
    if (c.type == MR_SHAPE_BOX) {
        const V3 he = c.scale;
        const V3 l{(i & 1u) == 0u ? -he.x : he.x, (i & 2u) == 0u ? -he.y : he.y, (i & 4u) == 0u ? -he.z : he.z};
        return add(c.center, qrot(c.rotation, l));
    }
    const mr_float4& v = c.verts[i];
    return add(c.center, qrot(c.rotation, V3{c.scale.x * v.x, c.scale.y * v.y, c.scale.z * v.z}));
}

// Index of the point with the largest projection on d (the lowest on ties).
uint32_t core_support(const Core& c, V3 d) {
    uint32_t best = 0u;
    float bv = dot(d, core_point(c, 0u));
    for (uint32_t i = 1u; i < c.n; ++i) {
        const float v = dot(d, core_point(c, i));
        if (v > bv) {
            bv = v;
            best = i;
        }
    }
    return best;
}

Core make_core(const StepContext& ctx, uint32_t s, const Collider& col) {
    Core c;
    c.type = col.type;
    c.radius = 0.0f;
    c.center = col.center;
    c.rotation = col.rotation;
    c.scale = col.half_extents;
    c.p0 = col.p0;
    c.p1 = col.p1;
    c.p2 = V3{0.0f, 0.0f, 0.0f};
    c.verts = nullptr;
    if (col.type == MR_SHAPE_SPHERE) {
        c.n = 1u;
        c.radius = col.radius;
    } else if (col.type == MR_SHAPE_CAPSULE) {
        c.n = 2u;
        c.radius = col.radius;
    } else if (col.type == MR_SHAPE_BOX) {
        c.n = 8u;
    } else {
        const MRGeometryC& g = ctx.m.geometries[ctx.m.shapes[s].geometry];
        c.n = g.vertex_count;
        c.verts = ctx.m.geometry_vertices + g.vertex_offset;
    }
    return c;
}

struct Mv {
    V3 w, a, b;
    uint32_t ia, ib;
};

Mv minkowski_support(const Core& A, const Core& B, V3 d) {
    Mv m;
    m.ia = core_support(A, d);
    m.ib = core_support(B, neg(d));
    m.a = core_point(A, m.ia);
    m.b = core_point(B, m.ib);
    m.w = sub(m.a, m.b);
    return m;
}

// Closest point to the origin of simplex s[0..n) (n = 1..3), reducing it in
// place to the sub-simplex whose interior holds that point; lam are the
// barycentric weights of the kept points. Voronoi-region tests.
V3 simplex_closest(Mv* s, uint32_t& n, float* lam) {
    if (n == 1u) {
        lam[0] = 1.0f;
        return s[0].w;
    }
    if (n == 2u) {
        const V3 a = s[0].w, b = s[1].w, ab = sub(b, a);
        const float den = dot(ab, ab);
        const float t = den > kTiny ? -dot(a, ab) / den : 0.0f;
        if (t <= 0.0f) {
            n = 1u;
            lam[0] = 1.0f;
            return a;
        }
        if (t >= 1.0f) {
            s[0] = s[1];
            n = 1u;
            lam[0] = 1.0f;
            return b;
        }
        lam[0] = 1.0f - t;
        lam[1] = t;
        return add(a, mul(t, ab));
    }
    // triangle (Ericson, Real-Time Collision Detection 5.1.5), point = origin
    const V3 a = s[0].w, b = s[1].w, c = s[2].w;
    const V3 ab = sub(b, a), ac = sub(c, a), ap = neg(a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        n = 1u;
        lam[0] = 1.0f;
        return a;
    }
    const V3 bp = neg(b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        s[0] = s[1];
        n = 1u;
        lam[0] = 1.0f;
        return b;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float t = d1 / (d1 - d3);
        n = 2u;
        lam[0] = 1.0f - t;
        lam[1] = t;
        return add(a, mul(t, ab));
    }
    const V3 cp = neg(c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        s[0] = s[2];
        n = 1u;
        lam[0] = 1.0f;
        return c;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        s[1] = s[2];
        n = 2u;
        lam[0] = 1.0f - t;
        lam[1] = t;
        return add(a, mul(t, ac));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        s[0] = s[1];
        s[1] = s[2];
        n = 2u;
        lam[0] = 1.0f - t;
        lam[1] = t;
        return add(b, mul(t, sub(c, b)));
    }
    const float den = 1.0f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    lam[0] = 1.0f - v - w;
    lam[1] = v;
    lam[2] = w;
    return add(a, add(mul(v, ab), mul(w, ac)));
}

// True when the origin lies inside (or on) tetrahedron s[0..4).
bool tetra_contains_origin(const Mv* s) {
    for (int f = 0; f < 4; ++f) {
        const V3 a = s[(f + 1) % 4].w, b = s[(f + 2) % 4].w, c = s[(f + 3) % 4].w, d = s[f].w;
        const V3 nrm = cross(sub(b, a), sub(c, a));
        const float sd = dot(nrm, sub(d, a)), so = dot(nrm, neg(a));
        if (sd * so < 0.0f) return false;
    }
    return true;
}

struct ConvexHit {
    uint32_t status;   // 0 ok, 1 no result
    bool overlap;
    V3 n;              // A -> B
    float sep;         // core separation (negative: depth)
    V3 pa, pb;         // core witness points
};

// GJK distance between two cores; overlap when the origin is inside the
// Minkowski difference (then pa, pb are meaningless).
ConvexHit gjk(const Core& A, const Core& B) {
    ConvexHit r;
    r.status = 0u;
    r.overlap = false;
    Mv s[4];
    float lam[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    uint32_t n = 0u;
    V3 v = sub(A.center, B.center);
    if (!(dot(v, v) > kTiny)) v = V3{1.0f, 0.0f, 0.0f};
    for (uint32_t it = 0; it < 64u; ++it) {
        const Mv m = minkowski_support(A, B, neg(v));
        const float vv = dot(v, v);
        if (n > 0u && vv - dot(v, m.w) <= 1.0e-6f * vv) break;  // no further progress
        bool repeat = false;
        for (uint32_t k = 0; k < n; ++k) repeat = repeat || (s[k].ia == m.ia && s[k].ib == m.ib);
        if (repeat) break;
        s[n++] = m;
        if (n == 4u) {
            if (tetra_contains_origin(s)) {
                r.overlap = true;
                return r;
            }
            // the closest face on the origin's side replaces the tetrahedron
            float best = INFINITY;
            Mv keep[3];
            float kl[3];
            uint32_t kn = 0u;
            for (int f = 0; f < 4; ++f) {
                Mv t[3] = {s[(f + 1) % 4], s[(f + 2) % 4], s[(f + 3) % 4]};
                uint32_t tn = 3u;
                float tl[3];
                const V3 p = simplex_closest(t, tn, tl);
                const float d = dot(p, p);
                if (d < best) {
                    best = d;
                    kn = tn;
                    for (uint32_t k = 0; k < tn; ++k) {
                        keep[k] = t[k];
                        kl[k] = tl[k];
                    }
                    v = p;
                }
            }
            n = kn;
            for (uint32_t k = 0; k < kn; ++k) {
                s[k] = keep[k];
                lam[k] = kl[k];
            }
        } else {
            v = simplex_closest(s, n, lam);
        }
        if (dot(v, v) <= 1.0e-14f) {
            r.overlap = true;
            return r;
        }
    }
    V3 pa{0.0f, 0.0f, 0.0f}, pb{0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < n; ++k) {
        pa = add(pa, mul(lam[k], s[k].a));
        pb = add(pb, mul(lam[k], s[k].b));
    }
    const V3 d = sub(pb, pa);
    const float dist = length(d);
    // closer than float resolution at these coordinates: touching, the normal is MPR's
    const V3 ca = vabs(A.center), cb = vabs(B.center);
    const float coord = fmaxd(fmaxd(fmaxd(ca.x, ca.y), ca.z), fmaxd(fmaxd(cb.x, cb.y), cb.z));
    if (!(dist > 1.0e-6f * (1.0f + coord))) {
        r.overlap = true;
        return r;
    }
    r.pa = pa;
    r.pb = pb;
    r.sep = dist;
    r.n = dist > kTiny ? mul(1.0f / dist, d) : normalize(neg(v));
    return r;
}

// Minkowski portal refinement (XenoCollide) for overlapping cores: the
// boundary of A - B along the ray from its interior point (centre A - centre
// B) through the origin gives the normal (A -> B) and the depth.
ConvexHit mpr(const Core& A, const Core& B) {
    ConvexHit r;
    r.status = 1u;
    r.overlap = true;
    Mv v0;
    v0.a = A.center;
    v0.b = B.center;
    v0.w = sub(A.center, B.center);
    v0.ia = v0.ib = 0u;
    if (!(dot(v0.w, v0.w) > 1.0e-12f)) v0.w = V3{1.0e-5f, 0.0f, 0.0f};
    V3 n = neg(v0.w);
    Mv v1 = minkowski_support(A, B, n);
    n = cross(v1.w, v0.w);
    if (!(dot(n, n) > 1.0e-14f)) {
        // the origin lies on the segment v0-v1: the normal is along it
        const V3 d = normalize(sub(v1.w, v0.w));
        r.status = 0u;
        r.n = d;
        r.sep = -dot(v1.w, d);
        r.pa = v1.a;
        r.pb = sub(v1.a, mul(-r.sep, d));
        return r;
    }
    Mv v2 = minkowski_support(A, B, n);
    if (dot(v2.w, n) < 0.0f) return r;  // zero: touching, refined to depth 0
    n = cross(sub(v1.w, v0.w), sub(v2.w, v0.w));
    if (dot(n, v0.w) > 0.0f) {
        const Mv t = v1;
        v1 = v2;
        v2 = t;
        n = neg(n);
    }
    Mv v3;
    bool portal = false;
    for (uint32_t it = 0; it < 32u; ++it) {
        v3 = minkowski_support(A, B, n);
        if (dot(v3.w, n) < 0.0f) return r;
        if (dot(cross(v1.w, v3.w), v0.w) < 0.0f) {
            v2 = v3;
            n = cross(sub(v1.w, v0.w), sub(v3.w, v0.w));
            continue;
        }
        if (dot(cross(v3.w, v2.w), v0.w) < 0.0f) {
            v1 = v3;
            n = cross(sub(v3.w, v0.w), sub(v2.w, v0.w));
            continue;
        }
        portal = true;
        break;
    }
    if (!portal) return r;
    for (uint32_t it = 0; it < 32u; ++it) {
        n = cross(sub(v2.w, v1.w), sub(v3.w, v1.w));
        const float nl = length(n);
        if (!(nl > kTiny)) return r;
        n = mul(1.0f / nl, n);
        const Mv v4 = minkowski_support(A, B, n);
        const float delta = dot(sub(v4.w, v3.w), n);
        const float depth = dot(v4.w, n);
        if (delta <= 1.0e-5f * fmaxd(1.0f, std::fabs(depth)) || it == 31u) {
            // the origin's projection on the portal, by barycentric weights
            const float dist = dot(v1.w, n);
            const V3 p = mul(dist, n);
            const V3 e1 = sub(v2.w, v1.w), e2 = sub(v3.w, v1.w), ep = sub(p, v1.w);
            const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
            const float d20 = dot(ep, e1), d21 = dot(ep, e2);
            const float den = d00 * d11 - d01 * d01;
            float b2 = 0.0f, b3 = 0.0f;
            if (std::fabs(den) > kTiny) {
                b2 = (d11 * d20 - d01 * d21) / den;
                b3 = (d00 * d21 - d01 * d20) / den;
            }
            const float b1 = 1.0f - b2 - b3;
            r.status = 0u;
            r.n = n;
            r.sep = -depth;
            r.pa = add(add(mul(b1, v1.a), mul(b2, v2.a)), mul(b3, v3.a));
            r.pb = sub(r.pa, mul(depth, n));
            return r;
        }
        if (dot(cross(v4.w, v1.w), v0.w) < 0.0f) {
            if (dot(cross(v4.w, v2.w), v0.w) < 0.0f) v1 = v4;
            else v3 = v4;
        } else {
            if (dot(cross(v4.w, v3.w), v0.w) < 0.0f) v2 = v4;
            else v1 = v4;
        }
    }
    return r;
}

void plane_basis(V3 n, V3& t1, V3& t2) {
    const V3 an = vabs(n);
    V3 ref;
    if (an.x <= an.y && an.x <= an.z) ref = V3{1.0f, 0.0f, 0.0f};
    else if (an.y <= an.z) ref = V3{0.0f, 1.0f, 0.0f};
    else ref = V3{0.0f, 0.0f, 1.0f};
    t1 = normalize(cross(n, ref));
    t2 = cross(n, t1);
}

constexpr uint32_t kFeatureCap = 16u;

// Points of core c within `slab` of its support plane along d, in index
// order (at most kFeatureCap), and the plane's offset.
uint32_t core_feature(const Core& c, V3 d, float slab, uint32_t* idx, float& top) {
    top = -INFINITY;
    for (uint32_t i = 0; i < c.n; ++i) top = fmaxd(top, dot(d, core_point(c, i)));
    uint32_t k = 0u;
    for (uint32_t i = 0; i < c.n && k < kFeatureCap; ++i)
        if (dot(d, core_point(c, i)) >= top - slab) idx[k++] = i;
    return k;
}

float core_extent(const Core& c) {
    float e = 0.0f;
    for (uint32_t i = 0; i < c.n; ++i) e = fmaxd(e, length(sub(core_point(c, i), c.center)));
    return e;
}

// Convex hull of planar points (2D, gift wrapping from the lowest index of the
// leftmost point), as indices into pts, counter-clockwise. Returns its size.
uint32_t hull2d(const float (*pts)[2], uint32_t n, uint32_t* out) {
    uint32_t start = 0u;
    for (uint32_t i = 1; i < n; ++i)
        if (pts[i][0] < pts[start][0] || (pts[i][0] == pts[start][0] && pts[i][1] < pts[start][1])) start = i;
    uint32_t h = 0u, p = start;
    do {
        out[h++] = p;
        uint32_t q = p == 0u ? 1u : 0u;
        for (uint32_t r = 0; r < n; ++r) {
            if (r == p) continue;
            const float cr = (pts[q][0] - pts[p][0]) * (pts[r][1] - pts[p][1]) -
                (pts[q][1] - pts[p][1]) * (pts[r][0] - pts[p][0]);
            const float dq = (pts[q][0] - pts[p][0]) * (pts[q][0] - pts[p][0]) +
                (pts[q][1] - pts[p][1]) * (pts[q][1] - pts[p][1]);
            const float dr = (pts[r][0] - pts[p][0]) * (pts[r][0] - pts[p][0]) +
                (pts[r][1] - pts[p][1]) * (pts[r][1] - pts[p][1]);
            if (cr < 0.0f || (cr == 0.0f && dr > dq)) q = r;
        }
        p = q;
    } while (p != start && h < n);
    return h;
}

// Clips segment (x0,y0)-(x1,y1) against a counter-clockwise convex polygon;
// returns the kept parameter interval [t0, t1] (empty when t0 > t1).
void clip_segment(const float (*pts)[2], const uint32_t* poly, uint32_t m, float x0, float y0, float x1, float y1,
                  float& t0, float& t1) {
    t0 = 0.0f;
    t1 = 1.0f;
    for (uint32_t k = 0; k < m && t0 <= t1; ++k) {
        const float* a = pts[poly[k]];
        const float* b = pts[poly[(k + 1) % m]];
        const float ex = b[0] - a[0], ey = b[1] - a[1];
        const float f0 = ex * (y0 - a[1]) - ey * (x0 - a[0]);
        const float f1 = ex * (y1 - a[1]) - ey * (x1 - a[0]);
        if (f0 < 0.0f && f1 < 0.0f) {
            t0 = 1.0f;
            t1 = 0.0f;
        } else if (f0 < 0.0f) {
            t0 = fmaxd(t0, f0 / (f0 - f1));
        } else if (f1 < 0.0f) {
            t1 = fmind(t1, f0 / (f0 - f1));
        }
    }
}

// Keeps the kCap deepest contacts (replacing the shallowest, earlier kept on ties).
void append_deepest(RawBatch& out, const RawContact& c) {
    if (out.count < MR_RAW_CONTACTS_PER_PAIR) {
        out.c[out.count++] = c;
        return;
    }
    uint32_t worst = 0u;
    for (uint32_t k = 1; k < out.count; ++k)
        if (out.c[k].separation > out.c[worst].separation) worst = k;
    if (c.separation < out.c[worst].separation) out.c[worst] = c;
}

uint32_t core_feature_key(const Core& c, uint32_t local, uint32_t tri_feature) {
    if (c.type == MR_SHAPE_TRIANGLE_MESH) return feature_key(MR_SHAPE_TRIANGLE_MESH, tri_feature);
    return feature_key(c.type, local);
}

// One vertex of the clipped patch polygon: plane coordinates, the point on
// A's feature, its feature keys, and the tag of the polygon edge leaving it
// (k < kFeatureCap: A's hull edge k; kFeatureCap + k: B's clip line k).
struct PatchVertex {
    float x, y;
    V3 p;
    uint32_t feat_a, feat_b, tag;
};

constexpr uint32_t kPatchCap = 2u * kFeatureCap;

// Contacts of two cores within `accepted`: the GJK (separated) or MPR
// (overlapping) normal, then the features of both cores within a slab of
// their support planes along it. Two faces give the corners of their
// intersection polygon (A's face clipped by B's, Sutherland-Hodgman, in the
// plane normal to the contact normal); an edge on a face gives the clipped
// segment's ends; otherwise the witness pair. Deepest MR_RAW_CONTACTS_PER_PAIR kept.
void convex_contacts(const Core& A, const Core& B, uint32_t tri_a, uint32_t tri_b, float accepted, RawBatch& out) {
    ConvexHit h = gjk(A, B);
    if (h.overlap) h = mpr(A, B);
    if (h.status != 0u) return;
    const float sep = h.sep - A.radius - B.radius;
    if (!(sep <= accepted)) return;
    V3 n = h.n;
    const float ext_a = fmaxd(core_extent(A), 1.0e-3f), ext_b = fmaxd(core_extent(B), 1.0e-3f);
    uint32_t fa[kFeatureCap], fb[kFeatureCap];
    float top_a, top_b;
    const uint32_t na = core_feature(A, n, 2.0e-3f * ext_a, fa, top_a);
    const uint32_t nb = core_feature(B, neg(n), 2.0e-3f * ext_b, fb, top_b);
    float plane_b = -top_b;  // B's support plane: dot(n, x) = plane_b
    auto emit = [&](V3 pa, V3 pb, float s, uint32_t feat_a, uint32_t feat_b) {
        append_deepest(out, make_contact(n, s - A.radius - B.radius, add(pa, mul(A.radius, n)),
                                         sub(pb, mul(B.radius, n)), feat_a, feat_b));
    };
    const uint32_t face_a = core_feature_key(A, 128u + fa[0], tri_a), face_b = core_feature_key(B, 128u + fb[0], tri_b);
    if (na == 1u || nb == 1u || (na == 2u && nb == 2u)) {
        const uint32_t feat_a = na == 1u ? core_feature_key(A, fa[0], tri_a) : face_a;
        const uint32_t feat_b = nb == 1u ? core_feature_key(B, fb[0], tri_b) : face_b;
        emit(h.pa, h.pb, h.sep, feat_a, feat_b);
        return;
    }
    V3 t1, t2;
    plane_basis(n, t1, t2);
    float pa2[kFeatureCap][2], pb2[kFeatureCap][2];
    V3 wa[kFeatureCap], wb[kFeatureCap];
    for (uint32_t k = 0; k < na; ++k) {
        wa[k] = core_point(A, fa[k]);
        pa2[k][0] = dot(t1, wa[k]);
        pa2[k][1] = dot(t2, wa[k]);
    }
    for (uint32_t k = 0; k < nb; ++k) {
        wb[k] = core_point(B, fb[k]);
        pb2[k][0] = dot(t1, wb[k]);
        pb2[k][1] = dot(t2, wb[k]);
    }
    uint32_t hull_a[kFeatureCap], hull_b[kFeatureCap];
    const uint32_t ma = na >= 3u ? hull2d(pa2, na, hull_a) : 0u;
    const uint32_t mb = nb >= 3u ? hull2d(pb2, nb, hull_b) : 0u;
    const uint32_t before = out.count;
    if (ma >= 3u && mb >= 3u) {
        // B's face normal (Newell) replaces the query normal when within ~8 degrees of it
        V3 fnrm{0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < mb; ++k)
            fnrm = add(fnrm, cross(wb[hull_b[k]], wb[hull_b[(k + 1u) % mb]]));
        const float fl = length(fnrm);
        if (fl > kTiny && dot(fnrm, n) >= 0.99f * fl) {
            n = mul(1.0f / fl, fnrm);
            plane_b = INFINITY;
            for (uint32_t k = 0; k < nb; ++k) plane_b = fmind(plane_b, dot(n, wb[k]));
        }
        PatchVertex poly[kPatchCap], next[kPatchCap];
        uint32_t m = ma;
        for (uint32_t k = 0; k < ma; ++k) {
            const uint32_t i = hull_a[k];
            poly[k] = PatchVertex{pa2[i][0], pa2[i][1], wa[i], core_feature_key(A, fa[i], tri_a), face_b, k};
        }
        for (uint32_t k = 0; k < mb && m > 0u; ++k) {
            const float* a = pb2[hull_b[k]];
            const float* b = pb2[hull_b[(k + 1u) % mb]];
            const float ex = b[0] - a[0], ey = b[1] - a[1];
            uint32_t q = 0u;
            for (uint32_t j = 0; j < m; ++j) {
                const PatchVertex& P = poly[j];
                const PatchVertex& Q = poly[(j + 1u) % m];
                const float fp = ex * (P.y - a[1]) - ey * (P.x - a[0]);
                const float fq = ex * (Q.y - a[1]) - ey * (Q.x - a[0]);
                const bool in_p = fp >= 0.0f, in_q = fq >= 0.0f;
                if (in_p != in_q && q < kPatchCap) {
                    const float t = fp / (fp - fq);
                    PatchVertex c;
                    c.x = P.x + t * (Q.x - P.x);
                    c.y = P.y + t * (Q.y - P.y);
                    c.p = add(P.p, mul(t, sub(Q.p, P.p)));
                    if (P.tag < kFeatureCap) {
                        c.feat_a = core_feature_key(A, 64u + fa[hull_a[P.tag]], tri_a);
                        c.feat_b = core_feature_key(B, 64u + fb[hull_b[k]], tri_b);
                    } else {
                        // on B's earlier line j: the vertex B's lines j and k share
                        const uint32_t j_line = P.tag - kFeatureCap;
                        const uint32_t shared = j_line + 1u == k ? hull_b[k]
                            : (j_line == 0u && k + 1u == mb) ? hull_b[0] : MR_INVALID_INDEX;
                        c.feat_a = face_a;
                        c.feat_b = shared != MR_INVALID_INDEX ? core_feature_key(B, fb[shared], tri_b)
                                                              : core_feature_key(B, 64u + fb[hull_b[k]], tri_b);
                    }
                    c.tag = in_p ? kFeatureCap + k : P.tag;
                    next[q++] = c;
                }
                if (in_q && q < kPatchCap) next[q++] = Q;
            }
            m = q;
            for (uint32_t j = 0; j < m; ++j) poly[j] = next[j];
        }
        const float tol = 4.0e-6f * (1.0f + fmaxd(ext_a, ext_b));
        for (uint32_t j = 0; j < m; ++j) {
            bool dup = false;
            for (uint32_t i = 0; i < j; ++i) {
                const float dx = poly[j].x - poly[i].x, dy = poly[j].y - poly[i].y;
                dup = dup || dx * dx + dy * dy <= tol * tol;
            }
            if (dup) continue;
            const float s = plane_b - dot(n, poly[j].p);
            emit(poly[j].p, add(poly[j].p, mul(s, n)), s, poly[j].feat_a, poly[j].feat_b);
        }
    } else if (na == 2u && mb >= 3u) {
        float t0, tt1;
        clip_segment(pb2, hull_b, mb, pa2[0][0], pa2[0][1], pa2[1][0], pa2[1][1], t0, tt1);
        if (t0 <= tt1)
            for (int e = 0; e < 2; ++e) {
                const float t = e == 0 ? t0 : tt1;
                const V3 p = add(wa[0], mul(t, sub(wa[1], wa[0])));
                const float s = plane_b - dot(n, p);
                emit(p, add(p, mul(s, n)), s, core_feature_key(A, 64u + fa[e], tri_a), face_b);
            }
    } else if (nb == 2u && ma >= 3u) {
        float t0, tt1;
        clip_segment(pa2, hull_a, ma, pb2[0][0], pb2[0][1], pb2[1][0], pb2[1][1], t0, tt1);
        if (t0 <= tt1)
            for (int e = 0; e < 2; ++e) {
                const float t = e == 0 ? t0 : tt1;
                const V3 p = add(wb[0], mul(t, sub(wb[1], wb[0])));
                const float s = dot(n, p) - top_a;
                emit(sub(p, mul(s, n)), p, s, face_a, core_feature_key(B, 64u + fb[e], tri_b));
            }
    }
    if (out.count == before) emit(h.pa, h.pb, h.sep, face_a, face_b);
}

// A convex hull against a plane: every vertex within `accepted` (deepest 8).
void convex_plane_contacts(const Core& hull, const Collider& plane, bool plane_is_a, float accepted, RawBatch& out) {
    for (uint32_t i = 0; i < hull.n; ++i) {
        const V3 p = core_point(hull, i);
        const float sep = dot(plane.plane_normal, sub(p, plane.center));
        if (!(sep <= accepted)) continue;
        const V3 sp = sub(p, mul(sep, plane.plane_normal));
        RawContact c = make_contact(neg(plane.plane_normal), sep, p, sp, feature_key(MR_SHAPE_CONVEX, i),
                                    feature_key(MR_SHAPE_PLANE, 0u));
        if (plane_is_a) c = swapped(c);
        append_deepest(out, c);
    }
}

// A static mesh against a finite collider: BVH4 traversal with the finite
// collider's bounds in the mesh's local frame, then every overlapping
// triangle as a 3-point core. One-sided meshes keep contacts whose normal
// (finite -> mesh) opposes the triangle's front normal.
void mesh_contacts(const StepContext& ctx, uint32_t mesh_shape, const Collider& mesh, const Core& fin,
                   const Collider& fin_col, bool mesh_is_a, float accepted, RawBatch& out) {
    const MRShapeC& ms = ctx.m.shapes[mesh_shape];
    const MRGeometryC& g = ctx.m.geometries[ms.geometry];
    const mr_float4* vt = ctx.m.geometry_vertices + g.vertex_offset;
    const MRTriangleC* tris = ctx.m.triangles + g.triangle_offset;
    const MRBvhNodeC* nodes = ctx.m.bvh_nodes + g.node_offset;
    const bool two_sided = (ms.flags & MR_SHAPE_FLAG_MESH_TWO_SIDED) != 0u;
    const V3 scale = mesh.half_extents;
    // the finite collider's bounds, padded by the accepted distance, in mesh-local unscaled coordinates
    V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = 0; k < 8u; ++k) {
        const V3 w{(k & 1u) ? fin_col.upper.x + accepted : fin_col.lower.x - accepted,
                   (k & 2u) ? fin_col.upper.y + accepted : fin_col.lower.y - accepted,
                   (k & 4u) ? fin_col.upper.z + accepted : fin_col.lower.z - accepted};
        const V3 l = qrot_inv(mesh.rotation, sub(w, mesh.center));
        const V3 u{l.x / scale.x, l.y / scale.y, l.z / scale.z};
        lo = vmin(lo, u);
        hi = vmax(hi, u);
    }
    uint32_t stack[48];
    uint32_t top = 0u;
    stack[top++] = 0u;
    while (top > 0u) {
        const MRBvhNodeC& nd = nodes[stack[--top]];
        const float* lx = &nd.lower_x.x;
        const float* ly = &nd.lower_y.x;
        const float* lz = &nd.lower_z.x;
        const float* ux = &nd.upper_x.x;
        const float* uy = &nd.upper_y.x;
        const float* uz = &nd.upper_z.x;
        for (int c = 3; c >= 0; --c) {  // pushed in reverse: children leave the stack in index order
            if (nd.child[c] == MR_INVALID_INDEX || nd.count[c] != 0u) continue;
            if (lx[c] <= hi.x && ux[c] >= lo.x && ly[c] <= hi.y && uy[c] >= lo.y && lz[c] <= hi.z && uz[c] >= lo.z &&
                top < 48u)
                stack[top++] = nd.child[c];
        }
        for (uint32_t c = 0; c < 4u; ++c) {
            if (nd.child[c] == MR_INVALID_INDEX || nd.count[c] == 0u) continue;
            if (!(lx[c] <= hi.x && ux[c] >= lo.x && ly[c] <= hi.y && uy[c] >= lo.y && lz[c] <= hi.z && uz[c] >= lo.z))
                continue;
            for (uint32_t t = nd.child[c]; t < nd.child[c] + nd.count[c]; ++t) {
                const MRTriangleC& tr = tris[t];
                Core tc;
                tc.type = MR_SHAPE_TRIANGLE_MESH;
                tc.n = 3u;
                tc.radius = 0.0f;
                tc.rotation = mesh.rotation;
                tc.scale = scale;
                tc.verts = nullptr;
                V3 p[3];
                for (int k = 0; k < 3; ++k) {
                    const mr_float4& v = vt[tr.vertices[k]];
                    p[k] = add(mesh.center, qrot(mesh.rotation, V3{scale.x * v.x, scale.y * v.y, scale.z * v.z}));
                }
                tc.p0 = p[0];
                tc.p1 = p[1];
                tc.p2 = p[2];
                tc.center = mul(1.0f / 3.0f, add(add(p[0], p[1]), p[2]));
                RawBatch one;
                one.count = 0u;
                convex_contacts(fin, tc, 0u, tr.feature, accepted, one);
                const V3 front = cross(sub(p[1], p[0]), sub(p[2], p[0]));
                for (uint32_t k = 0; k < one.count; ++k) {
                    const RawContact& rc = one.c[k];
                    if (!two_sided && !(dot(rc.normal, front) < 0.0f)) continue;
                    append_deepest(out, mesh_is_a ? swapped(rc) : rc);
                }
            }
        }
    }
}

void generate_contacts(const StepContext& ctx, uint32_t ia, uint32_t ib, uint32_t cls,
                       const Collider& A, const Collider& B, RawBatch& out) {
    out.count = 0;
    const float accepted = (A.contact_offset + B.contact_offset) + pair_query_padding(A, B);
    switch (cls) {
    case MR_PAIR_SPHERE_SPHERE: {
        const V3 d = sub(B.center, A.center);
        const float dist = length(d);
        const V3 n = dist > kTiny ? mul(1.0f / dist, d) : coincident_normal(ia, ib);
        const float sep = dist - A.radius - B.radius;
        if (sep > accepted) return;
        append(out, make_contact(n, sep, add(A.center, mul(A.radius, n)),
                                 sub(B.center, mul(B.radius, n)),
                                 feature_key(MR_SHAPE_SPHERE, 0u), feature_key(MR_SHAPE_SPHERE, 0u)));
        return;
    }
    case MR_PAIR_SPHERE_CAPSULE: {
        const bool sphere_a = A.type == MR_SHAPE_SPHERE;
        const Collider& sp = sphere_a ? A : B;
        const Collider& cp = sphere_a ? B : A;
        const uint32_t si = sphere_a ? ia : ib, ci = sphere_a ? ib : ia;
        const SegPoint cl = closest_on_segment(sp.center, cp.p0, cp.p1);
        const V3 d = sub(cl.point, sp.center);
        const float dist = length(d);
        const V3 n = dist > kTiny ? mul(1.0f / dist, d)
                                  : stable_segment_normal(sub(cp.p1, cp.p0), V3{0.0f, 0.0f, 0.0f}, si, ci);
        const float sep = dist - sp.radius - cp.radius;
        if (sep <= accepted) {
            RawContact c = make_contact(n, sep, add(sp.center, mul(sp.radius, n)),
                                        sub(cl.point, mul(cp.radius, n)),
                                        feature_key(MR_SHAPE_SPHERE, 0u),
                                        feature_key(MR_SHAPE_CAPSULE, capsule_feature(cl.t)));
            if (!sphere_a) c = swapped(c);
            append(out, c);
        }
        return;
    }
    case MR_PAIR_CAPSULE_CAPSULE: {
        const SegPair cl = closest_segments(A.p0, A.p1, B.p0, B.p1);
        const V3 d = sub(cl.pb, cl.pa);
        const float dist = length(d);
        const V3 n = dist > kTiny ? mul(1.0f / dist, d)
                                  : stable_segment_normal(sub(A.p1, A.p0), sub(B.p1, B.p0), ia, ib);
        const float sep = dist - A.radius - B.radius;
        if (sep <= accepted)
            append(out, make_contact(n, sep, add(cl.pa, mul(A.radius, n)), sub(cl.pb, mul(B.radius, n)),
                                     feature_key(MR_SHAPE_CAPSULE, capsule_feature(cl.ta)),
                                     feature_key(MR_SHAPE_CAPSULE, capsule_feature(cl.tb))));
        return;
    }
    case MR_PAIR_SPHERE_BOX: {
        const bool sphere_a = A.type == MR_SHAPE_SPHERE;
        const Collider& sp = sphere_a ? A : B;
        const Collider& bx = sphere_a ? B : A;
        V3 n, bp;
        float sep;
        uint32_t feat;
        sphere_box(sp, bx, n, bp, sep, feat);
        if (sep <= accepted) {
            RawContact c = make_contact(n, sep, add(sp.center, mul(sp.radius, n)), bp,
                                        feature_key(MR_SHAPE_SPHERE, 0u), feature_key(MR_SHAPE_BOX, feat));
            if (!sphere_a) c = swapped(c);
            append(out, c);
        }
        return;
    }
    case MR_PAIR_CAPSULE_BOX: {
        const bool cap_a = A.type == MR_SHAPE_CAPSULE;
        const Collider& cp = cap_a ? A : B;
        const Collider& bx = cap_a ? B : A;
        V3 n, cpp, bp;
        float sep, t;
        uint32_t feat;
        capsule_box(cp, bx, cap_a ? ia : ib, cap_a ? ib : ia, n, cpp, bp, sep, t, feat);
        if (sep <= accepted) {
            RawContact c = make_contact(n, sep, cpp, bp, feature_key(MR_SHAPE_CAPSULE, capsule_feature(t)),
                                        feature_key(MR_SHAPE_BOX, feat));
            if (!cap_a) c = swapped(c);
            append(out, c);
        }
        return;
    }
    case MR_PAIR_BOX_BOX:
        box_box(ia, ib, A, B, accepted, out);
        return;
    case MR_PAIR_SPHERE_PLANE:
    case MR_PAIR_CAPSULE_PLANE:
    case MR_PAIR_BOX_PLANE: {
        const bool plane_a = A.type == MR_SHAPE_PLANE;
        const Collider& pl = plane_a ? A : B;
        const Collider& fs = plane_a ? B : A;
        const uint32_t pf = feature_key(MR_SHAPE_PLANE, 0u);
        if (cls == MR_PAIR_SPHERE_PLANE) {
            const V3 surf = sub(fs.center, mul(fs.radius, pl.plane_normal));
            const float sep = dot(pl.plane_normal, sub(surf, pl.center));
            if (sep <= accepted)
                append_surface_contact(out, plane_a, pl.plane_normal, surf, sep,
                                       feature_key(MR_SHAPE_SPHERE, 0u), pf);
        } else if (cls == MR_PAIR_CAPSULE_PLANE) {
            for (uint32_t e = 0; e < 2u; ++e) {
                const V3 ep = e == 0u ? fs.p0 : fs.p1;
                const V3 surf = sub(ep, mul(fs.radius, pl.plane_normal));
                const float sep = dot(pl.plane_normal, sub(surf, pl.center));
                if (sep <= accepted)
                    append_surface_contact(out, plane_a, pl.plane_normal, surf, sep,
                                           feature_key(MR_SHAPE_CAPSULE, e), pf);
            }
        } else {
            for (uint32_t v = 0; v < 8u; ++v) {
                const V3 wv = box_vertex(fs, v);
                const float sep = dot(pl.plane_normal, sub(wv, pl.center));
                if (sep <= accepted)
                    append_surface_contact(out, plane_a, pl.plane_normal, wv, sep,
                                           feature_key(MR_SHAPE_BOX, v), pf);
            }
        }
        return;
    }
    case MR_PAIR_HEIGHTFIELD: {
        const bool hf_a = A.type == MR_SHAPE_HEIGHTFIELD;
        const Collider& fs = hf_a ? B : A;
        if (fs.type == MR_SHAPE_SPHERE) {
            heightfield_point(ctx, out, hf_a, fs.center, fs.radius, accepted,
                              feature_key(MR_SHAPE_SPHERE, 0u));
        } else if (fs.type == MR_SHAPE_CAPSULE) {
            heightfield_point(ctx, out, hf_a, fs.p0, fs.radius, accepted, feature_key(MR_SHAPE_CAPSULE, 0u));
            heightfield_point(ctx, out, hf_a, fs.p1, fs.radius, accepted, feature_key(MR_SHAPE_CAPSULE, 1u));
            heightfield_point(ctx, out, hf_a, fs.center, fs.radius, accepted,
                              feature_key(MR_SHAPE_CAPSULE, 2u));
        } else if (fs.type == MR_SHAPE_BOX) {
            for (uint32_t v = 0; v < 8u; ++v)
                heightfield_point(ctx, out, hf_a, box_vertex(fs, v), 0.0f, accepted,
                                  feature_key(MR_SHAPE_BOX, v));
        }
        return;
    }
    case MR_PAIR_CONVEX: {
        if (A.type == MR_SHAPE_PLANE || B.type == MR_SHAPE_PLANE) {
            const bool plane_a = A.type == MR_SHAPE_PLANE;
            convex_plane_contacts(make_core(ctx, plane_a ? ib : ia, plane_a ? B : A), plane_a ? A : B, plane_a,
                                  accepted, out);
        } else {
            convex_contacts(make_core(ctx, ia, A), make_core(ctx, ib, B), 0u, 0u, accepted, out);
        }
        return;
    }
    case MR_PAIR_MESH: {
        const bool mesh_a = A.type == MR_SHAPE_TRIANGLE_MESH;
        mesh_contacts(ctx, mesh_a ? ia : ib, mesh_a ? A : B, make_core(ctx, mesh_a ? ib : ia, mesh_a ? B : A),
                      mesh_a ? B : A, mesh_a, accepted, out);
        return;
    }
    default:
        return;
    }
}

// ------------------------------------------------------------------ manifold

uint64_t cand_feature(const Candidate& c) {
    return ((uint64_t)c.point.feature_and_life[0] << 32u) | (uint64_t)c.point.feature_and_life[1];
}

bool cand_less(const Candidate& l, const Candidate& r) {
    if (l.separation != r.separation) return l.separation < r.separation;
    return cand_feature(l) < cand_feature(r);
}

V3 stable_tangent(V3 n) {
    const V3 a = vabs(n);
    V3 ref;
    if (a.x <= a.y && a.x <= a.z) ref = V3{1.0f, 0.0f, 0.0f};
    else if (a.y <= a.z) ref = V3{0.0f, 1.0f, 0.0f};
    else ref = V3{0.0f, 0.0f, 1.0f};
    return normalize(cross(ref, n));
}

// Collision.metal:5170-5280 (sortCandidates + reduceCandidates).
void reduce_candidates(Candidate* c, uint32_t& count, V3 n) {
    for (uint32_t i = 1; i < count; ++i) {
        const Candidate val = c[i];
        uint32_t d = i;
        while (d > 0u && cand_less(val, c[d - 1u])) {
            c[d] = c[d - 1u];
            --d;
        }
        c[d] = val;
    }
    if (count > MR_MANIFOLD_POINT_CAPACITY) {
        bool selected[kMaxCandidates] = {};
        uint32_t chosen[MR_MANIFOLD_POINT_CAPACITY];
        chosen[0] = 0u;
        selected[0] = true;
        uint32_t nch = 1u;
        for (uint32_t stage = 1; stage < MR_MANIFOLD_POINT_CAPACITY; ++stage) {
            uint32_t best = MR_INVALID_INDEX;
            float best_score = -1.0f;
            uint64_t best_feat = ~(uint64_t)0;
            for (uint32_t i = 0; i < count; ++i) {
                if (selected[i]) continue;
                float score;
                if (stage == 1u) {
                    const V3 d = sub(c[i].world_point, c[chosen[0]].world_point);
                    const V3 t = sub(d, mul(dot(d, n), n));
                    score = dot(t, t);
                } else if (stage == 2u) {
                    const V3 a = sub(c[chosen[1]].world_point, c[chosen[0]].world_point);
                    const V3 b = sub(c[i].world_point, c[chosen[0]].world_point);
                    score = std::fabs(dot(cross(a, b), n));
                } else {
                    score = INFINITY;
                    for (uint32_t s = 0; s < nch; ++s) {
                        const V3 d = sub(c[i].world_point, c[chosen[s]].world_point);
                        const V3 t = sub(d, mul(dot(d, n), n));
                        score = fmind(score, dot(t, t));
                    }
                }
                const uint64_t feat = cand_feature(c[i]);
                if (score > best_score + 1.0e-12f ||
                    (std::fabs(score - best_score) <= 1.0e-12f && feat < best_feat)) {
                    best = i;
                    best_score = score;
                    best_feat = feat;
                }
            }
            if (best != MR_INVALID_INDEX) {
                chosen[nch++] = best;
                selected[best] = true;
            }
        }
        Candidate red[MR_MANIFOLD_POINT_CAPACITY];
        for (uint32_t i = 0; i < nch; ++i) red[i] = c[chosen[i]];
        count = nch;
        for (uint32_t i = 0; i < count; ++i) c[i] = red[i];
    }
    for (uint32_t i = 1; i < count; ++i) {
        const Candidate val = c[i];
        const uint64_t f = cand_feature(val);
        uint32_t d = i;
        while (d > 0u && f < cand_feature(c[d - 1u])) {
            c[d] = c[d - 1u];
            --d;
        }
        c[d] = val;
    }
}

V3 world_from_anchor(const MRBodyStateC& b, Q4 rot, const mr_float4& anchor) {
    return add(f3(b.position), qrot(rot, f3(anchor)));
}

mr_float4 anchor_from_world(const MRBodyStateC& b, Q4 rot, V3 x) {
    const V3 l = qrot_inv(rot, sub(x, f3(b.position)));
    return mr_float4{l.x, l.y, l.z, 0.0f};
}

// Collision.metal:5290-5530 (buildPersistentWorldManifold).
bool build_manifold(const StepContext& ctx, const Work& wk, uint32_t env, const MRCompiledPairC& pair,
                    const RawBatch& raw, const MRManifoldHeaderC* old_h,
                    const MRManifoldPointC* old_p, MRManifoldHeaderC& out_h,
                    MRManifoldPointC* out_p, uint32_t& retained, uint32_t& fresh) {
    if (raw.count == 0u) return false;
    const MRShapeC& sa = ctx.m.shapes[pair.collider_a];
    const MRShapeC& sb = ctx.m.shapes[pair.collider_b];
    const MRBodyStateC& ba = wk.bodies[sa.body];
    const MRBodyStateC& bb = wk.bodies[sb.body];
    Q4 ra, rb;
    if (!checked_quat(q4(ba.orientation), ra) || !checked_quat(q4(bb.orientation), rb)) return false;
    uint32_t deepest = 0;
    for (uint32_t i = 1; i < raw.count; ++i) {
        const RawContact& l = raw.c[i];
        const RawContact& r = raw.c[deepest];
        if (l.separation < r.separation ||
            (l.separation == r.separation &&
             (l.feature_a < r.feature_a || (l.feature_a == r.feature_a && l.feature_b < r.feature_b))))
            deepest = i;
    }
    const V3 n = normalize(raw.c[deepest].normal);
    if (!finite3(n) || dot(n, n) < 0.9999f) return false;
    const float* th = &ctx.cfg->manifold_breaking_separation;  // sep, tangential, merge, cosine
    Candidate cand[kMaxCandidates];
    uint32_t nc = 0;
    bool compatible = false;
    V3 transported{0.0f, 0.0f, 0.0f};
    float old_age = 0.0f;
    if (old_h != nullptr) {
        const bool gen = old_h->generations_and_flags[0] == sa.slot_generation &&
            old_h->generations_and_flags[1] == sb.slot_generation &&
            old_h->pair_and_count[3] <= MR_MANIFOLD_POINT_CAPACITY;
        if (gen) {
            const V3 on = normalize(qrot(ra, f3(old_h->normal_and_age)));
            compatible = finite3(on) && dot(on, n) >= th[3];
            if (compatible) {
                const V3 ot = qrot(ra, f3(old_h->tangent_and_metric));
                transported = sub(ot, mul(dot(ot, n), n));
                old_age = old_h->normal_and_age.w;
                for (uint32_t p = 0; p < old_h->pair_and_count[3]; ++p) {
                    const MRManifoldPointC& op = old_p[p];
                    const V3 pa = world_from_anchor(ba, ra, op.local_anchor_a);
                    const V3 pb = world_from_anchor(bb, rb, op.local_anchor_b);
                    const V3 d = sub(pb, pa);
                    const float sep = dot(d, n);
                    const V3 tan = sub(d, mul(sep, n));
                    const float drift = length(tan);
                    if (sep > th[0] || drift > th[1]) continue;
                    Candidate c;
                    c.point = op;
                    c.point.feature_and_life[2] =
                        op.feature_and_life[2] == 0xffffffffu ? 0xffffffffu : op.feature_and_life[2] + 1u;
                    c.world_point = mul(0.5f, add(pa, pb));
                    c.separation = sep;
                    c.drift = drift;
                    cand[nc++] = c;
                    ++retained;
                }
            }
        }
    }
    const float merge_sq = th[2] * th[2];
    for (uint32_t k = 0; k < raw.count; ++k) {
        const RawContact& rc = raw.c[k];
        const mr_float4 la = anchor_from_world(ba, ra, rc.point_a);
        const mr_float4 lb = anchor_from_world(bb, rb, rc.point_b);
        uint32_t match = MR_INVALID_INDEX;
        for (uint32_t i = 0; i < nc; ++i) {
            if (cand[i].point.feature_and_life[0] == rc.feature_a &&
                cand[i].point.feature_and_life[1] == rc.feature_b) {
                match = i;
                break;
            }
        }
        if (match == MR_INVALID_INDEX) {
            for (uint32_t i = 0; i < nc; ++i) {
                const V3 da = sub(f3(cand[i].point.local_anchor_a), f3(la));
                const V3 db = sub(f3(cand[i].point.local_anchor_b), f3(lb));
                if (dot(da, da) + dot(db, db) <= 2.0f * merge_sq) {
                    match = i;
                    break;
                }
            }
        }
        if (match == MR_INVALID_INDEX) {
            if (nc >= kMaxCandidates) return false;
            match = nc++;
            std::memset(&cand[match], 0, sizeof(Candidate));
            cand[match].point.feature_and_life[2] = 0u;
            ++fresh;
        }
        cand[match].point.local_anchor_a = la;
        cand[match].point.local_anchor_b = lb;
        cand[match].point.feature_and_life[0] = rc.feature_a;
        cand[match].point.feature_and_life[1] = rc.feature_b;
        cand[match].point.feature_and_life[3] = 0u;
        cand[match].world_point = mul(0.5f, add(rc.point_a, rc.point_b));
        cand[match].separation = rc.separation;
        cand[match].drift = 0.0f;
    }
    if (nc == 0u) return false;
    reduce_candidates(cand, nc, n);
    V3 t = transported;
    if (!compatible || !finite3(t) || dot(t, t) <= 1.0e-12f) t = stable_tangent(n);
    else t = normalize(t);
    std::memset(&out_h, 0, sizeof(out_h));
    out_h.pair_and_count[0] = env;
    out_h.pair_and_count[1] = pair.collider_a;
    out_h.pair_and_count[2] = pair.collider_b;
    out_h.pair_and_count[3] = nc;
    out_h.generations_and_flags[0] = sa.slot_generation;
    out_h.generations_and_flags[1] = sb.slot_generation;
    const V3 ln = qrot_inv(ra, n);
    out_h.normal_and_age = mr_float4{ln.x, ln.y, ln.z, compatible ? fmind(old_age + 1.0f, 65535.0f) : 0.0f};
    float metric = 0.0f;
    for (uint32_t i = 0; i < nc; ++i) {
        out_p[i] = cand[i].point;
        metric = fmaxd(metric, fmaxd(fmaxd(cand[i].separation, 0.0f), cand[i].drift));
    }
    for (uint32_t i = nc; i < MR_MANIFOLD_POINT_CAPACITY; ++i) std::memset(&out_p[i], 0, sizeof(MRManifoldPointC));
    const V3 lt = qrot_inv(ra, t);
    out_h.tangent_and_metric = mr_float4{lt.x, lt.y, lt.z, metric};
    return true;
}

// ------------------------------------------------------------- cone solve

// Dot product of two generalized vectors (length n <= 64) in a fixed
// pairwise order: x[i] = a[i] * b[i] (+0 past n), then x[i] += x[i ^ d] for
// d = 1, 2, 4, ..., 32, i.e. ((r0 + r1) + (r2 + r3)) over the four 16-lane
// row sums, each row summed pairwise from adjacent lanes outwards. This is
// the order of a 64-lane butterfly reduction, so the device reproduces it bit
// for bit with one DoF per lane.
float fold_dot(const float* a, const float* b, uint32_t n) {
    float x[64], y[64];
    for (uint32_t i = 0; i < 64u; ++i) x[i] = i < n ? a[i] * b[i] : 0.0f;
    for (uint32_t d = 1u; d < 64u; d <<= 1) {
        for (uint32_t i = 0; i < 64u; ++i) y[i] = x[i] + x[i ^ d];
        for (uint32_t i = 0; i < 64u; ++i) x[i] = y[i];
    }
    return x[0];
}

// MetalWorldContact.metal:482-509 (projectFrictionCone).
void project_cone(float lam[3], float mu_u, float mu_v, float max_normal) {
    lam[0] = fmaxd(lam[0], 0.0f);
    if (max_normal > 0.0f) lam[0] = fmind(lam[0], max_normal);
    const float lu = mu_u * lam[0];
    const float lv = mu_v * lam[0];
    if (!(lu > 0.0f) || !(lv > 0.0f)) {
        lam[1] = 0.0f;
        lam[2] = 0.0f;
        return;
    }
    const float ns = (lam[1] * lam[1]) / (lu * lu) + (lam[2] * lam[2]) / (lv * lv);
    if (ns > 1.0f) {
        const float s = 1.0f / std::sqrt(ns);
        lam[1] *= s;
        lam[2] *= s;
    }
}

// MetalWorldContact.metal:540-623 (invert3x3: symmetrize, scale-normalize,
// one-percent CFM floor, adjugate inverse).
bool invert3x3(const float m[3][3], float inv[3][3]) {
    float scale = 0.0f;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            if (!std::isfinite(m[r][c])) return false;
            scale = fmaxd(scale, std::fabs(m[r][c]));
        }
    if (!(scale > kMatrixFloor)) return false;
    float g[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            g[r][c] = 0.5f * (m[r][c] + m[c][r]) / scale;
            if (r == c) g[r][c] += kContactRegularization;
        }
    const float c00 = g[1][1] * g[2][2] - g[1][2] * g[2][1];
    const float c01 = g[1][2] * g[2][0] - g[1][0] * g[2][2];
    const float c02 = g[1][0] * g[2][1] - g[1][1] * g[2][0];
    const float det = g[0][0] * c00 + g[0][1] * c01 + g[0][2] * c02;
    if (!(det > kMatrixFloor) || !std::isfinite(det)) return false;
    const float rcp = 1.0f / (det * scale);
    inv[0][0] = c00 * rcp;
    inv[0][1] = (g[0][2] * g[2][1] - g[0][1] * g[2][2]) * rcp;
    inv[0][2] = (g[0][1] * g[1][2] - g[0][2] * g[1][1]) * rcp;
    inv[1][0] = c01 * rcp;
    inv[1][1] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) * rcp;
    inv[1][2] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) * rcp;
    inv[2][0] = c02 * rcp;
    inv[2][1] = (g[0][1] * g[2][0] - g[0][0] * g[2][1]) * rcp;
    inv[2][2] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) * rcp;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(inv[r][c])) return false;
    return true;
}

V3 inv_inertia_mul(const MRBodyStateC& b, V3 x) {
    return V3{dot(f3(b.inverse_inertia_world_row0), x), dot(f3(b.inverse_inertia_world_row1), x),
              dot(f3(b.inverse_inertia_world_row2), x)};
}

// Velocity change of point x per unit impulse d on a dynamic scene body
// (MetalWorldContact.metal:446-462 scenePointResponse).
V3 scene_response(const MRBodyStateC& b, V3 x, V3 d) {
    const V3 lever = sub(x, f3(b.position));
    const V3 dw = inv_inertia_mul(b, cross(lever, d));
    return add(mul(b.linear_velocity_inv_mass.w, d), cross(dw, lever));
}

// Velocity change at point xc of body b from a unit impulse along d at point xj
// (the scene part of typedNormalCrossContactResponse,
// MetalWorldContact.metal:6740-6784).
V3 scene_cross_response(const MRBodyStateC& b, V3 xc, V3 xj, V3 d) {
    const V3 dw = inv_inertia_mul(b, cross(sub(xj, f3(b.position)), d));
    return add(mul(b.linear_velocity_inv_mass.w, d), cross(dw, sub(xc, f3(b.position))));
}

V3 scene_point_velocity(const MRBodyStateC& b, V3 x) {
    return add(f3(b.linear_velocity_inv_mass), cross(f3(b.angular_velocity), sub(x, f3(b.position))));
}

void apply_scene_impulse(MRBodyStateC& b, V3 x, V3 imp) {
    const V3 lever = sub(x, f3(b.position));
    const V3 dv = mul(b.linear_velocity_inv_mass.w, imp);
    b.linear_velocity_inv_mass.x += dv.x;
    b.linear_velocity_inv_mass.y += dv.y;
    b.linear_velocity_inv_mass.z += dv.z;
    const V3 dw = inv_inertia_mul(b, cross(lever, imp));
    b.angular_velocity.x += dw.x;
    b.angular_velocity.y += dw.y;
    b.angular_velocity.z += dw.z;
}

bool is_dynamic_scene(const StepContext& ctx, uint32_t body) {
    return body >= ctx.m.header->link_count && ctx.m.bodies[body].motion_type == MR_MOTION_DYNAMIC;
}

bool is_link(const StepContext& ctx, uint32_t body) { return body < ctx.m.header->link_count; }

// ---------------------------------------------------------------- helpers

void world_inverse_inertia(const MRBodyC& b, Q4 rot, MRBodyStateC& s) {
    const V3 c0 = qrot(rot, V3{1.0f, 0.0f, 0.0f});
    const V3 c1 = qrot(rot, V3{0.0f, 1.0f, 0.0f});
    const V3 c2 = qrot(rot, V3{0.0f, 0.0f, 1.0f});
    const float R[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
    const float I[3][3] = {
        {b.inverse_inertia_row0.x, b.inverse_inertia_row0.y, b.inverse_inertia_row0.z},
        {b.inverse_inertia_row1.x, b.inverse_inertia_row1.y, b.inverse_inertia_row1.z},
        {b.inverse_inertia_row2.x, b.inverse_inertia_row2.y, b.inverse_inertia_row2.z}};
    float out[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            float acc = 0.0f;
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2) acc += R[r][a] * I[a][b2] * R[c][b2];
            out[r][c] = acc;
        }
    s.inverse_inertia_world_row0 = mr_float4{out[0][0], out[0][1], out[0][2], 0.0f};
    s.inverse_inertia_world_row1 = mr_float4{out[1][0], out[1][1], out[1][2], 0.0f};
    s.inverse_inertia_world_row2 = mr_float4{out[2][0], out[2][1], out[2][2], 0.0f};
}

// ------------------------------------------------------------- the microstep

struct EnvState {
    float* q;
    float* v;
    MRBodyStateC* scene;
    MRManifoldHeaderC* mh;   // [capacity]
    MRManifoldPointC* mp;    // [capacity * 4]
    uint32_t* mcount;
};

// ---- continuous collision: the hybrid mode's time-of-impact query
// (mr_world_resolve_ccd and the event select kernel, Collision.metal:6707-6950,
// 7157-7330, 7657-7800). Over the remaining time T of a microstep every body
// moves rigidly from the segment's start: links with the COM and angular
// velocity of the current kinematics, dynamic scene bodies with their velocity
// plus T * gravity, kinematic bodies with theirs, static bodies not at all;
// rotations follow exp(t w). Pairs starting within their contact distance are
// left to the segment's speculative contacts. Two spheres get the closed-form
// TOI of their linearly moving centres; other pairs conservative advancement,
// stepping t by distance / (closing-speed bound |v_a - v_b| + |w_a| r_a +
// |w_b| r_b, or the normal speed against a static plane) until the pair is
// within the distance tolerance. The earliest (TOI, pair) ends the segment.

struct CcdMotion {
    V3 p0;
    Q4 r0;
    V3 v, w;
};

CcdMotion ccd_motion(const StepContext& ctx, const Work& wk, const EnvState& st, uint32_t b, float T) {
    const uint32_t nlinks = ctx.m.header->link_count;
    CcdMotion m;
    if (b < nlinks) {
        m.p0 = wk.pos[b];
        m.r0 = wk.rot[b];
        m.v = wk.vl[b];
        m.w = wk.w[b];
        return m;
    }
    const MRBodyStateC& s = st.scene[b - nlinks];
    m.p0 = f3(s.position);
    m.r0 = q4(s.orientation);
    const uint32_t mt = ctx.m.bodies[b].motion_type;
    if (mt == MR_MOTION_DYNAMIC) {
        m.v = add(f3(s.linear_velocity_inv_mass), mul(T, ctx.m.gravity));
        m.w = f3(s.angular_velocity);
    } else if (mt == MR_MOTION_KINEMATIC) {
        m.v = f3(s.linear_velocity_inv_mass);
        m.w = f3(s.angular_velocity);
    } else {
        m.v = V3{0.0f, 0.0f, 0.0f};
        m.w = V3{0.0f, 0.0f, 0.0f};
    }
    return m;
}

bool ccd_collider(const StepContext& ctx, uint32_t s, const CcdMotion& m, float t, Collider& c) {
    const V3 p = add(m.p0, mul(t, m.v));
    Q4 r;
    if (!normalized_quat(qmul(rotation_increment(mul(t, m.w)), m.r0), r)) return false;
    return project_collider_pose(ctx, s, p, r, c);
}

float ccd_reach(const MRShapeC& sh) {
    const float offset = length(f3(sh.local_position));
    if (sh.shape_type == MR_SHAPE_SPHERE) return offset + sh.dimensions.x;
    if (sh.shape_type == MR_SHAPE_CAPSULE) return offset + sh.dimensions.x + sh.dimensions.y;
    if (sh.shape_type == MR_SHAPE_BOX) return offset + length(f3(sh.dimensions));
    return INFINITY;
}

float ccd_motion_bound(const CcdMotion& m, const MRShapeC& sh, float T) {
    const float wl = length(m.w), vl = length(m.v);
    if (!(wl > 0.0f) && !(vl > 0.0f)) return 0.0f;
    const float reach = wl > 0.0f ? ccd_reach(sh) : 0.0f;
    return T * (vl + wl * reach);
}

float box_box_separation(const Collider& A, const Collider& B) {
    const V3 d = sub(B.center, A.center);
    float best = -INFINITY;
    for (int i = 0; i < 3; ++i) {
        const V3 ax = box_axis(A, i);
        best = fmaxd(best, std::fabs(dot(d, ax)) - comp(A.half_extents, i) - box_projection_radius(B, ax));
    }
    for (int j = 0; j < 3; ++j) {
        const V3 ax = box_axis(B, j);
        best = fmaxd(best, std::fabs(dot(d, ax)) - box_projection_radius(A, ax) - comp(B.half_extents, j));
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const V3 c = cross(box_axis(A, i), box_axis(B, j));
            const float l = length(c);
            if (!(l > 1.0e-6f)) continue;
            const V3 ax = mul(1.0f / l, c);
            best = fmaxd(best, std::fabs(dot(d, ax)) - box_projection_radius(A, ax) - box_projection_radius(B, ax));
        }
    return best;
}

// Separation at any distance: the analytic pair functions without the
// contact-distance cut-off; false for pairs without such a query (heightfield).
bool ccd_separation(uint32_t ia, uint32_t ib, uint32_t cls, const Collider& A, const Collider& B, float& d) {
    if (cls == MR_PAIR_SPHERE_SPHERE) {
        d = length(sub(B.center, A.center)) - A.radius - B.radius;
    } else if (cls == MR_PAIR_SPHERE_CAPSULE) {
        const Collider& sp = A.type == MR_SHAPE_SPHERE ? A : B;
        const Collider& cp = A.type == MR_SHAPE_SPHERE ? B : A;
        const SegPoint cl = closest_on_segment(sp.center, cp.p0, cp.p1);
        d = length(sub(cl.point, sp.center)) - sp.radius - cp.radius;
    } else if (cls == MR_PAIR_CAPSULE_CAPSULE) {
        const SegPair cl = closest_segments(A.p0, A.p1, B.p0, B.p1);
        d = length(sub(cl.pb, cl.pa)) - A.radius - B.radius;
    } else if (cls == MR_PAIR_SPHERE_BOX) {
        const bool sphere_a = A.type == MR_SHAPE_SPHERE;
        V3 n, bp;
        uint32_t feat;
        sphere_box(sphere_a ? A : B, sphere_a ? B : A, n, bp, d, feat);
    } else if (cls == MR_PAIR_CAPSULE_BOX) {
        const bool cap_a = A.type == MR_SHAPE_CAPSULE;
        V3 n, cpp, bp;
        float t;
        uint32_t feat;
        capsule_box(cap_a ? A : B, cap_a ? B : A, cap_a ? ia : ib, cap_a ? ib : ia, n, cpp, bp, d, t, feat);
    } else if (cls == MR_PAIR_BOX_BOX) {
        d = box_box_separation(A, B);
    } else if (cls == MR_PAIR_SPHERE_PLANE || cls == MR_PAIR_CAPSULE_PLANE || cls == MR_PAIR_BOX_PLANE) {
        const Collider& pl = A.type == MR_SHAPE_PLANE ? A : B;
        const Collider& fs = A.type == MR_SHAPE_PLANE ? B : A;
        if (cls == MR_PAIR_SPHERE_PLANE) {
            d = dot(pl.plane_normal, sub(fs.center, pl.center)) - fs.radius;
        } else if (cls == MR_PAIR_CAPSULE_PLANE) {
            d = fmind(dot(pl.plane_normal, sub(fs.p0, pl.center)), dot(pl.plane_normal, sub(fs.p1, pl.center))) -
                fs.radius;
        } else {
            d = INFINITY;
            for (uint32_t vi = 0; vi < 8u; ++vi) d = fmind(d, dot(pl.plane_normal, sub(box_vertex(fs, vi), pl.center)));
        }
    } else {
        return false;
    }
    return std::isfinite(d);
}

enum : uint32_t { kCcdNone = 0u, kCcdImpact = 1u, kCcdUnsafe = 2u };

uint32_t ccd_pair(const StepContext& ctx, const Work& wk, const EnvState& st, uint32_t k, float T, float& toi) {
    const MRWorldStepConfigC& cfg = *ctx.cfg;
    const MRCompiledPairC& pr = ctx.m.pairs[k];
    const MRShapeC& sa = ctx.m.shapes[pr.collider_a];
    const MRShapeC& sb = ctx.m.shapes[pr.collider_b];
    if (((sa.flags | sb.flags) & MR_SHAPE_FLAG_ENABLE_CCD) == 0u ||
        ((sa.flags | sb.flags) & MR_SHAPE_FLAG_SIMULATION_DISABLED) != 0u)
        return kCcdNone;
    const CcdMotion ma = ccd_motion(ctx, wk, st, sa.body, T), mb = ccd_motion(ctx, wk, st, sb.body, T);
    const float tol = cfg.ccd_distance_tolerance;
    const float motion_a = ccd_motion_bound(ma, sa, T), motion_b = ccd_motion_bound(mb, sb, T);
    const bool speculative_safe = motion_a <= sa.contact_rest_radius.x + tol && motion_b <= sb.contact_rest_radius.x + tol;
    Collider a0, b0, a1, b1;
    if (!ccd_collider(ctx, pr.collider_a, ma, 0.0f, a0) || !ccd_collider(ctx, pr.collider_b, mb, 0.0f, b0) ||
        !ccd_collider(ctx, pr.collider_a, ma, T, a1) || !ccd_collider(ctx, pr.collider_b, mb, T, b1))
        return speculative_safe ? kCcdNone : kCcdUnsafe;
    Collider a_sw = a0, b_sw = b0;
    a_sw.lower = vmin(a0.lower, a1.lower);
    a_sw.upper = vmax(a0.upper, a1.upper);
    b_sw.lower = vmin(b0.lower, b1.lower);
    b_sw.upper = vmax(b0.upper, b1.upper);
    if (!pair_may_overlap(a_sw, b_sw)) return kCcdNone;
    float d;
    if (!ccd_separation(pr.collider_a, pr.collider_b, pr.pair_class, a0, b0, d))
        return speculative_safe ? kCcdNone : kCcdUnsafe;
    // within the contact distance the segment's speculative contacts hold the pair
    if (d <= fmaxd(tol, sa.contact_rest_radius.x + sb.contact_rest_radius.x)) return kCcdNone;
    if (pr.pair_class == MR_PAIR_SPHERE_SPHERE) {
        const V3 rel0 = sub(b0.center, a0.center);
        const V3 drel = sub(sub(b1.center, b0.center), sub(a1.center, a0.center));
        const float radius = a0.radius + b0.radius;
        const float aa = dot(drel, drel), bb = 2.0f * dot(rel0, drel), cc = dot(rel0, rel0) - radius * radius;
        const float disc = bb * bb - 4.0f * aa * cc;
        if (aa > kTiny && disc >= 0.0f) {
            const float alpha = (-bb - std::sqrt(disc)) / (2.0f * aa);
            if (alpha >= 0.0f && alpha <= 1.0f) {
                toi = alpha * T;
                return kCcdImpact;
            }
        }
        return kCcdNone;
    }
    // Closing-speed bound: |v_a - v_b| + |w_a| r_a + |w_b| r_b; against a
    // static plane only the motion along its normal closes the gap.
    const bool plane_a = a0.type == MR_SHAPE_PLANE, plane_b = b0.type == MR_SHAPE_PLANE;
    float total;
    if ((plane_a && !(motion_a > 0.0f)) || (plane_b && !(motion_b > 0.0f))) {
        const V3 n = plane_a ? a0.plane_normal : b0.plane_normal;
        const CcdMotion& fm = plane_a ? mb : ma;
        const float wl = length(fm.w);
        const float reach = wl > 0.0f ? ccd_reach(plane_a ? sb : sa) : 0.0f;
        total = T * (std::fabs(dot(fm.v, n)) + wl * reach);
    } else {
        const float wa = length(ma.w), wb = length(mb.w);
        const float spin = (wa > 0.0f ? wa * ccd_reach(sa) : 0.0f) + (wb > 0.0f ? wb * ccd_reach(sb) : 0.0f);
        total = T * (length(sub(ma.v, mb.v)) + spin);
    }
    total = fmaxd(total, tol);
    const float min_alpha = fmind(1.0f, cfg.ccd_minimum_advance / T);
    float alpha = 0.0f;
    for (uint32_t it = 0; it < cfg.max_ccd_iterations; ++it) {
        if (d <= tol) {
            toi = alpha * T;
            return kCcdImpact;
        }
        alpha += fmaxd(d / total, min_alpha);
        if (alpha > 1.0f) return kCcdNone;
        Collider a, b;
        if (!ccd_collider(ctx, pr.collider_a, ma, alpha * T, a) || !ccd_collider(ctx, pr.collider_b, mb, alpha * T, b) ||
            !ccd_separation(pr.collider_a, pr.collider_b, pr.pair_class, a, b, d))
            return speculative_safe ? kCcdNone : kCcdUnsafe;
    }
    if (d <= tol || !speculative_safe) {
        // Within the tolerance: the time of impact. Otherwise the iterations ran
        // out with the pair separated up to alpha T: the segment ends there and
        // the pair is queried again from that state.
        toi = alpha * T;
        return kCcdImpact;
    }
    return kCcdNone;
}

// The segment the hybrid mode runs next from the current state: its duration
// (the remaining time, or the earliest TOI clamped below by the time
// tolerance) and whether it ends at an impact. False with f set when a pair is
// unresolved and not covered by the speculative contacts, or when an impact
// would need more than max_ccd_events segments in this microstep; the failure
// of the smaller pair index wins.
bool ccd_segment(const StepContext& ctx, EnvState& st, Work& wk, float remaining, uint32_t segment, Failure& f,
                 bool& impact, float& duration) {
    impact = false;
    duration = remaining;
    if (!forward_kinematics(ctx, st.q, st.v, wk, f)) return false;
    uint32_t best_bits = ~0u, best_pair = ~0u, unsafe = ~0u;
    for (uint32_t k = 0; k < ctx.m.header->pair_count; ++k) {
        float toi = 0.0f;
        const uint32_t r = ccd_pair(ctx, wk, st, k, remaining, toi);
        if (r == kCcdImpact) {
            const float t = fmaxd(toi, 0.0f);
            uint32_t bits;
            std::memcpy(&bits, &t, sizeof(bits));
            if (bits < best_bits || (bits == best_bits && k < best_pair)) {
                best_bits = bits;
                best_pair = k;
            }
        } else if (r == kCcdUnsafe && unsafe == ~0u) {
            unsafe = k;
        }
    }
    uint32_t fail_pair = ~0u, fail_code = MR_STEP_SUCCESS;
    if (unsafe != ~0u) {
        fail_pair = unsafe;
        fail_code = MR_STEP_DID_NOT_CONVERGE;
    }
    if (best_bits != ~0u) {
        float t;
        std::memcpy(&t, &best_bits, sizeof(t));
        const float tau = fmaxd(t, ctx.cfg->ccd_time_tolerance);
        if (tau < remaining - ctx.cfg->ccd_time_tolerance) {
            if (segment >= ctx.cfg->max_ccd_events && best_pair < fail_pair) {
                fail_pair = best_pair;
                fail_code = MR_STEP_CCD_EVENT_BUDGET_EXHAUSTED;
            }
            impact = true;
            duration = tau;
        }
    }
    if (fail_code != MR_STEP_SUCCESS) {
        f.code = fail_code;
        f.index = fail_pair;
        return false;
    }
    return true;
}

struct MicroOutputs {
    MRContactStatusC* cstatus;
    MRPairRecordC* pairs;  // evidence (capacity = pair_capacity) or null
    uint32_t* pair_count;
};

// One microstep, or one segment of a hybrid-CCD microstep, of duration ctx.h;
// `damp` applies the scene bodies' per-microstep velocity damping (the first
// segment of a microstep).
bool microstep(const StepContext& ctx, EnvState& st, const float* command, bool first_sub, bool final_sub,
               bool capture, MicroOutputs& out, Work& wk, Failure& f, std::vector<Constraint>& cons,
               std::vector<float>& jrow, std::vector<float>& resp, bool damp) {
    const MRModelHeaderC& hd = *ctx.m.header;
    const uint32_t nv = hd.nv;
    const uint32_t nlinks = hd.link_count;
    const float h = ctx.h;
    MRContactStatusC& cs = *out.cstatus;

    // 1. drives: implicit drives from the current state every microstep;
    // effort commands once per control step (held in wk.effort)
    if (first_sub || ctx.cfg->actuation_mode == MR_ACTUATION_IMPLICIT_POSITION_DRIVE)
        refresh_drives(ctx, st.q, st.v, command, wk);
    // 2. kinematics + ABA candidate velocity
    if (!forward_kinematics(ctx, st.q, st.v, wk, f)) return false;
    if (!aba(ctx, st.v, wk, f)) return false;

    // 3. body states: links from FK, scene bodies predicted
    for (uint32_t i = 0; i < nlinks; ++i) {
        MRBodyStateC& b = wk.bodies[i];
        std::memset(&b, 0, sizeof(b));
        b.position = mr_float4{wk.pos[i].x, wk.pos[i].y, wk.pos[i].z, 1.0f};
        b.orientation = mr_float4{wk.rot[i].x, wk.rot[i].y, wk.rot[i].z, wk.rot[i].w};
        b.flags_indices[0] = ctx.m.bodies[i].motion_type;
        b.flags_indices[1] = 0u;
        b.flags_indices[2] = i;
    }
    for (uint32_t b = nlinks; b < hd.body_count; ++b) {
        const MRBodyC& bp = ctx.m.bodies[b];
        MRBodyStateC s = st.scene[b - nlinks];
        Q4 rq;
        if (!normalized_quat(q4(s.orientation), rq) || !finite3(f3(s.position)) ||
            !finite3(f3(s.linear_velocity_inv_mass)) || !finite3(f3(s.angular_velocity))) {
            f.code = MR_STEP_NONFINITE_INPUT;
            f.index = b;
            return false;
        }
        s.position.w = 1.0f;
        s.orientation = mr_float4{rq.x, rq.y, rq.z, rq.w};
        s.linear_velocity_inv_mass.w = bp.motion_type == MR_MOTION_DYNAMIC ? bp.mass.y : 0.0f;
        s.angular_velocity.w = 0.0f;
        s.flags_indices[0] = bp.motion_type;
        s.flags_indices[1] = MR_INVALID_INDEX;
        s.flags_indices[2] = b;
        world_inverse_inertia(bp, rq, s);
        if (bp.motion_type == MR_MOTION_DYNAMIC) {
            const float dl = damp ? bp.damping_scale.x : 1.0f;
            const float da = damp ? bp.damping_scale.y : 1.0f;
            const V3 lv = add(mul(dl, f3(s.linear_velocity_inv_mass)), mul(h, ctx.m.gravity));
            const V3 av = mul(da, f3(s.angular_velocity));
            s.linear_velocity_inv_mass.x = lv.x;
            s.linear_velocity_inv_mass.y = lv.y;
            s.linear_velocity_inv_mass.z = lv.z;
            s.angular_velocity.x = av.x;
            s.angular_velocity.y = av.y;
            s.angular_velocity.z = av.z;
        } else {
            s.linear_velocity_inv_mass.x = s.linear_velocity_inv_mass.y = s.linear_velocity_inv_mass.z = 0.0f;
            s.angular_velocity.x = s.angular_velocity.y = s.angular_velocity.z = 0.0f;
        }
        wk.bodies[b] = s;
    }
    // 4. collider projection
    for (uint32_t s = 0; s < hd.shape_count; ++s) {
        if (!project_collider(ctx, wk, s, wk.col[s])) {
            f.code = MR_STEP_NONFINITE_INPUT;
            f.index = s;
            return false;
        }
    }
    // 5. broadphase over the compiled pairs, canonical order
    std::vector<uint32_t> active;
    for (uint32_t p = 0; p < hd.pair_count; ++p) {
        const MRCompiledPairC& pr = ctx.m.pairs[p];
        if (pair_may_overlap(wk.col[pr.collider_a], wk.col[pr.collider_b])) active.push_back(p);
    }
    const uint32_t nact = (uint32_t)active.size();
    cs.required_pairs = nact;
    cs.active_pairs = nact;
    if (cs.max_active_pairs < nact) cs.max_active_pairs = nact;
    if (capture && out.pairs != nullptr) {
        const uint32_t nrec = nact < ctx.pair_capacity ? nact : ctx.pair_capacity;
        for (uint32_t k = 0; k < nrec; ++k) {
            const MRCompiledPairC& pr = ctx.m.pairs[active[k]];
            out.pairs[k] = MRPairRecordC{ctx.env, pr.collider_a, pr.collider_b, pr.pair_class};
        }
        *out.pair_count = nact;
    }
    if (nact > ctx.pair_capacity) {
        f.code = MR_STEP_PAIR_CAPACITY_OVERFLOW;
        f.index = active[ctx.pair_capacity];
        return false;
    }
    // 6+7. narrowphase and persistent manifolds in canonical pair order
    const uint32_t old_count = *st.mcount;
    std::vector<MRManifoldHeaderC> new_h;
    std::vector<MRManifoldPointC> new_p;
    uint32_t raw_total = 0, retained = 0, fresh = 0;
    float max_pen = 0.0f;
    for (uint32_t k = 0; k < nact; ++k) {
        const MRCompiledPairC& pr = ctx.m.pairs[active[k]];
        const Collider& A = wk.col[pr.collider_a];
        const Collider& B = wk.col[pr.collider_b];
        RawBatch raw;
        generate_contacts(ctx, pr.collider_a, pr.collider_b, pr.pair_class, A, B, raw);
        for (uint32_t c = 0; c < raw.count; ++c) {
            if (!finite3(raw.c[c].normal) || !std::isfinite(raw.c[c].separation) ||
                !finite3(raw.c[c].point_a) || !finite3(raw.c[c].point_b)) {
                f.code = MR_STEP_NONFINITE_RESULT;
                f.index = active[k];
                return false;
            }
        }
        raw_total += raw.count;
        if (raw.count == 0u) continue;
        const float rest = ctx.m.shapes[pr.collider_a].contact_rest_radius.y +
            ctx.m.shapes[pr.collider_b].contact_rest_radius.y;
        for (uint32_t c = 0; c < raw.count; ++c) max_pen = fmaxd(max_pen, fmaxd(rest - raw.c[c].separation, 0.0f));
        const MRManifoldHeaderC* oh = nullptr;
        const MRManifoldPointC* op = nullptr;
        for (uint32_t m = 0; m < old_count; ++m) {
            if (st.mh[m].pair_and_count[1] == pr.collider_a && st.mh[m].pair_and_count[2] == pr.collider_b) {
                oh = &st.mh[m];
                op = &st.mp[m * MR_MANIFOLD_POINT_CAPACITY];
                break;
            }
        }
        MRManifoldHeaderC mh;
        MRManifoldPointC mp[MR_MANIFOLD_POINT_CAPACITY];
        if (!build_manifold(ctx, wk, ctx.env, pr, raw, oh, op, mh, mp, retained, fresh)) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = active[k];
            return false;
        }
        new_h.push_back(mh);
        for (uint32_t i = 0; i < MR_MANIFOLD_POINT_CAPACITY; ++i) new_p.push_back(mp[i]);
    }
    const uint32_t nman = (uint32_t)new_h.size();
    cs.raw_contacts = raw_total;
    cs.required_manifolds = nman;
    cs.manifolds = nman;
    cs.retained_points = retained;
    cs.new_points = fresh;
    cs.diagnostics.x = max_pen;
    if (nman > ctx.manifold_capacity) {
        f.code = MR_STEP_MANIFOLD_CAPACITY_OVERFLOW;
        f.index = nman;
        return false;
    }
    uint32_t ncons = 0;
    for (uint32_t m = 0; m < nman; ++m) ncons += new_h[m].pair_and_count[3];
    cs.required_constraints = ncons;
    if (ncons > ctx.constraint_capacity) {
        f.code = MR_STEP_CONSTRAINT_CAPACITY_OVERFLOW;
        f.index = ncons;
        return false;
    }
    cs.constraints = ncons;
    if (cs.max_constraints < ncons) cs.max_constraints = ncons;
    // 8. constraint emission (Collision.metal:9180-9470)
    cons.resize(ncons);
    uint32_t ci = 0;
    for (uint32_t m = 0; m < nman; ++m) {
        const MRManifoldHeaderC& mh = new_h[m];
        const MRShapeC& sa = ctx.m.shapes[mh.pair_and_count[1]];
        const MRShapeC& sb = ctx.m.shapes[mh.pair_and_count[2]];
        const MRBodyStateC& ba = wk.bodies[sa.body];
        const MRBodyStateC& bb = wk.bodies[sb.body];
        Q4 ra, rb;
        checked_quat(q4(ba.orientation), ra);
        checked_quat(q4(bb.orientation), rb);
        const V3 n = normalize(qrot(ra, f3(mh.normal_and_age)));
        V3 t = qrot(ra, f3(mh.tangent_and_metric));
        t = sub(t, mul(dot(t, n), n));
        t = dot(t, t) > 1.0e-12f ? normalize(t) : stable_tangent(n);
        const V3 bt = cross(n, t);
        const MRMaterialC& ma = ctx.m.materials[sa.material];
        const MRMaterialC& mb = ctx.m.materials[sb.material];
        for (uint32_t p = 0; p < mh.pair_and_count[3]; ++p) {
            const MRManifoldPointC& mp = new_p[m * MR_MANIFOLD_POINT_CAPACITY + p];
            Constraint& c = cons[ci++];
            c.body_a = sa.body;
            c.body_b = sb.body;
            c.flags = mp.feature_and_life[2] == 0u ? MR_CONSTRAINT_FLAG_NEW_IMPACT : MR_CONSTRAINT_FLAG_WARM_STARTED;
            c.manifold = m;
            c.point = p;
            c.pa = world_from_anchor(ba, ra, mp.local_anchor_a);
            c.pb = world_from_anchor(bb, rb, mp.local_anchor_b);
            c.point_world = mul(0.5f, add(c.pa, c.pb));
            const float rest = sa.contact_rest_radius.y + sb.contact_rest_radius.y;
            c.separation = dot(sub(c.pb, c.pa), n) - rest;
            c.dir[0] = n;
            c.dir[1] = t;
            c.dir[2] = bt;
            const float fa[4] = {ma.friction.x, ma.friction.y, ma.friction.z, ma.friction.w};
            const float fb[4] = {mb.friction.x, mb.friction.y, mb.friction.z, mb.friction.w};
            for (int k = 0; k < 4; ++k) c.friction[k] = std::sqrt(fmaxd(0.0f, fa[k] * fb[k]));
            c.restitution = fmaxd(ma.response.x, mb.response.x);
            c.restitution_threshold = fmaxd(ma.response.y, mb.response.y);
            c.compliance = ma.response.z + mb.response.z;
            c.dissipation = ma.response.w + mb.response.w;
            if (ctx.cfg->warm_start != 0u && ctx.cfg->ccd_mode == MR_CCD_HYBRID && mp.impulses.w > 0.0f) {
                // impulses scale with the step: a segment warm-starts from the last
                // segment's impulses times the ratio of the two durations
                const float s = h / mp.impulses.w;
                c.lambda[0] = mp.impulses.x * s;
                c.lambda[1] = mp.impulses.y * s;
                c.lambda[2] = mp.impulses.z * s;
            } else if (ctx.cfg->warm_start != 0u) {
                c.lambda[0] = mp.impulses.x;
                c.lambda[1] = mp.impulses.y;
                c.lambda[2] = mp.impulses.z;
            } else {
                c.lambda[0] = c.lambda[1] = c.lambda[2] = 0.0f;
            }
        }
    }
    // 9. Jacobian rows and inverse-ABA response columns per constraint row
    jrow.assign((size_t)ncons * 3u * nv, 0.0f);
    resp.assign((size_t)ncons * 3u * nv, 0.0f);
    for (uint32_t c = 0; c < ncons; ++c) {
        Constraint& k = cons[c];
        const bool la = is_link(ctx, k.body_a), lb = is_link(ctx, k.body_b);
        if (!la && !lb) continue;
        for (int r = 0; r < 3; ++r) {
            float* jr = &jrow[((size_t)c * 3u + (size_t)r) * nv];
            for (uint32_t j = 0; j < nv; ++j) {
                const V3 cb = lb ? point_jacobian_column(ctx, wk, k.body_b, k.pb, j) : V3{0.0f, 0.0f, 0.0f};
                const V3 ca = la ? point_jacobian_column(ctx, wk, k.body_a, k.pa, j) : V3{0.0f, 0.0f, 0.0f};
                jr[j] = dot(k.dir[r], sub(cb, ca));
            }
            PointForce forces[2];
            uint32_t nf = 0;
            if (la) forces[nf++] = PointForce{k.body_a, k.pa, neg(k.dir[r])};
            if (lb) forces[nf++] = PointForce{k.body_b, k.pb, k.dir[r]};
            inverse_aba(ctx, wk, forces, nf, &resp[((size_t)c * 3u + (size_t)r) * nv]);
        }
    }
    // 10. IR evaluation (MetalWorldContact.metal:3800-3995) and block inverse
    auto relative_rows = [&](const Constraint& k, uint32_t c, const float* vel, float rel[3]) {
        const bool sa_dyn = !is_link(ctx, k.body_a) && is_dynamic_scene(ctx, k.body_a);
        const bool sb_dyn = !is_link(ctx, k.body_b) && is_dynamic_scene(ctx, k.body_b);
        const V3 va = sa_dyn ? scene_point_velocity(wk.bodies[k.body_a], k.point_world) : V3{0.0f, 0.0f, 0.0f};
        const V3 vb = sb_dyn ? scene_point_velocity(wk.bodies[k.body_b], k.point_world) : V3{0.0f, 0.0f, 0.0f};
        const bool art = is_link(ctx, k.body_a) || is_link(ctx, k.body_b);
        for (int r = 0; r < 3; ++r) {
            // articulation part, plus the point velocity of a dynamic body B,
            // minus that of a dynamic body A (other endpoints add nothing)
            float a = art ? fold_dot(&jrow[((size_t)c * 3u + (size_t)r) * nv], vel, nv) : 0.0f;
            if (sb_dyn) a = a + dot(k.dir[r], vb);
            if (sa_dyn) a = a - dot(k.dir[r], va);
            rel[r] = a;
        }
    };
    const float slop = ctx.cfg->contact_slop;
    const float maxdep = ctx.cfg->max_depenetration_velocity;
    for (uint32_t c = 0; c < ncons; ++c) {
        Constraint& k = cons[c];
        float rel[3];
        relative_rows(k, c, wk.vcand, rel);
        const float slip = std::sqrt(rel[1] * rel[1] + rel[2] * rel[2]);
        const bool stat = slip <= 1.0e-3f;
        k.mu_u = stat ? k.friction[0] : k.friction[2];
        k.mu_v = stat ? k.friction[1] : k.friction[3];
        const float src_target = k.separation > 0.0f ? -k.separation / h : 0.0f;
        const float tau = 2.0f * h;
        const float err = fmind(k.separation + slop, 0.0f);
        const float ratio = h / tau;
        const float den = 1.0f + 2.0f * 1.0f * ratio + ratio * ratio;
        float stab = -h * err / (tau * tau * den);
        stab = clampd(fmaxd(stab, 0.0f), 0.0f, maxdep);
        if ((k.flags & MR_CONSTRAINT_FLAG_NEW_IMPACT) != 0u) {
            const float incoming = rel[0] - src_target;
            if (incoming < -k.restitution_threshold) stab = fmaxd(stab, -k.restitution * incoming);
        }
        k.target[0] = src_target + stab;
        k.target[1] = 0.0f;
        k.target[2] = 0.0f;
        k.reg[0] = fmaxd(k.compliance / (h * h) + k.dissipation / h, 0.0f);
        k.reg[1] = 0.0f;
        k.reg[2] = 0.0f;
        if (!std::isfinite(k.target[0]) || !std::isfinite(k.reg[0])) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = c;
            return false;
        }
        // Delassus diagonal block W = J M^-1 J^T + scene responses + R
        const bool art = is_link(ctx, k.body_a) || is_link(ctx, k.body_b);
        const bool sa_dyn = is_dynamic_scene(ctx, k.body_a);
        const bool sb_dyn = is_dynamic_scene(ctx, k.body_b);
        float W[3][3];
        for (int r = 0; r < 3; ++r)
            for (int s = 0; s < 3; ++s) {
                float val = art ? fold_dot(&jrow[((size_t)c * 3u + (size_t)r) * nv],
                                           &resp[((size_t)c * 3u + (size_t)s) * nv], nv) : 0.0f;
                if (sa_dyn) val += dot(k.dir[r], scene_response(wk.bodies[k.body_a], k.point_world, k.dir[s]));
                if (sb_dyn) val += dot(k.dir[r], scene_response(wk.bodies[k.body_b], k.point_world, k.dir[s]));
                if (r == s) val += k.reg[r];
                W[r][s] = val;
            }
        if (!invert3x3(W, k.winv)) {
            f.code = MR_STEP_FACTORIZATION_FAILED;
            f.index = c;
            return false;
        }
    }
    // problem capture: header, J rows, response rows, free velocity,
    // targets, regularisation, cone coefficients, warm-start impulses
    const bool capture_problem = final_sub && g_capture.out != nullptr && ctx.env == g_capture.env;
    if (capture_problem) {
        const uint32_t nscene_c = hd.body_count - nlinks;
        const size_t need = 8u + 2u * (size_t)ncons * 3u * nv + 2u * nv + (size_t)ncons * (3u * 4u + 2u + 16u) +
                            25u * (size_t)nscene_c;
        g_capture.used = 0;
        if (need <= g_capture.capacity) {
            float* o = g_capture.out;
            bool dynamic_scene = false;
            for (uint32_t c = 0; c < ncons; ++c)
                dynamic_scene = dynamic_scene || is_dynamic_scene(ctx, cons[c].body_a) || is_dynamic_scene(ctx, cons[c].body_b);
            const uint32_t hdr[8] = {nv, ncons, dynamic_scene ? 1u : 0u, nscene_c, 0u, 0u, 0u, 0u};
            std::memcpy(o, hdr, sizeof(hdr));
            size_t at = 8;
            for (uint32_t c = 0; c < ncons; ++c) {
                const bool art = is_link(ctx, cons[c].body_a) || is_link(ctx, cons[c].body_b);
                for (uint32_t r = 0; r < 3u; ++r)
                    for (uint32_t j = 0; j < nv; ++j) o[at++] = art ? jrow[((size_t)c * 3u + r) * nv + j] : 0.0f;
            }
            for (uint32_t c = 0; c < ncons; ++c) {
                const bool art = is_link(ctx, cons[c].body_a) || is_link(ctx, cons[c].body_b);
                for (uint32_t r = 0; r < 3u; ++r)
                    for (uint32_t j = 0; j < nv; ++j) o[at++] = art ? resp[((size_t)c * 3u + r) * nv + j] : 0.0f;
            }
            for (uint32_t j = 0; j < nv; ++j) o[at++] = wk.vcand[j];
            for (uint32_t c = 0; c < ncons; ++c)
                for (int r = 0; r < 3; ++r) o[at++] = cons[c].target[r];
            for (uint32_t c = 0; c < ncons; ++c)
                for (int r = 0; r < 3; ++r) o[at++] = cons[c].reg[r];
            for (uint32_t c = 0; c < ncons; ++c) {
                o[at++] = cons[c].mu_u;
                o[at++] = cons[c].mu_v;
            }
            for (uint32_t c = 0; c < ncons; ++c)
                for (int r = 0; r < 3; ++r) o[at++] = cons[c].lambda[r];
            // contact geometry: dynamic flags, scene indices, point, directions
            for (uint32_t c = 0; c < ncons; ++c) {
                const Constraint& k = cons[c];
                const bool da = is_dynamic_scene(ctx, k.body_a), db = is_dynamic_scene(ctx, k.body_b);
                o[at++] = da ? 1.0f : 0.0f;
                o[at++] = db ? 1.0f : 0.0f;
                o[at++] = da ? (float)(k.body_a - nlinks) : -1.0f;
                o[at++] = db ? (float)(k.body_b - nlinks) : -1.0f;
                o[at++] = k.point_world.x;
                o[at++] = k.point_world.y;
                o[at++] = k.point_world.z;
                for (int r = 0; r < 3; ++r) {
                    o[at++] = k.dir[r].x;
                    o[at++] = k.dir[r].y;
                    o[at++] = k.dir[r].z;
                }
            }
            // scene bodies: position, inverse mass, world inverse inertia, twist
            for (uint32_t b = 0; b < nscene_c; ++b) {
                const MRBodyStateC& sb = wk.bodies[nlinks + b];
                o[at++] = sb.position.x;
                o[at++] = sb.position.y;
                o[at++] = sb.position.z;
                o[at++] = is_dynamic_scene(ctx, nlinks + b) ? sb.linear_velocity_inv_mass.w : 0.0f;
                const mr_float4 rows[3] = {sb.inverse_inertia_world_row0, sb.inverse_inertia_world_row1,
                                           sb.inverse_inertia_world_row2};
                for (int r = 0; r < 3; ++r) {
                    o[at++] = rows[r].x;
                    o[at++] = rows[r].y;
                    o[at++] = rows[r].z;
                }
                o[at++] = sb.linear_velocity_inv_mass.x;
                o[at++] = sb.linear_velocity_inv_mass.y;
                o[at++] = sb.linear_velocity_inv_mass.z;
                o[at++] = sb.angular_velocity.x;
                o[at++] = sb.angular_velocity.y;
                o[at++] = sb.angular_velocity.z;
            }
            g_capture.used = (uint32_t)at;
        }
    }
    // 11. warm start, then ordered block Gauss-Seidel sweeps
    auto apply_delta = [&](const Constraint& k, uint32_t c, const float d[3]) {
        const bool art = is_link(ctx, k.body_a) || is_link(ctx, k.body_b);
        if (art) {
            const float* r0 = &resp[((size_t)c * 3u + 0u) * nv];
            const float* r1 = &resp[((size_t)c * 3u + 1u) * nv];
            const float* r2 = &resp[((size_t)c * 3u + 2u) * nv];
            for (uint32_t j = 0; j < nv; ++j) wk.vcand[j] += (r0[j] * d[0] + r1[j] * d[1]) + r2[j] * d[2];
        }
        const V3 imp = add(add(mul(d[0], k.dir[0]), mul(d[1], k.dir[1])), mul(d[2], k.dir[2]));
        if (is_dynamic_scene(ctx, k.body_a)) apply_scene_impulse(wk.bodies[k.body_a], k.point_world, neg(imp));
        if (is_dynamic_scene(ctx, k.body_b)) apply_scene_impulse(wk.bodies[k.body_b], k.point_world, imp);
    };
    const uint32_t base_iters = ctx.cfg->velocity_iterations + (final_sub ? ctx.cfg->final_velocity_iterations : 0u);
    const bool adaptive = ctx.cfg->residual_tolerance > 0.0f;
    const uint32_t max_iters = adaptive ? ctx.cfg->max_iterations : base_iters;
    float sweep_delta = 0.0f;
    uint32_t iters = 0;
    const bool jacobi = ctx.cfg->solver == MR_SOLVER_RELAXED_JACOBI;
    if (jacobi) {
        // Relaxed block Jacobi (mrWorldWave32SolvePacket, MetalWorldContact.metal:6142-7213).
        // Articulated dot products here are serial in DoF order; sums over
        // contacts run in constraint order (contacts of other islands add
        // exact zeros, so these are the island sums).
        auto is_art = [&](const Constraint& k) { return is_link(ctx, k.body_a) || is_link(ctx, k.body_b); };
        auto serial_dot = [&](const float* x, const float* y) {
            float acc = 0.0f;
            for (uint32_t j = 0; j < nv; ++j) acc = acc + x[j] * y[j];
            return acc;
        };
        // W_nn(c, j): normal velocity at contact c per unit normal impulse of contact j
        auto normal_cross = [&](uint32_t c, uint32_t j) {
            const Constraint& kc = cons[c];
            const Constraint& kj = cons[j];
            float w = (is_art(kc) && is_art(kj)) ? serial_dot(&jrow[(size_t)c * 3u * nv], &resp[(size_t)j * 3u * nv])
                                                 : 0.0f;
            const uint32_t cb[2] = {kc.body_a, kc.body_b}, jb[2] = {kj.body_a, kj.body_b};
            for (int x = 0; x < 2; ++x) {
                if (!is_dynamic_scene(ctx, cb[x])) continue;
                for (int y = 0; y < 2; ++y) {
                    if (jb[y] != cb[x]) continue;
                    const float t = dot(kc.dir[0], scene_cross_response(wk.bodies[cb[x]], kc.point_world,
                                                                        kj.point_world, kj.dir[0]));
                    w = (x == y) ? w + t : w - t;
                }
            }
            return w;
        };
        // relaxation = diag / sum_j |W_nn(c, j)| clamped to [1/32, 1] (:6722-6798)
        std::vector<float> relax(ncons, 1.0f);
        for (uint32_t c = 0; c < ncons; ++c) {
            float sum = 0.0f, diag = 0.0f;
            for (uint32_t j = 0; j < ncons; ++j) {
                const float w = std::fabs(normal_cross(c, j));
                sum = sum + w;
                if (j == c) diag = w + cons[c].reg[0];
            }
            relax[c] = sum > fmaxd(diag, kMatrixFloor) ? clampd(diag / sum, 1.0f / 32.0f, 1.0f) : 1.0f;
        }
        // all deltas at once: DoF sums in constraint order, applied to a DoF that any
        // response touches; scene bodies accumulate linear and angular impulse (:6858-6990)
        std::vector<float> delta((size_t)ncons * 3u, 0.0f);
        auto apply_all = [&]() {
            for (uint32_t j = 0; j < nv; ++j) {
                float vd = 0.0f;
                bool part = false;
                for (uint32_t c = 0; c < ncons; ++c) {
                    if (!is_art(cons[c])) continue;
                    const float r0 = resp[((size_t)c * 3u + 0u) * nv + j];
                    const float r1 = resp[((size_t)c * 3u + 1u) * nv + j];
                    const float r2 = resp[((size_t)c * 3u + 2u) * nv + j];
                    part = part || r0 != 0.0f || r1 != 0.0f || r2 != 0.0f;
                    vd = vd + ((r0 * delta[3u * c] + r1 * delta[3u * c + 1u]) + r2 * delta[3u * c + 2u]);
                }
                if (part) wk.vcand[j] = wk.vcand[j] + vd;
            }
            for (uint32_t b = nlinks; b < hd.body_count; ++b) {
                if (!is_dynamic_scene(ctx, b)) continue;
                MRBodyStateC& st = wk.bodies[b];
                V3 lin{0.0f, 0.0f, 0.0f}, ang{0.0f, 0.0f, 0.0f};
                bool touched = false;
                for (uint32_t c = 0; c < ncons; ++c) {
                    const Constraint& k = cons[c];
                    if (k.body_a != b && k.body_b != b) continue;
                    touched = true;
                    const float* d = &delta[3u * c];
                    const V3 imp = add(add(mul(d[0], k.dir[0]), mul(d[1], k.dir[1])), mul(d[2], k.dir[2]));
                    const V3 lever = sub(k.point_world, f3(st.position));
                    if (k.body_a == b) {
                        lin = sub(lin, imp);
                        ang = add(ang, cross(lever, neg(imp)));
                    }
                    if (k.body_b == b) {
                        lin = add(lin, imp);
                        ang = add(ang, cross(lever, imp));
                    }
                }
                if (!touched) continue;
                const V3 dv = mul(st.linear_velocity_inv_mass.w, lin);
                st.linear_velocity_inv_mass.x += dv.x;
                st.linear_velocity_inv_mass.y += dv.y;
                st.linear_velocity_inv_mass.z += dv.z;
                const V3 dw = inv_inertia_mul(st, ang);
                st.angular_velocity.x += dw.x;
                st.angular_velocity.y += dw.y;
                st.angular_velocity.z += dw.z;
            }
        };
        // warm start: project, then apply every warm impulse once (:6402-6470)
        for (uint32_t c = 0; c < ncons; ++c) {
            Constraint& k = cons[c];
            project_cone(k.lambda, k.mu_u, k.mu_v, ctx.cfg->maximum_normal_impulse);
            for (int r = 0; r < 3; ++r) delta[3u * c + r] = k.lambda[r];
        }
        apply_all();
        float largest = 0.0f;
        for (uint32_t it = 0; it < max_iters; ++it) {
            float it_delta = 0.0f;
            for (uint32_t c = 0; c < ncons; ++c) {
                Constraint& k = cons[c];
                float rel[3] = {0.0f, 0.0f, 0.0f};
                const bool sa_dyn = is_dynamic_scene(ctx, k.body_a), sb_dyn = is_dynamic_scene(ctx, k.body_b);
                const V3 va = sa_dyn ? scene_point_velocity(wk.bodies[k.body_a], k.point_world) : V3{0.0f, 0.0f, 0.0f};
                const V3 vb = sb_dyn ? scene_point_velocity(wk.bodies[k.body_b], k.point_world) : V3{0.0f, 0.0f, 0.0f};
                for (int r = 0; r < 3; ++r) {
                    float a = is_art(k) ? serial_dot(&jrow[((size_t)c * 3u + (size_t)r) * nv], wk.vcand) : 0.0f;
                    if (sb_dyn) a = a + dot(k.dir[r], vb);
                    if (sa_dyn) a = a - dot(k.dir[r], va);
                    rel[r] = a;
                }
                float rhs[3];
                for (int r = 0; r < 3; ++r) rhs[r] = (k.target[r] - rel[r]) - k.reg[r] * k.lambda[r];
                float prop[3] = {k.lambda[0], k.lambda[1], k.lambda[2]};
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) prop[r] += k.winv[r][q] * rhs[q];
                project_cone(prop, k.mu_u, k.mu_v, ctx.cfg->maximum_normal_impulse);
                float d[3], cand[3];
                for (int r = 0; r < 3; ++r) {
                    d[r] = relax[c] * (prop[r] - k.lambda[r]);
                    cand[r] = k.lambda[r] + d[r];
                }
                if (!std::isfinite(cand[0]) || !std::isfinite(cand[1]) || !std::isfinite(cand[2]) ||
                    !std::isfinite(d[0]) || !std::isfinite(d[1]) || !std::isfinite(d[2])) {
                    f.code = MR_STEP_NONFINITE_RESULT;
                    f.index = c;
                    return false;
                }
                for (int r = 0; r < 3; ++r) {
                    k.lambda[r] = cand[r];
                    delta[3u * c + r] = d[r];
                }
                it_delta = fmaxd(it_delta, fmaxd(std::fabs(d[0]), fmaxd(std::fabs(d[1]), std::fabs(d[2]))));
            }
            apply_all();
            largest = fmaxd(largest, it_delta);
            ++iters;
            if (adaptive && it_delta <= ctx.cfg->residual_tolerance) break;
        }
        sweep_delta = largest;
    } else {
    for (uint32_t c = 0; c < ncons; ++c) {
        Constraint& k = cons[c];
        project_cone(k.lambda, k.mu_u, k.mu_v, ctx.cfg->maximum_normal_impulse);
        apply_delta(k, c, k.lambda);
    }
    for (uint32_t it = 0; it < max_iters; ++it) {
        sweep_delta = 0.0f;
        for (uint32_t c = 0; c < ncons; ++c) {
            Constraint& k = cons[c];
            float rel[3];
            relative_rows(k, c, wk.vcand, rel);
            float rhs[3];
            for (int r = 0; r < 3; ++r) rhs[r] = (k.target[r] - rel[r]) - k.reg[r] * k.lambda[r];
            float cand[3] = {k.lambda[0], k.lambda[1], k.lambda[2]};
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) cand[r] += k.winv[r][s] * rhs[s];
            project_cone(cand, k.mu_u, k.mu_v, ctx.cfg->maximum_normal_impulse);
            const float d[3] = {cand[0] - k.lambda[0], cand[1] - k.lambda[1], cand[2] - k.lambda[2]};
            if (!std::isfinite(cand[0]) || !std::isfinite(cand[1]) || !std::isfinite(cand[2]) ||
                !std::isfinite(d[0]) || !std::isfinite(d[1]) || !std::isfinite(d[2])) {
                f.code = MR_STEP_NONFINITE_RESULT;
                f.index = c;
                return false;
            }
            sweep_delta = fmaxd(sweep_delta, fmaxd(std::fabs(d[0]), fmaxd(std::fabs(d[1]), std::fabs(d[2]))));
            k.lambda[0] = cand[0];
            k.lambda[1] = cand[1];
            k.lambda[2] = cand[2];
            apply_delta(k, c, d);
        }
        ++iters;
        if (adaptive && sweep_delta <= ctx.cfg->residual_tolerance) break;
    }
    }
    if (capture_problem && g_capture.used != 0u) {
        float* o = g_capture.out;
        size_t at = g_capture.used;
        for (uint32_t c = 0; c < ncons; ++c)
            for (int r = 0; r < 3; ++r) o[at++] = cons[c].lambda[r];
        for (uint32_t j = 0; j < nv; ++j) o[at++] = wk.vcand[j];
        for (uint32_t b = nlinks; b < hd.body_count; ++b) {
            const MRBodyStateC& sb = wk.bodies[b];
            o[at++] = sb.linear_velocity_inv_mass.x;
            o[at++] = sb.linear_velocity_inv_mass.y;
            o[at++] = sb.linear_velocity_inv_mass.z;
            o[at++] = sb.angular_velocity.x;
            o[at++] = sb.angular_velocity.y;
            o[at++] = sb.angular_velocity.z;
        }
        g_capture.used = (uint32_t)at;
    }
    // 12. residuals (MetalWorldContact.metal:8220-8315) and impulse write-back
    float max_normal = 0.0f, max_cone = 0.0f;
    for (uint32_t c = 0; c < ncons; ++c) {
        Constraint& k = cons[c];
        float rel[3];
        relative_rows(k, c, wk.vcand, rel);
        const float eq = (rel[0] - k.target[0]) + k.reg[0] * k.lambda[0];
        const float nres = k.lambda[0] > kConeEpsilon ? std::fabs(eq) : fmaxd(-eq, 0.0f);
        max_normal = fmaxd(max_normal, nres);
        const float lu = k.mu_u * k.lambda[0], lv = k.mu_v * k.lambda[0];
        float viol;
        if (lu > 0.0f && lv > 0.0f)
            viol = fmaxd(std::sqrt((k.lambda[1] * k.lambda[1]) / (lu * lu) + (k.lambda[2] * k.lambda[2]) / (lv * lv)) - 1.0f, 0.0f);
        else
            viol = std::sqrt(k.lambda[1] * k.lambda[1] + k.lambda[2] * k.lambda[2]);
        max_cone = fmaxd(max_cone, viol);
        MRManifoldPointC& mp = new_p[k.manifold * MR_MANIFOLD_POINT_CAPACITY + k.point];
        mp.impulses = mr_float4{k.lambda[0], k.lambda[1], k.lambda[2], ctx.cfg->ccd_mode == MR_CCD_HYBRID ? h : 0.0f};
    }
    if (!std::isfinite(sweep_delta) || !std::isfinite(max_normal) || !std::isfinite(max_cone)) {
        f.code = MR_STEP_NONFINITE_RESULT;
        f.index = MR_INVALID_INDEX;
        return false;
    }
    cs.solver_iterations = ncons > 0u ? iters : 0u;
    cs.residuals = mr_float4{ncons > 0u ? sweep_delta : 0.0f, max_normal, max_cone,
                             cs.residuals.w + (float)(ncons > 0u ? iters : 0u)};
    // islands over dynamic nodes: node 0 = articulation, 1 + s = scene body s
    {
        uint32_t parent[MR_MAX_BODIES + 1];
        bool used[MR_MAX_BODIES + 1] = {};
        const uint32_t nodes = 1u + (hd.body_count - nlinks);
        for (uint32_t i = 0; i < nodes; ++i) parent[i] = i;
        auto find = [&](uint32_t x) {
            while (parent[x] != x) x = parent[x];
            return x;
        };
        auto node_of = [&](uint32_t body) -> uint32_t {
            if (is_link(ctx, body)) return 0u;
            if (is_dynamic_scene(ctx, body)) return 1u + (body - nlinks);
            return MR_INVALID_INDEX;
        };
        for (uint32_t c = 0; c < ncons; ++c) {
            const uint32_t a = node_of(cons[c].body_a), b = node_of(cons[c].body_b);
            if (a != MR_INVALID_INDEX) used[a] = true;
            if (b != MR_INVALID_INDEX) used[b] = true;
            if (a != MR_INVALID_INDEX && b != MR_INVALID_INDEX) {
                const uint32_t ra = find(a), rb = find(b);
                if (ra != rb) {
                    if (ra < rb) parent[rb] = ra;
                    else parent[ra] = rb;
                }
            }
        }
        uint32_t islands = 0;
        for (uint32_t i = 0; i < nodes; ++i)
            if (used[i] && find(i) == i) ++islands;
        cs.islands = islands;
    }
    // 13. joint limits (MetalWorldContact.metal:2878-2960)
    for (uint32_t j = 0; j < nv; ++j) {
        const MRDofC& dof = ctx.m.dofs[j];
        float vel = wk.vcand[j];
        if ((dof.flags & MR_DOF_FLAG_VELOCITY_LIMIT) != 0u) vel = clampd(vel, -dof.limits.z, dof.limits.z);
        if ((dof.flags & MR_DOF_FLAG_POSITION_LIMIT) != 0u && dof.q_index != MR_INVALID_INDEX) {
            const float pos = st.q[dof.q_index];
            vel = clampd(vel, (dof.limits.x - pos) / h, (dof.limits.y - pos) / h);
        }
        if (!std::isfinite(vel)) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = j;
            return false;
        }
        wk.vcand[j] = vel;
    }
    // 14. integration (MetalWorldContact.metal:8876-9043)
    float qn[MR_MAX_Q];
    for (uint32_t i = 0; i < hd.nq; ++i) qn[i] = st.q[i];
    if (hd.root_type == MR_ROOT_FLOATING) {
        qn[0] = st.q[0] + h * wk.vcand[0];
        qn[1] = st.q[1] + h * wk.vcand[1];
        qn[2] = st.q[2] + h * wk.vcand[2];
        const Q4 inc = rotation_increment(V3{h * wk.vcand[3], h * wk.vcand[4], h * wk.vcand[5]});
        Q4 o;
        if (!normalized_quat(qmul(inc, Q4{st.q[3], st.q[4], st.q[5], st.q[6]}), o)) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = 3u;
            return false;
        }
        qn[3] = o.x; qn[4] = o.y; qn[5] = o.z; qn[6] = o.w;
    }
    for (uint32_t i = 1; i < nlinks; ++i) {
        const MRBodyC& b = ctx.m.bodies[i];
        if (b.q_index != MR_INVALID_INDEX) qn[b.q_index] = st.q[b.q_index] + h * wk.vcand[b.v_index];
    }
    for (uint32_t i = 0; i < hd.nq; ++i) {
        if (!std::isfinite(qn[i])) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = i;
            return false;
        }
    }
    for (uint32_t b = nlinks; b < hd.body_count; ++b) {
        const MRBodyC& bp = ctx.m.bodies[b];
        MRBodyStateC& s = wk.bodies[b];
        if (bp.motion_type != MR_MOTION_DYNAMIC) continue;
        V3 lin = f3(s.linear_velocity_inv_mass);
        V3 ang = f3(s.angular_velocity);
        const float ll = length(lin), al = length(ang);
        if (bp.damping.z > 0.0f && ll > bp.damping.z) lin = mul(bp.damping.z / ll, lin);
        if (bp.damping.w > 0.0f && al > bp.damping.w) ang = mul(bp.damping.w / al, ang);
        s.linear_velocity_inv_mass.x = lin.x;
        s.linear_velocity_inv_mass.y = lin.y;
        s.linear_velocity_inv_mass.z = lin.z;
        s.angular_velocity.x = ang.x;
        s.angular_velocity.y = ang.y;
        s.angular_velocity.z = ang.z;
        const V3 np = add(f3(s.position), mul(h, lin));
        s.position = mr_float4{np.x, np.y, np.z, 1.0f};
        const Q4 inc = rotation_increment(mul(h, ang));
        Q4 o;
        if (!normalized_quat(qmul(inc, q4(s.orientation)), o) || !finite3(np)) {
            f.code = MR_STEP_NONFINITE_RESULT;
            f.index = b;
            return false;
        }
        s.orientation = mr_float4{o.x, o.y, o.z, o.w};
        world_inverse_inertia(bp, o, s);
    }
    // 15. commit the microstep into the working state
    for (uint32_t i = 0; i < hd.nq; ++i) st.q[i] = qn[i];
    for (uint32_t j = 0; j < nv; ++j) st.v[j] = wk.vcand[j];
    for (uint32_t b = nlinks; b < hd.body_count; ++b) st.scene[b - nlinks] = wk.bodies[b];
    for (uint32_t m = 0; m < nman; ++m) {
        st.mh[m] = new_h[m];
        for (uint32_t p = 0; p < MR_MANIFOLD_POINT_CAPACITY; ++p)
            st.mp[m * MR_MANIFOLD_POINT_CAPACITY + p] = new_p[m * MR_MANIFOLD_POINT_CAPACITY + p];
    }
    *st.mcount = nman;
    return true;
}

} // namespace

extern "C" {

// One control step for environments [env_begin, env_end): reset, checkpoint,
// `physics_substeps` microsteps, transactional commit (MetalWorld.metal:113-519).
// State arrays are environment-major with the strides of the model descriptor
// (max_nq, max_nv, max_scene_bodies, manifold_capacity). `command` holds one
// control step: [environment][max_nv]. Evidence arrays may be null.
int mr_oracle_control_step(const MRWorldModelDescC* desc, const MRWorldStepConfigC* cfg,
                           uint32_t env_begin, uint32_t env_end, uint32_t control_step,
                           float* q, float* v, MRBodyStateC* scene, MRManifoldHeaderC* mh,
                           MRManifoldPointC* mp, uint32_t* mcount, const float* command,
                           const uint32_t* reset_mask, const float* reset_q, const float* reset_v,
                           const MRBodyStateC* reset_scene, MRWorldStatusC* status,
                           MRContactStatusC* cstatus, float* acceleration,
                           MRPairRecordC* evidence_pairs, uint32_t* evidence_counts) {
    if (desc == nullptr || cfg == nullptr || cfg->physics_substeps == 0u) return -1;
    const uint32_t max_nq = desc->max_nq, max_nv = desc->max_nv;
    const uint32_t max_sc = desc->max_scene_bodies;
    const uint32_t mcap = cfg->manifold_capacity;
    Work* wk = new Work();
    std::vector<Constraint> cons;
    std::vector<float> jrow, resp;
    for (uint32_t env = env_begin; env < env_end; ++env) {
        StepContext ctx;
        ctx.desc = desc;
        ctx.cfg = cfg;
        ctx.m = env_model(desc, env);
        ctx.env = env;
        ctx.h = cfg->control_timestep / (float)cfg->physics_substeps;
        ctx.manifold_capacity = mcap;
        ctx.pair_capacity = cfg->pair_capacity;
        ctx.constraint_capacity = cfg->constraint_capacity;
        const MRModelHeaderC& hd = *ctx.m.header;
        float* eq = q + (size_t)env * max_nq;
        float* ev = v + (size_t)env * max_nv;
        MRBodyStateC* es = scene + (size_t)env * max_sc;
        MRManifoldHeaderC* emh = mh + (size_t)env * mcap;
        MRManifoldPointC* emp = mp + (size_t)env * mcap * MR_MANIFOLD_POINT_CAPACITY;
        uint32_t* emc = mcount + env;
        if (reset_mask != nullptr && reset_mask[env] != 0u) {
            for (uint32_t i = 0; i < hd.nq; ++i) eq[i] = reset_q[(size_t)env * max_nq + i];
            for (uint32_t i = 0; i < hd.nv; ++i) ev[i] = reset_v[(size_t)env * max_nv + i];
            for (uint32_t s = 0; s < hd.scene_body_count; ++s) es[s] = reset_scene[(size_t)env * max_sc + s];
            *emc = 0u;
        }
        // working copies (the checkpoint is the untouched caller state)
        std::vector<float> wq(eq, eq + hd.nq), wv(ev, ev + hd.nv);
        std::vector<MRBodyStateC> ws(es, es + hd.scene_body_count);
        std::vector<MRManifoldHeaderC> wmh(emh, emh + mcap);
        std::vector<MRManifoldPointC> wmp(emp, emp + (size_t)mcap * MR_MANIFOLD_POINT_CAPACITY);
        uint32_t wmc = *emc < mcap ? *emc : mcap;
        EnvState st{wq.data(), wv.data(), ws.data(), wmh.data(), wmp.data(), &wmc};
        MRContactStatusC cs;
        std::memset(&cs, 0, sizeof(cs));
        cs.first_failing_pair = MR_INVALID_INDEX;
        cs.first_failing_constraint = MR_INVALID_INDEX;
        MRWorldStatusC ws_out;
        std::memset(&ws_out, 0, sizeof(ws_out));
        ws_out.environment = env;
        ws_out.control_step = control_step;
        ws_out.failing_substep = MR_INVALID_INDEX;
        ws_out.failing_index = MR_INVALID_INDEX;
        float min_piv = INFINITY, max_piv = 0.0f, max_acc = 0.0f;
        uint32_t ev_count = 0;
        MicroOutputs mo;
        mo.cstatus = &cs;
        mo.pairs = evidence_pairs != nullptr ? evidence_pairs + (size_t)env * cfg->pair_capacity : nullptr;
        mo.pair_count = &ev_count;
        const float* cmd = command + (size_t)env * max_nv;
        Failure fail;
        const bool hybrid = cfg->ccd_mode == MR_CCD_HYBRID;
        const float h_full = ctx.h;
        uint32_t impact_segments = 0u;
        for (uint32_t sub = 0; sub < cfg->physics_substeps; ++sub) {
            const bool capture = cfg->capture_evidence != 0u && sub == cfg->capture_substep;
            // hybrid CCD: segments ending at successive times of impact, each a
            // full microstep of its own duration
            float remaining = h_full;
            bool failed = false;
            for (uint32_t seg = 0;; ++seg) {
                bool impact = false;
                float duration = remaining;
                if (hybrid && !ccd_segment(ctx, st, *wk, remaining, seg, fail, impact, duration)) {
                    failed = true;
                    break;
                }
                ctx.h = duration;
                const bool ok = microstep(ctx, st, cmd, sub == 0u && seg == 0u,
                                          sub + 1u == cfg->physics_substeps && !impact, capture && seg == 0u, mo,
                                          *wk, fail, cons, jrow, resp, seg == 0u);
                ctx.h = h_full;
                if (!ok) {
                    failed = true;
                    break;
                }
                min_piv = fmind(min_piv, wk->minPivot);
                max_piv = fmaxd(max_piv, wk->maxPivot);
                for (uint32_t j = 0; j < hd.nv; ++j) max_acc = fmaxd(max_acc, std::fabs(wk->accel[j]));
                if (!impact) break;
                ++impact_segments;
                remaining -= duration;
            }
            if (failed) {
                ws_out.failing_substep = sub;
                ws_out.failing_index = fail.index;
                break;
            }
            ws_out.successful_substeps = sub + 1u;
        }
        ws_out.code = fail.code;
        cs.code = fail.code;
        if (hybrid) cs.diagnostics.w = (float)impact_segments;
        if (fail.code == MR_STEP_SUCCESS) {
            for (uint32_t i = 0; i < hd.nq; ++i) eq[i] = wq[i];
            for (uint32_t i = 0; i < hd.nv; ++i) ev[i] = wv[i];
            for (uint32_t s = 0; s < hd.scene_body_count; ++s) es[s] = ws[s];
            for (uint32_t m = 0; m < wmc; ++m) emh[m] = wmh[m];
            for (size_t p = 0; p < (size_t)wmc * MR_MANIFOLD_POINT_CAPACITY; ++p) emp[p] = wmp[p];
            *emc = wmc;
            if (acceleration != nullptr)
                for (uint32_t j = 0; j < hd.nv; ++j) acceleration[(size_t)env * max_nv + j] = wk->accel[j];
            float qerr = 0.0f;
            if (hd.root_type == MR_ROOT_FLOATING) {
                const float n2 = eq[3] * eq[3] + eq[4] * eq[4] + eq[5] * eq[5] + eq[6] * eq[6];
                qerr = std::fabs(std::sqrt(n2) - 1.0f);
            }
            ws_out.diagnostics = mr_float4{std::isfinite(min_piv) ? min_piv : 0.0f, max_piv, max_acc, qerr};
        } else {
            if (acceleration != nullptr)
                for (uint32_t j = 0; j < hd.nv; ++j) acceleration[(size_t)env * max_nv + j] = 0.0f;
            ws_out.diagnostics = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
            if (fail.code == MR_STEP_PAIR_CAPACITY_OVERFLOW) cs.first_failing_pair = fail.index;
            else cs.first_failing_constraint = fail.index;
        }
        status[env] = ws_out;
        cstatus[env] = cs;
        if (evidence_counts != nullptr) evidence_counts[env] = ev_count;
    }
    delete wk;
    return 0;
}

// Multi-threaded runner for the CPU baseline: splits [0, env_count) into
// contiguous blocks over `threads` std::threads.
int mr_oracle_control_step_parallel(const MRWorldModelDescC* desc, const MRWorldStepConfigC* cfg,
                                    uint32_t env_count, uint32_t control_step, uint32_t threads,
                                    float* q, float* v, MRBodyStateC* scene, MRManifoldHeaderC* mh,
                                    MRManifoldPointC* mp, uint32_t* mcount, const float* command,
                                    MRWorldStatusC* status, MRContactStatusC* cstatus,
                                    float* acceleration) {
    if (threads <= 1u) {
        return mr_oracle_control_step(desc, cfg, 0u, env_count, control_step, q, v, scene, mh, mp, mcount,
                                      command, nullptr, nullptr, nullptr, nullptr, status, cstatus,
                                      acceleration, nullptr, nullptr);
    }
    std::vector<std::thread> pool;
    std::vector<int> rc(threads, 0);
    for (uint32_t t = 0; t < threads; ++t) {
        const uint32_t b = (uint32_t)(((uint64_t)env_count * t) / threads);
        const uint32_t e = (uint32_t)(((uint64_t)env_count * (t + 1u)) / threads);
        pool.emplace_back([=, &rc]() {
            rc[t] = mr_oracle_control_step(desc, cfg, b, e, control_step, q, v, scene, mh, mp, mcount, command,
                                           nullptr, nullptr, nullptr, nullptr, status, cstatus, acceleration,
                                           nullptr, nullptr);
        });
    }
    for (auto& th : pool) th.join();
    for (int r : rc)
        if (r != 0) return r;
    return 0;
}

// Unconstrained FP32 forward dynamics of one environment (ABA only): returns
// generalized accelerations for state (q, v) and effort `tau`; status code
// as return value. Used by the free-body known-answer and FP64 parity tests.
int mr_oracle_forward_dynamics(const MRWorldModelDescC* desc, const MRWorldStepConfigC* cfg, uint32_t env,
                               const float* q, const float* v, const float* tau, float* acc_out) {
    StepContext ctx;
    ctx.desc = desc;
    ctx.cfg = cfg;
    ctx.m = env_model(desc, env);
    ctx.env = env;
    ctx.h = cfg->control_timestep / (float)cfg->physics_substeps;
    Work* wk = new Work();
    for (uint32_t j = 0; j < ctx.m.header->nv; ++j) wk->effort[j] = tau[j];
    Failure f;
    int code = 0;
    if (!forward_kinematics(ctx, q, v, *wk, f) || !aba(ctx, v, *wk, f)) code = (int)f.code;
    else
        for (uint32_t j = 0; j < ctx.m.header->nv; ++j) acc_out[j] = wk->accel[j];
    delete wk;
    return code;
}

// Inverse-ABA response M_eff^{-1} J^T f for one point force on link `link`
// at world point `point` (3 floats) with force `force` (3 floats); state
// (q, v) fixes the configuration. Used to check the response against a dense
// FP64 mass matrix.
int mr_oracle_point_response(const MRWorldModelDescC* desc, const MRWorldStepConfigC* cfg, uint32_t env,
                             const float* q, const float* v, uint32_t link, const float* point,
                             const float* force, float* out) {
    StepContext ctx;
    ctx.desc = desc;
    ctx.cfg = cfg;
    ctx.m = env_model(desc, env);
    ctx.env = env;
    ctx.h = cfg->control_timestep / (float)cfg->physics_substeps;
    Work* wk = new Work();
    for (uint32_t j = 0; j < ctx.m.header->nv; ++j) wk->effort[j] = 0.0f;
    Failure f;
    int code = 0;
    if (!forward_kinematics(ctx, q, v, *wk, f) || !aba(ctx, v, *wk, f)) {
        code = (int)f.code;
    } else {
        PointForce pf{link, V3{point[0], point[1], point[2]}, V3{force[0], force[1], force[2]}};
        inverse_aba(ctx, *wk, &pf, 1u, out);
    }
    delete wk;
    return code;
}

// Link world poses (position xyz + quaternion xyzw per link) for state q.
int mr_oracle_link_poses(const MRWorldModelDescC* desc, const MRWorldStepConfigC* cfg, uint32_t env,
                         const float* q, const float* v, float* out) {
    StepContext ctx;
    ctx.desc = desc;
    ctx.cfg = cfg;
    ctx.m = env_model(desc, env);
    ctx.env = env;
    ctx.h = cfg->control_timestep / (float)cfg->physics_substeps;
    Work* wk = new Work();
    Failure f;
    int code = 0;
    if (!forward_kinematics(ctx, q, v, *wk, f)) code = (int)f.code;
    else
        for (uint32_t i = 0; i < ctx.m.header->link_count; ++i) {
            out[7 * i + 0] = wk->pos[i].x; out[7 * i + 1] = wk->pos[i].y; out[7 * i + 2] = wk->pos[i].z;
            out[7 * i + 3] = wk->rot[i].x; out[7 * i + 4] = wk->rot[i].y; out[7 * i + 5] = wk->rot[i].z;
            out[7 * i + 6] = wk->rot[i].w;
        }
    delete wk;
    return code;
}

// Broadphase flags and narrowphase witnesses of every compiled pair of one
// environment for state (q, v, scene): overlap[npairs], counts[npairs],
// reals[npairs][8][10] = (normal xyz, separation, point A xyz, point B xyz),
// features[npairs][8][2]. Returns the number of compiled pairs, or -1.
int mr_oracle_collide(const MRWorldModelDescC* desc, const MRWorldStepConfigC* cfg, uint32_t env, const float* q,
                      const float* v, const MRBodyStateC* scene, uint32_t* overlap, uint32_t* counts, float* reals,
                      uint32_t* features) {
    StepContext ctx;
    ctx.desc = desc;
    ctx.cfg = cfg;
    ctx.m = env_model(desc, env);
    ctx.env = env;
    ctx.h = cfg->control_timestep / (float)cfg->physics_substeps;
    const MRModelHeaderC& hd = *ctx.m.header;
    Work* wk = new Work();
    Failure f;
    int rc = (int)hd.pair_count;
    if (!forward_kinematics(ctx, q, v, *wk, f)) {
        rc = -1;
    } else {
        for (uint32_t i = 0; i < hd.link_count; ++i) {
            MRBodyStateC& b = wk->bodies[i];
            std::memset(&b, 0, sizeof(b));
            b.position = mr_float4{wk->pos[i].x, wk->pos[i].y, wk->pos[i].z, 1.0f};
            b.orientation = mr_float4{wk->rot[i].x, wk->rot[i].y, wk->rot[i].z, wk->rot[i].w};
        }
        for (uint32_t b = hd.link_count; b < hd.body_count; ++b) wk->bodies[b] = scene[b - hd.link_count];
        for (uint32_t s = 0; s < hd.shape_count && rc >= 0; ++s)
            if (!project_collider(ctx, *wk, s, wk->col[s])) rc = -1;
        for (uint32_t p = 0; p < hd.pair_count && rc >= 0; ++p) {
            const MRCompiledPairC& pr = ctx.m.pairs[p];
            overlap[p] = pair_may_overlap(wk->col[pr.collider_a], wk->col[pr.collider_b]) ? 1u : 0u;
            RawBatch raw;
            raw.count = 0;
            generate_contacts(ctx, pr.collider_a, pr.collider_b, pr.pair_class, wk->col[pr.collider_a],
                              wk->col[pr.collider_b], raw);
            counts[p] = raw.count;
            for (uint32_t c = 0; c < raw.count; ++c) {
                float* o = reals + ((size_t)p * MR_RAW_CONTACTS_PER_PAIR + c) * 10u;
                const RawContact& r = raw.c[c];
                o[0] = r.normal.x; o[1] = r.normal.y; o[2] = r.normal.z; o[3] = r.separation;
                o[4] = r.point_a.x; o[5] = r.point_a.y; o[6] = r.point_a.z;
                o[7] = r.point_b.x; o[8] = r.point_b.y; o[9] = r.point_b.z;
                features[((size_t)p * MR_RAW_CONTACTS_PER_PAIR + c) * 2u] = r.feature_a;
                features[((size_t)p * MR_RAW_CONTACTS_PER_PAIR + c) * 2u + 1u] = r.feature_b;
            }
        }
    }
    delete wk;
    return rc;
}

// Captures the final-microstep contact problem of environment `env` during the
// next serial mr_oracle_control_step calls (buffer == null turns it off).
// Returns the number of floats written by the last capture through `used`.
void mr_oracle_set_problem_capture(uint32_t env, float* buffer, uint32_t capacity_floats) {
    g_capture.env = env;
    g_capture.out = buffer;
    g_capture.capacity = capacity_floats;
    g_capture.used = 0;
}

uint32_t mr_oracle_problem_capture_used(void) { return g_capture.used; }

void mr_oracle_sincos(const float* x, float* s, float* c, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) det_sincos(x[i], s[i], c[i]);
}

uint32_t mr_oracle_abi_version(void) { return MR_WORLD_ABI_VERSION; }

} // extern "C"
