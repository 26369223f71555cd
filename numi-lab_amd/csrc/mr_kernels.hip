// Fused temporalCone control-step kernel for gfx950.
//
// One 64-lane workgroup (one wavefront) owns one world for a whole control
// step: reset/checkpoint, `physics_substeps` microsteps and the transactional
// commit. The world's articulation, colliders, active pairs and solver rows
// live in LDS; persistent state (q, v, scene bodies, manifolds) is read once
// at the start and written once at the end of the control step, so a failed
// step leaves HBM untouched (the checkpoint is the unmodified state). Lanes
// work on independent items (links of one tree level, colliders, compiled
// pairs, manifolds, constraint rows, generalized velocities); every item is
// computed with a fixed scalar evaluation order, and every order-dependent
// accumulation (children into a parent, contact impulses into a velocity)
// is performed by one owner lane in a fixed sequence. Results are therefore
// deterministic and independent of launch shape.
//
// Stage map to the reference (paths relative to the reference repository):
//   drive refresh        src/metal/MetalWorld.metal:280-366
//   kinematics + ABA     src/metal/ArticulatedABA.metal:352-1440
//   scene prediction     src/metal/MetalWorldContact.metal:2509-2567
//   collider projection  src/metal/Collision.metal:321-600, 5666-5712
//   pair flags / scan    src/metal/Collision.metal:796-858, 7905-8176
//   narrowphase          src/metal/Collision.metal:1003-1720, 4379-4830
//   manifolds            src/metal/Collision.metal:5108-5530, 9557-9830
//   emission             src/metal/Collision.metal:9180-9470
//   responses            src/metal/ArticulatedInverseMass.metal:1652-1700
//   IR evaluation        src/metal/MetalWorldContact.metal:3800-3995
//   cone solve           src/metal/MetalWorldContact.metal:482-720, 7856-8316
//   joint limits         src/metal/MetalWorldContact.metal:2878-2960
//   integration          src/metal/MetalWorldContact.metal:8876-9043
//   latch / commit       src/metal/MetalWorldContact.metal:9051-9082, MetalWorld.metal:372-519

#include "mr_device_math.h"
#include "mr_kernel_args.h"
#include "../../include/mr_c_api.h"
#include "mr_mlp_tile.h"

#include <mutex>

using namespace mrd;
using mr::ColdConstraint;
using mr::HotConstraint;
using mr::SolveBlock;
using mr::GeoBlock;
using mr::KernelArgs;

namespace {

constexpr float kTiny = 1.0e-14f;
constexpr float kFloatEps = 1.1920928955078125e-7f;
constexpr float kAbsolutePivotFloor = 1.0e-12f;
constexpr float kMatrixFloor = 1.0e-10f;
constexpr float kConeEpsilon = 1.0e-7f;
constexpr float kContactRegularization = 1.0e-2f;
constexpr uint32_t kMaxCandidates = 12u;
constexpr uint32_t kLanes = 64u;

MRD_INL V3 f3(const mr_float4& v) { return V3{v.x, v.y, v.z}; }
MRD_INL Q4 q4(const mr_float4& v) { return Q4{v.x, v.y, v.z, v.w}; }
MRD_INL uint32_t feature_key(uint32_t t, uint32_t l) { return ((t & 0x0fu) << 28u) | (l & 0x0fffffffu); }

struct SpI {
    float A[6];
    float B[9];
    float C[6];
};

struct Collider {
    uint32_t type, body, disabled;
    V3 center;
    Q4 rotation;
    V3 lower, upper;
    float radius, half_length, contact_offset;
    V3 half_extents, p0, p1, plane_normal;
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

MRD_INL V3 sym_mul(const float* s, V3 w) {
    return V3{s[0] * w.x + s[1] * w.y + s[2] * w.z, s[1] * w.x + s[3] * w.y + s[4] * w.z,
              s[2] * w.x + s[4] * w.y + s[5] * w.z};
}
MRD_INL V3 mat_mul(const float* b, V3 v) {
    return V3{b[0] * v.x + b[1] * v.y + b[2] * v.z, b[3] * v.x + b[4] * v.y + b[5] * v.z,
              b[6] * v.x + b[7] * v.y + b[8] * v.z};
}
MRD_INL V3 matT_mul(const float* b, V3 w) {
    return V3{b[0] * w.x + b[3] * w.y + b[6] * w.z, b[1] * w.x + b[4] * w.y + b[7] * w.z,
              b[2] * w.x + b[5] * w.y + b[8] * w.z};
}
MRD_INL int sym_index(int i, int j) {
    return (i == 0) ? j : ((i == 1) ? (j == 0 ? 1 : (j == 1 ? 3 : 4)) : (j == 0 ? 2 : (j == 1 ? 4 : 5)));
}

// I' = X^T I X for a translation r: A' = A - (Q + Q^T) - r~ C r~, B' = B + r~ C.
MRD_INL void shift_inertia(const SpI& in, V3 r, SpI& out) {
    float Cm[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Cm[i][j] = in.C[sym_index(i, j)];
    float T[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T[0][j] = r.y * Cm[2][j] - r.z * Cm[1][j];
        T[1][j] = r.z * Cm[0][j] - r.x * Cm[2][j];
        T[2][j] = r.x * Cm[1][j] - r.y * Cm[0][j];
    }
    float Q[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float b0 = in.B[3 * i + 0], b1 = in.B[3 * i + 1], b2 = in.B[3 * i + 2];
        Q[i][0] = b1 * r.z - b2 * r.y;
        Q[i][1] = b2 * r.x - b0 * r.z;
        Q[i][2] = b0 * r.y - b1 * r.x;
    }
    float P[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        P[i][0] = T[i][1] * r.z - T[i][2] * r.y;
        P[i][1] = T[i][2] * r.x - T[i][0] * r.z;
        P[i][2] = T[i][0] * r.y - T[i][1] * r.x;
    }
    const int si[6] = {0, 0, 0, 1, 1, 2};
    const int sj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; ++k) out.A[k] = in.A[k] - (Q[si[k]][sj[k]] + Q[sj[k]][si[k]]) - P[si[k]][sj[k]];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out.B[3 * i + j] = in.B[3 * i + j] + T[i][j];
#pragma unroll
    for (int k = 0; k < 6; ++k) out.C[k] = in.C[k];
}

MRD_INL float spi_max_abs(const SpI& I) {
    float m = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) m = fmaxd(m, fabs_(I.A[k]));
#pragma unroll
    for (int k = 0; k < 9; ++k) m = fmaxd(m, fabs_(I.B[k]));
#pragma unroll
    for (int k = 0; k < 6; ++k) m = fmaxd(m, fabs_(I.C[k]));
    return m;
}

MRD_INL float spi6(const SpI& I, int r, int c) {
    if (r < 3 && c < 3) return I.A[sym_index(r, c)];
    if (r < 3) return I.B[3 * r + (c - 3)];
    if (c < 3) return I.B[3 * c + (r - 3)];
    return I.C[sym_index(r - 3, c - 3)];
}

// ------------------------------------------------------------ environment

struct Env {
    const KernelArgs* a;
    uint32_t env;
    const MRModelHeaderC* hd;
    const MRBodyC* bodies;
    const MRDofC* dofs;
    const MRActuatorProfileC* acts;  // [nv] or null: every DoF inactive
    const MRShapeC* shapes;
    const MRMaterialC* mats;
    const MRCompiledPairC* pairs;
    uint32_t nlinks, nbodies, nv, nq, nshapes, npairs, nscene, levels, root_type;
    float h;
    V3 gravity;
    // LDS
    mr::TopoCache* topo;
    mr::LinkConst* lconst;
    float* arm;  // effective armature per DoF (constant for the launch)
    float *q, *v, *vcand, *effort, *accel, *cmd;
    V3 *pos, *w, *vl, *sa, *sl, *ca, *cl, *r;
    Q4* rot;
    SpI* I;
    V3 *pt, *pb, *Ut, *Ub, *alpha, *acc;
    float *D, *u, *L;
    MRBodyStateC* scene;
    Collider* col;
    uint32_t *active, *npts, *moff, *okey;
    uint32_t* hfitem;  // [kHfSlots]: active pair | query << 16
    mr::HfHit* hfq;    // [kHfSlots]: heightfield query results
    HotConstraint* hot;  // global scratch
    float* lam;          // LDS [ccap][3]
    uint32_t* cpair;     // LDS [ccap]: body_a | body_b << 8 | kind << 16
    float *rslot, *rstack;
    uint8_t* isl;
    float* resid;
    float *jrelax, *jdelta;  // relaxed-Jacobi damping per contact and impulse deltas
    // MLP instance only
    unsigned char* mlp_lds;  // warm-start MLP tile scratch
    float* mlp_lam;          // [ccap][3] plain warm start, kept for the fallback
    MRBodyStateC* mlp_scene; // scene bodies before the warm start
    // global scratch of this environment
    MRManifoldHeaderC* stage_h;
    MRManifoldPointC* stage_p;
    float* jrow;
    float* resp;
    ColdConstraint* cold;
};

template <typename T>
MRD_INL T* lds_at(unsigned char* base, uint32_t off) {
    return reinterpret_cast<T*>(base + off);
}

MRD_INL bool is_link(const Env& e, uint32_t b) { return b < e.nlinks; }
MRD_INL uint32_t link_parent(const Env& e, uint32_t i) {
    const uint32_t p = e.topo->parent[i];
    return p == 0xffu ? MR_INVALID_INDEX : p;
}
MRD_INL uint32_t link_vidx(const Env& e, uint32_t i) {
    const uint32_t v = e.topo->vidx[i];
    return v == 0xffu ? MR_INVALID_INDEX : v;
}
MRD_INL bool is_dynamic_scene(const Env& e, uint32_t b) {
    return b >= e.nlinks && e.topo->motion[b] == MR_MOTION_DYNAMIC;
}

MRD_INL void body_pose(const Env& e, uint32_t b, V3& p, Q4& q) {
    if (b < e.nlinks) {
        p = e.pos[b];
        q = e.rot[b];
    } else {
        const MRBodyStateC& s = e.scene[b - e.nlinks];
        p = f3(s.position);
        q = q4(s.orientation);
    }
}

// Packed failure key: (order << 32) | (code << 24) | index (24 bits).
MRD_INL void report_failure(unsigned long long* slot, uint32_t order, uint32_t code, uint32_t index) {
    const unsigned long long key = ((unsigned long long)order << 32) | ((unsigned long long)(code & 0xffu) << 24) |
        (unsigned long long)(index & 0xffffffu);
    atomicMin(slot, key);
}

// -------------------------------------------------------------- wave ops

MRD_INL uint32_t wave_prefix_sum(uint32_t value, uint32_t lane, uint32_t& total) {
    uint32_t x = value;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, (unsigned)d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - value;
}

// One butterfly step through a DPP lane permutation: x + x[perm(lane)].
template <int kCtrl>
MRD_INL float dpp_fold_step(float x) {
    const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xf, 0xf, false);
    return x + __int_as_float(y);
}

// Canonical 64-lane pairwise sum x[i] + x[i ^ d] for d = 1, 2, 4, ..., 32.
// d = 1 and 2 are quad permutations; after them every lane of a quad holds
// the same bits (IEEE addition commutes), so the half-row mirror pairs quad 0
// with quad 1 exactly as xor 4 would, and the row mirror then pairs the two
// halves of a row as xor 8. A 16-lane row swap then gives every lane the even
// and the odd row of its row pair (xor 16), and a 32-lane swap the two halves
// (xor 32): the four row sums combine as (r0 + r1) + (r2 + r3).
// Must be called with all 64 lanes active; every lane returns the same value.
MRD_INL float wave_fold(float x) {
    x = dpp_fold_step<0xB1>(x);   // quad_perm [1,0,3,2]: xor 1
    x = dpp_fold_step<0x4E>(x);   // quad_perm [2,3,0,1]: xor 2
    x = dpp_fold_step<0x141>(x);  // row_half_mirror: xor 4 on quad-uniform data
    x = dpp_fold_step<0x140>(x);  // row_mirror: xor 8 on half-row-uniform data
    const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    x = __uint_as_float(r16[0]) + __uint_as_float(r16[1]);  // even row + odd row
    const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(r32[0]) + __uint_as_float(r32[1]);  // lower half + upper half
}

// wave_fold of N independent values, step by step across all of them so that
// the N dependency chains interleave; each result has the bits wave_fold gives.
template <int N> MRD_INL void wave_fold_n(float (&x)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = dpp_fold_step<0xB1>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = dpp_fold_step<0x4E>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = dpp_fold_step<0x141>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = dpp_fold_step<0x140>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x[i]), __float_as_uint(x[i]), false, false);
        x[i] = __uint_as_float(r16[0]) + __uint_as_float(r16[1]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[i]), __float_as_uint(x[i]), false, false);
        x[i] = __uint_as_float(r32[0]) + __uint_as_float(r32[1]);
    }
}

// True on every lane when `bad` holds on any active lane: a uniform branch
// condition for values every lane computes alike.
MRD_INL bool wave_any(bool bad) { return __builtin_amdgcn_ballot_w64(bad) != 0ull; }

// A compile-time flag passed to a generic lambda.
template <bool B> struct BoolTag { static constexpr bool value = B; };

MRD_INL float wave_max(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxd(v, __shfl_xor(v, d, 64));
    return v;
}

MRD_INL uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

MRD_INL uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(v, d, 64);
        v = v > o ? v : o;
    }
    return v;
}

MRD_INL uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(v, d, 64);
        v = v < o ? v : o;
    }
    return v;
}

// ----------------------------------------------------------------- drives

MRD_INL float effective_armature(const Env& e, uint32_t j) {
    const MRDofC& dof = e.dofs[j];
    float value = dof.drive.z;
    if (e.a->cfg.actuation_mode == MR_ACTUATION_IMPLICIT_POSITION_DRIVE && (dof.flags & MR_DOF_FLAG_DRIVE) != 0u)
        value += e.h * dof.drive.y + e.h * e.h * dof.drive.x;
    return value;
}

// Torque-speed envelope of DoF j at joint velocity vel (MetalWorld.metal:
// 252-269): min(effort limit, stall x efficiency x (1 - speed fraction)). An
// inactive profile executes as no-load speed FLT_MAX, efficiency 1 and stall
// torque = the effort limit, or FLT_MAX without one (MetalWorld.mm:1376-1410).
MRD_INL float actuator_envelope(const MRDofC& dof, const MRActuatorProfileC* act, float vel) {
    constexpr float kFloatMax = 3.40282346638528859812e38f;
    float no_load = kFloatMax, efficiency = 1.0f;
    float stall = ((dof.flags & MR_DOF_FLAG_EFFORT_LIMIT) != 0u && dof.limits.w > 0.0f) ? dof.limits.w : kFloatMax;
    if (act != nullptr && (act->identity[1] & MR_ACTUATOR_PROFILE_ACTIVE) != 0u) {
        no_load = act->motor_and_speed.z;
        efficiency = act->motor_and_speed.w;
        stall = act->transmission_and_envelope.z;
    }
    const float speed_fraction = clampd(fabs_(vel) / fmaxd(no_load, 1.175494351e-38f), 0.0f, 1.0f);
    return fmind(dof.limits.w, stall * efficiency * (1.0f - speed_fraction));
}

// Actuator command of DoF j from the current q / v, clamped to the envelope
// (MetalWorld.metal:184-270 prepare, 280-366 refresh).
MRD_INL float drive_effort(const Env& e, uint32_t j) {
    const MRDofC& dof = e.dofs[j];
    const float command = e.cmd[j];
    const float envelope = actuator_envelope(dof, e.acts != nullptr ? &e.acts[j] : nullptr, e.v[j]);
    if (e.a->cfg.actuation_mode == MR_ACTUATION_EFFORT) return clampd(command, -envelope, envelope);
    float cmd = 0.0f;
    if ((dof.flags & MR_DOF_FLAG_DRIVE) != 0u && dof.q_index != MR_INVALID_INDEX) {
        float target = command;
        if ((dof.flags & MR_DOF_FLAG_POSITION_LIMIT) != 0u) target = clampd(target, dof.limits.x, dof.limits.y);
        const float vel = e.v[j];
        cmd = dof.drive.x * (target - e.q[dof.q_index] - e.h * vel) - dof.drive.y * vel;
        const float dry = dof.drive.w;
        if (dry > 0.0f) {
            if (fabs_(vel) > 1.0e-4f) cmd -= vel > 0.0f ? dry : -dry;
            else cmd -= clampd(cmd, -dry, dry);
        }
        if ((dof.flags & MR_DOF_FLAG_EFFORT_LIMIT) != 0u && dof.limits.w > 0.0f)
            cmd = clampd(cmd, -dof.limits.w, dof.limits.w);
    }
    return clampd(cmd, -envelope, envelope);
}

// ------------------------------------------------------------- kinematics

// Forward kinematics of one non-root link from its parent (lane-local).
MRD_INL bool fk_link(const Env& e, uint32_t i) {
    const mr::LinkConst lc = e.lconst[i];
    const uint32_t p = link_parent(e, i);
    const uint32_t jt = e.topo->jtype[i];
    const Q4 pjr = qmul(e.rot[p], Q4{lc.prot[0], lc.prot[1], lc.prot[2], lc.prot[3]});
    const bool moving = jt == MR_JOINT_REVOLUTE || jt == MR_JOINT_CONTINUOUS;
    const V3 axis_n{lc.axis[0], lc.axis[1], lc.axis[2]};
    float qi = 0.0f, qd = 0.0f;
    Q4 motion{0.0f, 0.0f, 0.0f, 1.0f};
    if (moving) {
        qi = e.q[e.topo->qidx[i]];
        qd = e.v[link_vidx(e, i)];
        motion = axis_angle(axis_n, qi);
    }
    Q4 child;
    if (!normalized_quat(qmul(qmul(pjr, motion), Q4{lc.crot_conj[0], lc.crot_conj[1], lc.crot_conj[2], lc.crot_conj[3]}),
                         child))
        return false;
    e.rot[i] = child;
    const V3 axis_w = qrot(pjr, axis_n);
    const V3 joint_pos = add(e.pos[p], qrot(e.rot[p], V3{lc.panchor[0], lc.panchor[1], lc.panchor[2]}));
    const V3 child_anchor = qrot(child, V3{lc.canchor[0], lc.canchor[1], lc.canchor[2]});
    const V3 pos = sub(joint_pos, child_anchor);
    e.pos[i] = pos;
    e.r[i] = sub(pos, e.pos[p]);
    const V3 p2j = sub(joint_pos, e.pos[p]);
    const V3 wp = e.w[p];
    const V3 zero{0.0f, 0.0f, 0.0f};
    const V3 wi = moving ? add(wp, mul(qd, axis_w)) : wp;
    e.w[i] = wi;
    const V3 vli = sub(add(e.vl[p], cross(wp, p2j)), cross(wi, child_anchor));
    e.vl[i] = vli;
    e.sa[i] = moving ? axis_w : zero;
    e.sl[i] = moving ? neg(cross(axis_w, child_anchor)) : zero;
    const V3 cai = moving ? mul(qd, cross(wp, axis_w)) : zero;
    e.ca[i] = cai;
    const V3 cli = sub(sub(cross(wp, cross(wp, p2j)), cross(cai, child_anchor)), cross(wi, cross(wi, child_anchor)));
    e.cl[i] = cli;
    return finite3(pos) && finite4(child) && finite3(wi) && finite3(vli) && finite3(cli) && finite3(cai);
}

MRD_INL void world_inertia(const MRBodyC& b, Q4 rot, float out[6]) {
    const V3 c0 = qrot(rot, V3{1.0f, 0.0f, 0.0f});
    const V3 c1 = qrot(rot, V3{0.0f, 1.0f, 0.0f});
    const V3 c2 = qrot(rot, V3{0.0f, 0.0f, 1.0f});
    const float R[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
    const float Ib[3][3] = {{b.inertia_row0.x, b.inertia_row0.y, b.inertia_row0.z},
                            {b.inertia_row1.x, b.inertia_row1.y, b.inertia_row1.z},
                            {b.inertia_row2.x, b.inertia_row2.y, b.inertia_row2.z}};
    const int si[6] = {0, 0, 0, 1, 1, 2};
    const int sj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int rr = si[k], cc = sj[k];
        float acc = 0.0f;
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y) acc += R[rr][x] * Ib[x][y] * R[cc][y];
        out[k] = acc;
    }
}

MRD_INL void world_inverse_inertia(const MRBodyC& b, Q4 rot, MRBodyStateC& s) {
    const V3 c0 = qrot(rot, V3{1.0f, 0.0f, 0.0f});
    const V3 c1 = qrot(rot, V3{0.0f, 1.0f, 0.0f});
    const V3 c2 = qrot(rot, V3{0.0f, 0.0f, 1.0f});
    const float R[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
    const float Ib[3][3] = {{b.inverse_inertia_row0.x, b.inverse_inertia_row0.y, b.inverse_inertia_row0.z},
                            {b.inverse_inertia_row1.x, b.inverse_inertia_row1.y, b.inverse_inertia_row1.z},
                            {b.inverse_inertia_row2.x, b.inverse_inertia_row2.y, b.inverse_inertia_row2.z}};
    float o[3][3];
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            float acc = 0.0f;
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) acc += R[rr][x] * Ib[x][y] * R[cc][y];
            o[rr][cc] = acc;
        }
    s.inverse_inertia_world_row0 = mr_float4{o[0][0], o[0][1], o[0][2], 0.0f};
    s.inverse_inertia_world_row1 = mr_float4{o[1][0], o[1][1], o[1][2], 0.0f};
    s.inverse_inertia_world_row2 = mr_float4{o[2][0], o[2][1], o[2][2], 0.0f};
}

// ABA initialisation of one link (rigid inertia + bias force).
MRD_INL void aba_init_link(const Env& e, uint32_t i) {
    const MRBodyC& b = e.bodies[i];
    SpI I;
    world_inertia(b, e.rot[i], I.A);
#pragma unroll
    for (int k = 0; k < 9; ++k) I.B[k] = 0.0f;
    const float m = b.mass.x;
    I.C[0] = m; I.C[1] = 0.0f; I.C[2] = 0.0f;
    I.C[3] = m; I.C[4] = 0.0f; I.C[5] = m;
    const V3 wi = e.w[i];
    V3 bt = cross(wi, sym_mul(I.A, wi));
    V3 bf = neg(mul(m, e.gravity));
    if (e.a->cfg.apply_body_damping != 0u) {
        bt = add(bt, mul(b.damping.y, wi));
        bf = add(bf, mul(b.damping.x, e.vl[i]));
    }
    e.I[i] = I;
    e.pt[i] = bt;
    e.pb[i] = bf;
}

// Joint projection of link i and its contribution to the parent, written in
// place into I[i], pt[i], pb[i]. Returns false on a pivot failure.
MRD_INL bool aba_project_link(const Env& e, uint32_t i, float& pivot_out) {
    const MRBodyC& b = e.bodies[i];
    const SpI I = e.I[i];
    const V3 Ica_t = add(sym_mul(I.A, e.ca[i]), mat_mul(I.B, e.cl[i]));
    const V3 Ica_b = add(matT_mul(I.B, e.ca[i]), sym_mul(I.C, e.cl[i]));
    V3 pt = add(e.pt[i], Ica_t);
    V3 pb = add(e.pb[i], Ica_b);
    SpI Ip = I;
    pivot_out = -1.0f;
    bool ok = true;
    if (b.v_index != MR_INVALID_INDEX) {
        const V3 sa = e.sa[i], sl = e.sl[i];
        const V3 Ut = add(sym_mul(I.A, sa), mat_mul(I.B, sl));
        const V3 Ub = add(matT_mul(I.B, sa), sym_mul(I.C, sl));
        const float D = (dot(sa, Ut) + dot(sl, Ub)) + e.arm[b.v_index];
        const float floor_ = fmaxd(kAbsolutePivotFloor, spi_max_abs(I) * 6.0f * kFloatEps);
        if (!(D > floor_) || !isfin(D)) ok = false;
        const float u = (e.effort[b.v_index] - dot(sa, pt)) - dot(sl, pb);
        const float uD = u / D;
        pt = add(pt, mul(uD, Ut));
        pb = add(pb, mul(uD, Ub));
        const float ut[3] = {Ut.x, Ut.y, Ut.z};
        const float ub[3] = {Ub.x, Ub.y, Ub.z};
        const int si[6] = {0, 0, 0, 1, 1, 2};
        const int sj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int k = 0; k < 6; ++k) Ip.A[k] = I.A[k] - (ut[si[k]] * ut[sj[k]]) / D;
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y) Ip.B[3 * x + y] = I.B[3 * x + y] - (ut[x] * ub[y]) / D;
#pragma unroll
        for (int k = 0; k < 6; ++k) Ip.C[k] = I.C[k] - (ub[si[k]] * ub[sj[k]]) / D;
        e.Ut[i] = Ut;
        e.Ub[i] = Ub;
        e.D[i] = D;
        e.u[i] = u;
        pivot_out = sqrt_rn(D);
    }
    SpI shifted;
    shift_inertia(Ip, e.r[i], shifted);
    e.I[i] = shifted;
    e.pt[i] = add(pt, cross(e.r[i], pb));
    e.pb[i] = pb;
    return ok;
}

// Parent accumulation: children of `p` (at the level just processed) are
// added in decreasing link index order.
MRD_INL void aba_accumulate(const Env& e, uint32_t p) {
    SpI P = e.I[p];
    V3 pt = e.pt[p], pb = e.pb[p];
    unsigned long long kids = e.topo->children[p];
    while (kids != 0ull) {
        const uint32_t c = 63u - (uint32_t)__builtin_clzll(kids);  // decreasing index
        kids &= ~(1ull << c);
        const SpI& S = e.I[c];
#pragma unroll
        for (int k = 0; k < 6; ++k) P.A[k] += S.A[k];
#pragma unroll
        for (int k = 0; k < 9; ++k) P.B[k] += S.B[k];
#pragma unroll
        for (int k = 0; k < 6; ++k) P.C[k] += S.C[k];
        pt = add(pt, e.pt[c]);
        pb = add(pb, e.pb[c]);
    }
    e.I[p] = P;
    e.pt[p] = pt;
    e.pb[p] = pb;
}

MRD_INL bool root_factor(const Env& e, float& minp, float& maxp, uint32_t& fail_row) {
    const SpI I = e.I[0];
    const float floor_ = fmaxd(kAbsolutePivotFloor, spi_max_abs(I) * 6.0f * kFloatEps);
    float* L = e.L;
    for (int k = 0; k < 36; ++k) L[k] = 0.0f;
    for (int row = 0; row < 6; ++row) {
        for (int col = 0; col <= row; ++col) {
            float val = spi6(I, row, col);
            for (int k = 0; k < col; ++k) val -= L[row * 6 + k] * L[col * 6 + k];
            if (row == col) {
                if (!(val > floor_) || !isfin(val)) {
                    fail_row = (uint32_t)row;
                    return false;
                }
                const float piv = sqrt_rn(val);
                L[row * 6 + row] = piv;
                minp = fmind(minp, piv);
                maxp = fmaxd(maxp, piv);
            } else {
                L[row * 6 + col] = val / L[col * 6 + col];
            }
        }
    }
    return true;
}

MRD_INL void root_solve(const float* L, const float rhs[6], float x[6]) {
    float y[6];
    for (int row = 0; row < 6; ++row) {
        float val = rhs[row];
        for (int col = 0; col < row; ++col) val -= L[row * 6 + col] * y[col];
        y[row] = val / L[row * 6 + row];
    }
    for (int row = 5; row >= 0; --row) {
        float val = y[row];
        for (int col = row + 1; col < 6; ++col) val -= L[col * 6 + row] * x[col];
        x[row] = val / L[row * 6 + row];
    }
}

// Inverse ABA for the point forces of one constraint row on up to two links:
// out[nv] = M_eff^{-1} J^T f, one row per lane.
//
// Inward pass: only links on the root paths of `la` / `lb` carry a bias. Every
// other link's bias and contribution is exactly +0 in every component (the
// accumulators start at +0, +0 - y and x + y never produce -0 from operands
// without -0, and u = (0 - s.t) - s.f is +0 for a +0 bias), so visiting only
// path links in decreasing index order, and adding the at most two pending
// child contributions of a merge link in decreasing child order, reproduces
// the dense pass bit for bit. The joint terms u of path links are pushed on a
// per-lane LDS stack; the outward pass visits all links in increasing index
// order and pops them in reverse. Link accelerations live in the LDS slots of
// the topology cache's plan.
MRD_INL void inverse_aba(const Env& e, uint32_t lane, uint32_t la, V3 xa, V3 fa, uint32_t lb, V3 xb, V3 fb,
                         float* out) {
    const mr::TopoCache& T = *e.topo;
    const uint32_t n = e.nlinks;
    const V3 zero{0.0f, 0.0f, 0.0f};
    const unsigned long long path = (la != MR_INVALID_INDEX ? T.anc[la] : 0ull) |
                                    (lb != MR_INVALID_INDEX ? T.anc[lb] : 0ull);
    float* ustack = e.rstack + lane;
    uint32_t sp = 0;
    // pending child contributions: target link, child index, torque, force
    uint32_t tg0 = MR_INVALID_INDEX, ch0 = 0u, tg1 = MR_INVALID_INDEX, ch1 = 0u;
    V3 c0t = zero, c0f = zero, c1t = zero, c1f = zero;
    auto gather = [&](uint32_t i, V3& t, V3& f) {
        t = zero;
        f = zero;
        if (i == la) {
            t = sub(t, cross(sub(xa, e.pos[i]), fa));
            f = sub(f, fa);
        }
        if (i == lb) {
            t = sub(t, cross(sub(xb, e.pos[i]), fb));
            f = sub(f, fb);
        }
        const bool m0 = tg0 == i, m1 = tg1 == i;
        if (m0 && m1) {
            const bool first0 = ch0 > ch1;
            t = add(t, first0 ? c0t : c1t);
            f = add(f, first0 ? c0f : c1f);
            t = add(t, first0 ? c1t : c0t);
            f = add(f, first0 ? c1f : c0f);
        } else if (m0) {
            t = add(t, c0t);
            f = add(f, c0f);
        } else if (m1) {
            t = add(t, c1t);
            f = add(f, c1f);
        }
        if (m0) tg0 = MR_INVALID_INDEX;
        if (m1) tg1 = MR_INVALID_INDEX;
    };
    unsigned long long rest = path;
    while (rest != 0ull) {
        const uint32_t i = 63u - (uint32_t)__builtin_clzll(rest);
        rest &= ~(1ull << i);
        V3 t, f;
        gather(i, t, f);
        if (link_vidx(e, i) != MR_INVALID_INDEX) {
            const float u = (0.0f - dot(e.sa[i], t)) - dot(e.sl[i], f);
            const float uD = u / e.D[i];
            t = add(t, mul(uD, e.Ut[i]));
            f = add(f, mul(uD, e.Ub[i]));
            ustack[sp * 64u] = u;
            ++sp;
        }
        const V3 ct = add(t, cross(e.r[i], f));
        if (tg0 == MR_INVALID_INDEX) {
            tg0 = link_parent(e, i);
            ch0 = i;
            c0t = ct;
            c0f = f;
        } else {
            tg1 = link_parent(e, i);
            ch1 = i;
            c1t = ct;
            c1f = f;
        }
    }
    V3 ap, lp;
    if (e.root_type == MR_ROOT_FLOATING) {
        V3 t0, f0;
        gather(0u, t0, f0);
        const float rhs[6] = {-t0.x, -t0.y, -t0.z, -f0.x, -f0.y, -f0.z};
        float x[6];
        root_solve(e.L, rhs, x);
        ap = V3{x[0], x[1], x[2]};
        lp = V3{x[3], x[4], x[5]};
        out[0] = x[3]; out[1] = x[4]; out[2] = x[5];
        out[3] = x[0]; out[4] = x[1]; out[5] = x[2];
    } else {
        ap = zero;
        lp = zero;
    }
    auto slot_store = [&](uint32_t s, V3 a, V3 l) {
        float* p = e.rslot + s * 6u * 64u + lane;
        p[0] = a.x; p[64] = a.y; p[128] = a.z;
        p[192] = l.x; p[256] = l.y; p[320] = l.z;
    };
    if (n > 1u) slot_store(T.slot[0], ap, lp);
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t parent = link_parent(e, i), vi = link_vidx(e, i);
        const float* p = e.rslot + (uint32_t)T.slot[parent] * 6u * 64u + lane;
        const V3 a_par{p[0], p[64], p[128]};
        const V3 l_par{p[192], p[256], p[320]};
        const V3 lpi = add(l_par, cross(a_par, e.r[i]));
        V3 ai = a_par, li = lpi;
        if (vi != MR_INVALID_INDEX) {
            float uu = 0.0f;
            if (((path >> i) & 1ull) != 0ull) {
                --sp;
                uu = ustack[sp * 64u];
            }
            const float proj = dot(e.Ut[i], a_par) + dot(e.Ub[i], lpi);
            const float qdd = (uu - proj) / e.D[i];
            out[vi] = qdd;
            ai = add(a_par, mul(qdd, e.sa[i]));
            li = add(lpi, mul(qdd, e.sl[i]));
        }
        if (T.slot[i] != 0xffu) slot_store(T.slot[i], ai, li);
    }
}

MRD_INL V3 point_jacobian_column(const Env& e, uint32_t link, V3 x, uint32_t j) {
    const V3 zero{0.0f, 0.0f, 0.0f};
    if (e.root_type == MR_ROOT_FLOATING && j < 6u) {
        V3 ev = zero;
        if (j < 3u) {
            setcomp(ev, (int)j, 1.0f);
            return ev;
        }
        setcomp(ev, (int)j - 3, 1.0f);
        return cross(ev, sub(x, e.pos[0]));
    }
    const uint32_t owner = e.topo->dof_body[j];
    if (owner < 64u && ((e.topo->anc[link] >> owner) & 1ull) != 0ull)
        return add(e.sl[owner], cross(e.sa[owner], sub(x, e.pos[owner])));
    return zero;
}

// ------------------------------------------------------------ colliders

MRD_INL void inflate_bounds(Collider& c) {
    const V3 scale = add(vmax(vabs(c.lower), vabs(c.upper)), V3{1.0f, 1.0f, 1.0f});
    const V3 pad = mul(MR_COLLISION_AABB_RELATIVE_PAD, scale);
    c.lower = sub(c.lower, pad);
    c.upper = add(c.upper, pad);
}

// Collider of shape s with its body at (bpos, borient). Convex hulls and
// meshes are projected by the kGeo instances only.
template <bool kGeo = false>
MRD_INL bool project_collider_pose(const Env& e, uint32_t s, V3 bpos, Q4 borient, Collider& c) {
    const MRShapeC& sh = e.shapes[s];
    Q4 br, lr;
    if (!checked_quat(borient, br) || !checked_quat(q4(sh.local_rotation), lr)) return false;
    c.type = sh.shape_type;
    c.body = sh.body;
    c.disabled = (sh.flags & MR_SHAPE_FLAG_SIMULATION_DISABLED) != 0u ? 1u : 0u;
    c.rotation = qmul(br, lr);
    c.center = add(bpos, qrot(br, f3(sh.local_position)));
    c.contact_offset = sh.contact_rest_radius.x;
    c.radius = 0.0f;
    c.half_length = 0.0f;
    const V3 zero{0.0f, 0.0f, 0.0f};
    c.half_extents = zero;
    c.p0 = zero;
    c.p1 = zero;
    c.plane_normal = zero;
    c.lower = zero;
    c.upper = zero;
    if (!finite3(c.center) || !finite4(c.rotation)) return false;
    if (c.disabled) return true;
    const float co = c.contact_offset;
    if (sh.shape_type == MR_SHAPE_SPHERE) {
        c.radius = sh.dimensions.x;
        const float ex = c.radius + co;
        c.lower = sub(c.center, V3{ex, ex, ex});
        c.upper = add(c.center, V3{ex, ex, ex});
        inflate_bounds(c);
    } else if (sh.shape_type == MR_SHAPE_CAPSULE) {
        c.radius = sh.dimensions.x;
        c.half_length = sh.dimensions.y;
        const V3 axis = qrot(c.rotation, V3{0.0f, c.half_length, 0.0f});
        c.p0 = sub(c.center, axis);
        c.p1 = add(c.center, axis);
        const float ex = c.radius + co;
        c.lower = sub(vmin(c.p0, c.p1), V3{ex, ex, ex});
        c.upper = add(vmax(c.p0, c.p1), V3{ex, ex, ex});
        inflate_bounds(c);
    } else if (sh.shape_type == MR_SHAPE_BOX) {
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
    } else if (sh.shape_type == MR_SHAPE_PLANE) {
        c.plane_normal = normalize(qrot(c.rotation, V3{0.0f, 1.0f, 0.0f}));
    } else if (sh.shape_type == MR_SHAPE_HEIGHTFIELD) {
        const MRHeightfieldC& hf = e.a->heightfield;
        c.lower = sub(f3(hf.bounds_lower), V3{co, co, co});
        c.upper = add(f3(hf.bounds_upper), V3{co, co, co});
        inflate_bounds(c);
    } else if (kGeo && (sh.shape_type == MR_SHAPE_CONVEX || sh.shape_type == MR_SHAPE_TRIANGLE_MESH)) {
        // half_extents carries the vertex scale; bounds of the transformed
        // vertices (convex) or of the transformed local bounds (mesh)
        c.half_extents = f3(sh.dimensions);
        const MRGeometryC g = e.a->geometries[sh.geometry];
        const V3 sc = c.half_extents;
        V3 lo{__builtin_inff(), __builtin_inff(), __builtin_inff()};
        V3 hi{-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        const bool convex = sh.shape_type == MR_SHAPE_CONVEX;
        const uint32_t count = convex ? g.vertex_count : 8u;
        for (uint32_t i = 0; i < count; ++i) {
            V3 l;
            if (convex) {
                const mr_float4 v = e.a->geometry_vertices[g.vertex_offset + i];
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
    } else {
        return false;
    }
    return finite3(c.lower) && finite3(c.upper);
}

template <bool kGeo = false>
MRD_INL bool project_collider(const Env& e, uint32_t s, Collider& c) {
    V3 bpos;
    Q4 borient;
    body_pose(e, e.shapes[s].body, bpos, borient);
    return project_collider_pose<kGeo>(e, s, bpos, borient, c);
}

MRD_INL float collider_scale(const Collider& c) {
    return fmaxd(fmaxd(maxabs3(c.center), maxabs3(c.half_extents)),
                 fmaxd(fmaxd(maxabs3(c.p0), maxabs3(c.p1)),
                       fmaxd(fmaxd(fabs_(c.radius), fabs_(c.half_length)),
                             fmaxd(fabs_(c.contact_offset), fmaxd(maxabs3(c.lower), maxabs3(c.upper))))));
}

MRD_INL float pair_query_padding(const Collider& a, const Collider& b) {
    return (fmaxd(collider_scale(a), collider_scale(b)) + 1.0f) * MR_COLLISION_QUERY_RELATIVE_PAD;
}

MRD_INL bool pair_may_overlap(const Collider& a, const Collider& b) {
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
                              fmaxd(maxabs3(fin.upper), fmaxd(fabs_(plane.contact_offset), fabs_(fin.contact_offset))));
    const float pad = (scale + 1.0f) * MR_COLLISION_QUERY_RELATIVE_PAD;
    const float dist = dot(plane.plane_normal, sub(center, plane.center)) - dot(vabs(plane.plane_normal), half);
    return dist <= plane.contact_offset + pad;
}

// ------------------------------------------------------------ narrowphase

MRD_INL RawContact make_contact(V3 n, float sep, V3 pa, V3 pb, uint32_t fa, uint32_t fb) {
    RawContact c;
    c.normal = n;
    c.separation = sep;
    c.point_a = pa;
    c.point_b = pb;
    c.feature_a = fa;
    c.feature_b = fb;
    return c;
}

MRD_INL RawContact swapped(const RawContact& in) {
    RawContact c = in;
    c.normal = neg(in.normal);
    c.point_a = in.point_b;
    c.point_b = in.point_a;
    c.feature_a = in.feature_b;
    c.feature_b = in.feature_a;
    return c;
}

MRD_INL V3 coincident_normal(uint32_t a, uint32_t b) {
    const uint32_t hash = (a * 73856093u) ^ (b * 19349663u);
    V3 r{0.0f, 0.0f, 0.0f};
    setcomp(r, (int)(hash % 3u), (hash & 4u) == 0u ? 1.0f : -1.0f);
    return r;
}

MRD_INL V3 stable_segment_normal(V3 da, V3 db, uint32_t ca, uint32_t cb) {
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

MRD_INL SegPoint closest_on_segment(V3 p, V3 e0, V3 e1) {
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

MRD_INL SegPair closest_segments(V3 a0, V3 a1, V3 b0, V3 b1) {
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

MRD_INL uint32_t capsule_feature(float t) {
    if (t <= 0.0f) return 0u;
    return t >= 1.0f ? 1u : 2u;
}

MRD_INL uint32_t box_region(float c, float h) {
    if (c < -h) return 0u;
    return c > h ? 2u : 1u;
}

MRD_INL void sphere_box(const Collider& sphere, const Collider& box, V3& normal, V3& box_point, float& sep,
                        uint32_t& feature) {
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

MRD_INL void consider_capsule_box(V3 seg_point, V3 box_point, float t, uint32_t feat, float& best_sq, V3& best_seg,
                                  V3& best_box, float& best_t, uint32_t& best_feat) {
    const V3 d = sub(box_point, seg_point);
    const float sq = dot(d, d);
    const float tol = 1.0e-12f * (1.0f + fmaxd(best_sq, sq));
    if (!isfin(best_sq) || sq < best_sq - tol || (fabs_(sq - best_sq) <= tol && feat < best_feat)) {
        best_sq = sq;
        best_seg = seg_point;
        best_box = box_point;
        best_t = t;
        best_feat = feat;
    }
}

MRD_INL void capsule_box(const Collider& cap, const Collider& box, uint32_t cap_index, uint32_t box_index,
                         V3& normal, V3& cap_point, V3& box_point, float& sep, float& t_out, uint32_t& feature) {
    const V3 s0 = qrot_inv(box.rotation, sub(cap.p0, box.center));
    const V3 s1 = qrot_inv(box.rotation, sub(cap.p1, box.center));
    const V3 dir = sub(s1, s0);
    const V3 he = box.half_extents;
    float enter = 0.0f, exit_ = 1.0f;
    bool inter = true;
    for (int axis = 0; axis < 3; ++axis) {
        const float da = comp(dir, axis), sa = comp(s0, axis), ha = comp(he, axis);
        if (fabs_(da) <= kTiny) {
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
        exit_ = fmind(exit_, second);
        if (enter > exit_) inter = false;
    }
    if (inter) {
        t_out = clampd(0.5f * (enter + exit_), 0.0f, 1.0f);
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
    float best_sq = __builtin_inff();
    V3 best_seg{0.0f, 0.0f, 0.0f}, best_box{0.0f, 0.0f, 0.0f};
    float best_t = 0.0f;
    uint32_t best_feat = 0xffffffffu;
    for (uint32_t en = 0; en < 2u; ++en) {
        const V3 sp = en == 0u ? s0 : s1;
        const V3 bp{clampd(sp.x, -he.x, he.x), clampd(sp.y, -he.y, he.y), clampd(sp.z, -he.z, he.z)};
        consider_capsule_box(sp, bp, (float)en, en, best_sq, best_seg, best_box, best_t, best_feat);
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
            consider_capsule_box(cl.pa, cl.pb, cl.ta, edge_feature, best_sq, best_seg, best_box, best_t, best_feat);
            ++edge_feature;
        }
    }
    const float dist = sqrt_rn(fmaxd(best_sq, 0.0f));
    const V3 ln = dist > kTiny ? mul(1.0f / dist, sub(best_box, best_seg))
                               : qrot_inv(box.rotation, coincident_normal(cap_index, box_index));
    normal = qrot(box.rotation, ln);
    cap_point = add(box.center, qrot(box.rotation, add(best_seg, mul(cap.radius, ln))));
    box_point = add(box.center, qrot(box.rotation, best_box));
    sep = dist - cap.radius;
    t_out = best_t;
    feature = 64u + best_feat;
}

MRD_INL V3 box_axis(const Collider& box, int axis) {
    V3 l{0.0f, 0.0f, 0.0f};
    setcomp(l, axis, 1.0f);
    return qrot(box.rotation, l);
}

MRD_INL float box_projection_radius(const Collider& box, V3 axis) {
    return fabs_(dot(axis, box_axis(box, 0))) * box.half_extents.x +
        fabs_(dot(axis, box_axis(box, 1))) * box.half_extents.y +
        fabs_(dot(axis, box_axis(box, 2))) * box.half_extents.z;
}

MRD_INL V3 box_vertex(const Collider& box, uint32_t v) {
    const V3 he = box.half_extents;
    const V3 l{(v & 1u) == 0u ? -he.x : he.x, (v & 2u) == 0u ? -he.y : he.y, (v & 4u) == 0u ? -he.z : he.z};
    return add(box.center, qrot(box.rotation, l));
}

MRD_INL bool point_in_inflated_box(V3 p, const Collider& box, float infl) {
    const V3 l = vabs(qrot_inv(box.rotation, sub(p, box.center)));
    return l.x <= box.half_extents.x + infl && l.y <= box.half_extents.y + infl && l.z <= box.half_extents.z + infl;
}

MRD_INL V3 box_support(const Collider& box, V3 d) {
    V3 res = box.center;
    for (int axis = 0; axis < 3; ++axis) {
        const V3 basis = box_axis(box, axis);
        const float he = comp(box.half_extents, axis);
        res = add(res, mul(dot(d, basis) >= 0.0f ? he : -he, basis));
    }
    return res;
}

MRD_INL void append(RawBatch& out, const RawContact& c) {
    if (out.count < MR_RAW_CONTACTS_PER_PAIR) out.c[out.count++] = c;
}

MRD_INL void box_box(uint32_t ia, uint32_t ib, const Collider& A, const Collider& B, float accepted, RawBatch& out) {
    V3 axes[15];
    uint32_t nax = 0;
    for (int x = 0; x < 3; ++x) axes[nax++] = box_axis(A, x);
    for (int x = 0; x < 3; ++x) axes[nax++] = box_axis(B, x);
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) {
            const V3 cr = cross(box_axis(A, x), box_axis(B, y));
            const float sq = dot(cr, cr);
            if (sq > 1.0e-12f) axes[nax++] = mul(1.0f / sqrt_rn(sq), cr);
        }
    const V3 delta = sub(B.center, A.center);
    float best_sep = -__builtin_inff();
    V3 best_n = coincident_normal(ia, ib);
    uint32_t best_axis = 0;
    for (uint32_t k = 0; k < nax; ++k) {
        V3 axis = axes[k];
        const float proj = dot(delta, axis);
        if (proj < 0.0f || (proj == 0.0f && dot(axis, best_n) < 0.0f)) axis = neg(axis);
        const float sep = fabs_(proj) - box_projection_radius(A, axis) - box_projection_radius(B, axis);
        if (sep > accepted) return;
        const float tie = 1.0e-7f * (1.0f + fmaxd(fabs_(best_sep), fabs_(sep)));
        if (!isfin(best_sep) || sep > best_sep + tie || (fabs_(sep - best_sep) <= tie && k < best_axis)) {
            best_sep = sep;
            best_n = axis;
            best_axis = k;
        }
    }
    const float ra = box_projection_radius(A, best_n);
    const float rb = box_projection_radius(B, best_n);
    const float near_b = dot(B.center, best_n) - rb;
    const float far_a = dot(A.center, best_n) + ra;
    for (uint32_t vi = 0; vi < 8u; ++vi) {
        const V3 pa = box_vertex(A, vi);
        if (!point_in_inflated_box(pa, B, accepted)) continue;
        const float sep = near_b - dot(pa, best_n);
        if (sep <= accepted)
            append(out, make_contact(best_n, sep, pa, add(pa, mul(sep, best_n)), feature_key(MR_SHAPE_BOX, vi),
                                     feature_key(MR_SHAPE_BOX, 128u + best_axis)));
    }
    for (uint32_t vi = 0; vi < 8u; ++vi) {
        const V3 pb = box_vertex(B, vi);
        if (!point_in_inflated_box(pb, A, accepted)) continue;
        const float sep = dot(pb, best_n) - far_a;
        if (sep <= accepted)
            append(out, make_contact(best_n, sep, sub(pb, mul(sep, best_n)), pb,
                                     feature_key(MR_SHAPE_BOX, 128u + best_axis), feature_key(MR_SHAPE_BOX, vi)));
    }
    if (out.count == 0u) {
        const V3 pa = box_support(A, best_n);
        const V3 pb = box_support(B, neg(best_n));
        append(out, make_contact(best_n, dot(sub(pb, pa), best_n), pa, pb, feature_key(MR_SHAPE_BOX, 256u + best_axis),
                                 feature_key(MR_SHAPE_BOX, 256u + best_axis)));
    }
}

MRD_INL void append_surface_contact(RawBatch& out, bool surface_is_a, V3 n, V3 finite_point, float sep,
                                    uint32_t finite_feature, uint32_t surface_feature) {
    const V3 sp = sub(finite_point, mul(sep, n));
    RawContact c = make_contact(neg(n), sep, finite_point, sp, finite_feature, surface_feature);
    if (surface_is_a) c = swapped(c);
    append(out, c);
}

// One query point of a finite shape against the heightfield: the triangle
// under x, its unit normal n, the signed distance d of the shape surface
// (x minus rho along n) and that surface point. d is NaN when x lies outside
// the representable grid, so no threshold accepts it.
MRD_INL mr::HfHit heightfield_query(const Env& e, V3 x, float rho) {
    mr::HfHit h;
    h.d = __builtin_nanf("");
    const MRHeightfieldC& hf = e.a->heightfield;
    const float s = hf.origin_spacing.z;
    const float fx = (x.x - hf.origin_spacing.x) / s;
    const float fy = (x.y - hf.origin_spacing.y) / s;
    if (!isfin(fx) || !isfin(fy)) return h;
    const int ix = (int)__builtin_floorf(clampd(fx, 0.0f, (float)(hf.nx - 2u)));
    const int iy = (int)__builtin_floorf(clampd(fy, 0.0f, (float)(hf.ny - 2u)));
    const float tx = fx - (float)ix;
    const float ty = fy - (float)iy;
    const uint32_t nx = hf.nx;
    const float* H = e.a->heights;
    const float h00 = H[(uint32_t)iy * nx + (uint32_t)ix];
    const float h10 = H[(uint32_t)iy * nx + (uint32_t)ix + 1u];
    const float h01 = H[((uint32_t)iy + 1u) * nx + (uint32_t)ix];
    const float h11 = H[((uint32_t)iy + 1u) * nx + (uint32_t)ix + 1u];
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
    const V3 p = sub(x, mul(rho, n));
    const uint32_t cell = (uint32_t)iy * (nx - 1u) + (uint32_t)ix;
    h.n[0] = n.x; h.n[1] = n.y; h.n[2] = n.z;
    h.d = d;
    h.p[0] = p.x; h.p[1] = p.y; h.p[2] = p.z;
    h.feature = feature_key(MR_SHAPE_HEIGHTFIELD, 2u * cell + tri);
    return h;
}

// Query points of a finite shape against the heightfield (Collision.metal
// meshContacts on a heightfield): a sphere's center; a capsule's two segment
// ends and its center; a box's eight vertices.
MRD_INL uint32_t hf_query_count(uint32_t type) {
    return type == MR_SHAPE_SPHERE ? 1u : type == MR_SHAPE_CAPSULE ? 3u : type == MR_SHAPE_BOX ? 8u : 0u;
}

MRD_INL V3 hf_query_point(const Collider& fs, uint32_t q) {
    if (fs.type == MR_SHAPE_CAPSULE) return q == 0u ? fs.p0 : q == 1u ? fs.p1 : fs.center;
    if (fs.type == MR_SHAPE_BOX) return box_vertex(fs, q);
    return fs.center;
}

MRD_INL float hf_query_radius(const Collider& fs) { return fs.type == MR_SHAPE_BOX ? 0.0f : fs.radius; }

// ------------------------------------------- convex hulls and triangle meshes
//
// Shapes with a support-mapped core (a point set plus a radius): sphere
// (centre), capsule (segment ends), box (8 vertices), convex hull (its
// vertices, scaled) and a mesh triangle (3 vertices). Separated cores get
// GJK's distance and witness points (Voronoi-region simplex reduction, Ericson
// 5.1.5); overlapping cores, or cores closer than float resolution at their
// coordinates, get the Minkowski portal refinement (XenoCollide) along the
// ray from the centre difference through the origin. The contact patch then
// takes both cores' features within a slab of their support planes along the
// normal: two faces give the corners of their intersection polygon (A's face
// clipped by B's, Sutherland-Hodgman in the plane normal to the contact, with
// B's face normal replacing the query normal when within ~8 degrees), an edge
// on a face the clipped segment's ends, anything else the witness pair; the
// deepest MR_RAW_CONTACTS_PER_PAIR points are kept. Per lane: the cores,
// simplices and patch polygons live in private memory, so this code is
// compiled into its own kernel instances, used only by worlds with convex or
// mesh shapes (Collision.metal:2298-3740 is the reference's GJK / EPA / MPR /
// support-mapped patch; the algorithm here is the published one named above).

struct Core {
    uint32_t type;  // MR_SHAPE_* of the collider (MR_SHAPE_TRIANGLE_MESH for a triangle)
    uint32_t n;     // points
    float radius;
    V3 center;
    Q4 rotation;
    V3 scale;       // convex: vertex scale; box: half extents
    V3 p0, p1, p2;  // capsule endpoints / triangle vertices
    const mr_float4* verts;  // convex: the geometry's vertices
};

MRD_INL V3 core_point(const Core& c, uint32_t i) {
    if (c.type == MR_SHAPE_SPHERE) return c.center;
    if (c.type == MR_SHAPE_CAPSULE) return i == 0u ? c.p0 : c.p1;
    if (c.type == MR_SHAPE_TRIANGLE_MESH) return i == 0u ? c.p0 : (i == 1u ? c.p1 : c.p2);
    if (c.type == MR_SHAPE_BOX) {
        const V3 he = c.scale;
        const V3 l{(i & 1u) == 0u ? -he.x : he.x, (i & 2u) == 0u ? -he.y : he.y, (i & 4u) == 0u ? -he.z : he.z};
        return add(c.center, qrot(c.rotation, l));
    }
    const mr_float4 v = c.verts[i];
    return add(c.center, qrot(c.rotation, V3{c.scale.x * v.x, c.scale.y * v.y, c.scale.z * v.z}));
}

// Index of the point with the largest projection on d (the lowest on ties).
MRD_INL uint32_t core_support(const Core& c, V3 d) {
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

MRD_INL Core make_core(const Env& e, uint32_t s, const Collider& col) {
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
        const MRGeometryC& g = e.a->geometries[e.shapes[s].geometry];
        c.n = g.vertex_count;
        c.verts = e.a->geometry_vertices + g.vertex_offset;
    }
    return c;
}

struct Mv {
    V3 w, a, b;
    uint32_t ia, ib;
};

MRD_INL Mv minkowski_support(const Core& A, const Core& B, V3 d) {
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
// barycentric weights of the kept points.
MRD_INL V3 simplex_closest(Mv* s, uint32_t& n, float* lam) {
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
MRD_INL bool tetra_contains_origin(const Mv* s) {
    for (int f = 0; f < 4; ++f) {
        const V3 a = s[(f + 1) % 4].w, b = s[(f + 2) % 4].w, c = s[(f + 3) % 4].w, d = s[f].w;
        const V3 nrm = cross(sub(b, a), sub(c, a));
        const float sd = dot(nrm, sub(d, a)), so = dot(nrm, neg(a));
        if (sd * so < 0.0f) return false;
    }
    return true;
}

struct ConvexHit {
    uint32_t status;  // 0 ok, 1 no result
    bool overlap;
    V3 n;             // A -> B
    float sep;        // core separation (negative: depth)
    V3 pa, pb;        // core witness points
};

MRD_INL ConvexHit gjk(const Core& A, const Core& B) {
    ConvexHit r;
    r.status = 0u;
    r.overlap = false;
    r.n = r.pa = r.pb = V3{0.0f, 0.0f, 0.0f};
    r.sep = 0.0f;
    Mv s[4];
    float lam[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    uint32_t n = 0u;
    V3 v = sub(A.center, B.center);
    if (!(dot(v, v) > kTiny)) v = V3{1.0f, 0.0f, 0.0f};
    for (uint32_t it = 0; it < 64u; ++it) {
        const Mv m = minkowski_support(A, B, neg(v));
        const float vv = dot(v, v);
        if (n > 0u && vv - dot(v, m.w) <= 1.0e-6f * vv) break;
        bool repeat = false;
        for (uint32_t k = 0; k < n; ++k) repeat = repeat || (s[k].ia == m.ia && s[k].ib == m.ib);
        if (repeat) break;
        s[n++] = m;
        if (n == 4u) {
            if (tetra_contains_origin(s)) {
                r.overlap = true;
                return r;
            }
            float best = __builtin_inff();
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

MRD_INL ConvexHit mpr(const Core& A, const Core& B) {
    ConvexHit r;
    r.status = 1u;
    r.overlap = true;
    r.n = r.pa = r.pb = V3{0.0f, 0.0f, 0.0f};
    r.sep = 0.0f;
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
        const V3 d = normalize(sub(v1.w, v0.w));
        r.status = 0u;
        r.n = d;
        r.sep = -dot(v1.w, d);
        r.pa = v1.a;
        r.pb = sub(v1.a, mul(-r.sep, d));
        return r;
    }
    Mv v2 = minkowski_support(A, B, n);
    if (dot(v2.w, n) < 0.0f) return r;
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
        if (delta <= 1.0e-5f * fmaxd(1.0f, fabs_(depth)) || it == 31u) {
            const float dist = dot(v1.w, n);
            const V3 p = mul(dist, n);
            const V3 e1 = sub(v2.w, v1.w), e2 = sub(v3.w, v1.w), ep = sub(p, v1.w);
            const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
            const float d20 = dot(ep, e1), d21 = dot(ep, e2);
            const float den = d00 * d11 - d01 * d01;
            float b2 = 0.0f, b3 = 0.0f;
            if (fabs_(den) > kTiny) {
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

MRD_INL void plane_basis(V3 n, V3& t1, V3& t2) {
    const V3 an = vabs(n);
    V3 ref;
    if (an.x <= an.y && an.x <= an.z) ref = V3{1.0f, 0.0f, 0.0f};
    else if (an.y <= an.z) ref = V3{0.0f, 1.0f, 0.0f};
    else ref = V3{0.0f, 0.0f, 1.0f};
    t1 = normalize(cross(n, ref));
    t2 = cross(n, t1);
}

constexpr uint32_t kFeatureCap = 16u;
constexpr uint32_t kPatchCap = 2u * kFeatureCap;

// Points of core c within `slab` of its support plane along d, in index
// order (at most kFeatureCap), and the plane's offset.
MRD_INL uint32_t core_feature(const Core& c, V3 d, float slab, uint32_t* idx, float& top) {
    top = -__builtin_inff();
    for (uint32_t i = 0; i < c.n; ++i) top = fmaxd(top, dot(d, core_point(c, i)));
    uint32_t k = 0u;
    for (uint32_t i = 0; i < c.n && k < kFeatureCap; ++i)
        if (dot(d, core_point(c, i)) >= top - slab) idx[k++] = i;
    return k;
}

MRD_INL float core_extent(const Core& c) {
    float e = 0.0f;
    for (uint32_t i = 0; i < c.n; ++i) e = fmaxd(e, length(sub(core_point(c, i), c.center)));
    return e;
}

// Convex hull of planar points (gift wrapping from the lowest index of the
// leftmost point), counter-clockwise indices into pts; returns its size.
MRD_INL uint32_t hull2d(const float (*pts)[2], uint32_t n, uint32_t* out) {
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
// the kept parameter interval is [t0, t1] (empty when t0 > t1).
MRD_INL void clip_segment(const float (*pts)[2], const uint32_t* poly, uint32_t m, float x0, float y0, float x1,
                          float y1, float& t0, float& t1) {
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

// Keeps the deepest MR_RAW_CONTACTS_PER_PAIR contacts (replacing the
// shallowest, the earlier kept on ties).
MRD_INL void append_deepest(RawBatch& out, const RawContact& c) {
    if (out.count < MR_RAW_CONTACTS_PER_PAIR) {
        out.c[out.count++] = c;
        return;
    }
    uint32_t worst = 0u;
    for (uint32_t k = 1; k < out.count; ++k)
        if (out.c[k].separation > out.c[worst].separation) worst = k;
    if (c.separation < out.c[worst].separation) out.c[worst] = c;
}

MRD_INL uint32_t core_feature_key(const Core& c, uint32_t local, uint32_t tri_feature) {
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

MRD_INL void convex_contacts(const Core& A, const Core& B, uint32_t tri_a, uint32_t tri_b, float accepted,
                             RawBatch& out) {
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
    const uint32_t face_a = core_feature_key(A, 128u + fa[0], tri_a), face_b = core_feature_key(B, 128u + fb[0], tri_b);
    if (na == 1u || nb == 1u || (na == 2u && nb == 2u)) {
        const uint32_t feat_a = na == 1u ? core_feature_key(A, fa[0], tri_a) : face_a;
        const uint32_t feat_b = nb == 1u ? core_feature_key(B, fb[0], tri_b) : face_b;
        append_deepest(out, make_contact(n, h.sep - A.radius - B.radius, add(h.pa, mul(A.radius, n)),
                                         sub(h.pb, mul(B.radius, n)), feat_a, feat_b));
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
        V3 fnrm{0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < mb; ++k) fnrm = add(fnrm, cross(wb[hull_b[k]], wb[hull_b[(k + 1u) % mb]]));
        const float fl = length(fnrm);
        if (fl > kTiny && dot(fnrm, n) >= 0.99f * fl) {
            n = mul(1.0f / fl, fnrm);
            plane_b = __builtin_inff();
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
                const PatchVertex P = poly[j];
                const PatchVertex Q = poly[(j + 1u) % m];
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
            append_deepest(out, make_contact(n, s - A.radius - B.radius, add(poly[j].p, mul(A.radius, n)),
                                             sub(add(poly[j].p, mul(s, n)), mul(B.radius, n)), poly[j].feat_a,
                                             poly[j].feat_b));
        }
    } else if (na == 2u && mb >= 3u) {
        float t0, tt1;
        clip_segment(pb2, hull_b, mb, pa2[0][0], pa2[0][1], pa2[1][0], pa2[1][1], t0, tt1);
        if (t0 <= tt1)
            for (int ei = 0; ei < 2; ++ei) {
                const float t = ei == 0 ? t0 : tt1;
                const V3 p = add(wa[0], mul(t, sub(wa[1], wa[0])));
                const float s = plane_b - dot(n, p);
                append_deepest(out, make_contact(n, s - A.radius - B.radius, add(p, mul(A.radius, n)),
                                                 sub(add(p, mul(s, n)), mul(B.radius, n)),
                                                 core_feature_key(A, 64u + fa[ei], tri_a), face_b));
            }
    } else if (nb == 2u && ma >= 3u) {
        float t0, tt1;
        clip_segment(pa2, hull_a, ma, pb2[0][0], pb2[0][1], pb2[1][0], pb2[1][1], t0, tt1);
        if (t0 <= tt1)
            for (int ei = 0; ei < 2; ++ei) {
                const float t = ei == 0 ? t0 : tt1;
                const V3 p = add(wb[0], mul(t, sub(wb[1], wb[0])));
                const float s = dot(n, p) - top_a;
                append_deepest(out, make_contact(n, s - A.radius - B.radius, add(sub(p, mul(s, n)), mul(A.radius, n)),
                                                 sub(p, mul(B.radius, n)), face_a,
                                                 core_feature_key(B, 64u + fb[ei], tri_b)));
            }
    }
    if (out.count == before)
        append_deepest(out, make_contact(n, h.sep - A.radius - B.radius, add(h.pa, mul(A.radius, n)),
                                         sub(h.pb, mul(B.radius, n)), face_a, face_b));
}

// A convex hull against a plane: every vertex within `accepted` (deepest 8).
MRD_INL void convex_plane_contacts(const Core& hull, const Collider& plane, bool plane_is_a, float accepted,
                                   RawBatch& out) {
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

constexpr uint32_t kBvhStack = 48u;

// A static mesh against a finite collider: BVH4 traversal with the finite
// collider's bounds (padded by the accepted distance) in the mesh's unscaled
// local frame, then every overlapping triangle as a 3-point core. One-sided
// meshes keep the contacts whose normal (finite -> mesh) opposes the
// triangle's front normal (counter-clockwise vertices).
MRD_INL void mesh_contacts(const Env& e, uint32_t mesh_shape, const Collider& mesh, const Core& fin,
                           const Collider& fin_col, bool mesh_is_a, float accepted, RawBatch& out) {
    const MRShapeC& ms = e.shapes[mesh_shape];
    const MRGeometryC g = e.a->geometries[ms.geometry];
    const mr_float4* vt = e.a->geometry_vertices + g.vertex_offset;
    const MRTriangleC* tris = e.a->triangles + g.triangle_offset;
    const MRBvhNodeC* nodes = e.a->bvh_nodes + g.node_offset;
    const bool two_sided = (ms.flags & MR_SHAPE_FLAG_MESH_TWO_SIDED) != 0u;
    const V3 scale = mesh.half_extents;
    V3 lo{__builtin_inff(), __builtin_inff(), __builtin_inff()};
    V3 hi{-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint32_t k = 0; k < 8u; ++k) {
        const V3 w{(k & 1u) ? fin_col.upper.x + accepted : fin_col.lower.x - accepted,
                   (k & 2u) ? fin_col.upper.y + accepted : fin_col.lower.y - accepted,
                   (k & 4u) ? fin_col.upper.z + accepted : fin_col.lower.z - accepted};
        const V3 l = qrot_inv(mesh.rotation, sub(w, mesh.center));
        const V3 u{l.x / scale.x, l.y / scale.y, l.z / scale.z};
        lo = vmin(lo, u);
        hi = vmax(hi, u);
    }
    uint32_t stack[kBvhStack];
    uint32_t top = 0u;
    stack[top++] = 0u;
    while (top > 0u) {
        const MRBvhNodeC& nd = nodes[stack[--top]];
        const float lx[4] = {nd.lower_x.x, nd.lower_x.y, nd.lower_x.z, nd.lower_x.w};
        const float ly[4] = {nd.lower_y.x, nd.lower_y.y, nd.lower_y.z, nd.lower_y.w};
        const float lz[4] = {nd.lower_z.x, nd.lower_z.y, nd.lower_z.z, nd.lower_z.w};
        const float ux[4] = {nd.upper_x.x, nd.upper_x.y, nd.upper_x.z, nd.upper_x.w};
        const float uy[4] = {nd.upper_y.x, nd.upper_y.y, nd.upper_y.z, nd.upper_y.w};
        const float uz[4] = {nd.upper_z.x, nd.upper_z.y, nd.upper_z.z, nd.upper_z.w};
        uint32_t child[4], count[4];
        for (int c = 0; c < 4; ++c) {
            child[c] = nd.child[c];
            count[c] = nd.count[c];
        }
        for (int c = 3; c >= 0; --c) {  // pushed in reverse: children leave the stack in index order
            if (child[c] == MR_INVALID_INDEX || count[c] != 0u) continue;
            if (lx[c] <= hi.x && ux[c] >= lo.x && ly[c] <= hi.y && uy[c] >= lo.y && lz[c] <= hi.z && uz[c] >= lo.z &&
                top < kBvhStack)
                stack[top++] = child[c];
        }
        for (uint32_t c = 0; c < 4u; ++c) {
            if (child[c] == MR_INVALID_INDEX || count[c] == 0u) continue;
            if (!(lx[c] <= hi.x && ux[c] >= lo.x && ly[c] <= hi.y && uy[c] >= lo.y && lz[c] <= hi.z && uz[c] >= lo.z))
                continue;
            for (uint32_t t = child[c]; t < child[c] + count[c]; ++t) {
                const MRTriangleC tr = tris[t];
                Core tc;
                tc.type = MR_SHAPE_TRIANGLE_MESH;
                tc.n = 3u;
                tc.radius = 0.0f;
                tc.rotation = mesh.rotation;
                tc.scale = scale;
                tc.verts = nullptr;
                V3 p[3];
                for (int k = 0; k < 3; ++k) {
                    const mr_float4 v = vt[tr.vertices[k]];
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

// hf_first: index of this pair's first precomputed heightfield query in e.hfq,
// or ~0 to evaluate the queries here. Convex and mesh pairs are generated by
// the kGeo instances only.
template <bool kGeo = false>
MRD_INL void generate_contacts(const Env& e, uint32_t ia, uint32_t ib, uint32_t cls, const Collider& A,
                               const Collider& B, RawBatch& out, uint32_t hf_first) {
    out.count = 0;
    const float accepted = (A.contact_offset + B.contact_offset) + pair_query_padding(A, B);
    if (cls == MR_PAIR_SPHERE_SPHERE) {
        const V3 d = sub(B.center, A.center);
        const float dist = length(d);
        const V3 n = dist > kTiny ? mul(1.0f / dist, d) : coincident_normal(ia, ib);
        const float sep = dist - A.radius - B.radius;
        if (sep > accepted) return;
        append(out, make_contact(n, sep, add(A.center, mul(A.radius, n)), sub(B.center, mul(B.radius, n)),
                                 feature_key(MR_SHAPE_SPHERE, 0u), feature_key(MR_SHAPE_SPHERE, 0u)));
    } else if (cls == MR_PAIR_SPHERE_CAPSULE) {
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
            RawContact c = make_contact(n, sep, add(sp.center, mul(sp.radius, n)), sub(cl.point, mul(cp.radius, n)),
                                        feature_key(MR_SHAPE_SPHERE, 0u),
                                        feature_key(MR_SHAPE_CAPSULE, capsule_feature(cl.t)));
            if (!sphere_a) c = swapped(c);
            append(out, c);
        }
    } else if (cls == MR_PAIR_CAPSULE_CAPSULE) {
        const SegPair cl = closest_segments(A.p0, A.p1, B.p0, B.p1);
        const V3 d = sub(cl.pb, cl.pa);
        const float dist = length(d);
        const V3 n = dist > kTiny ? mul(1.0f / dist, d) : stable_segment_normal(sub(A.p1, A.p0), sub(B.p1, B.p0), ia, ib);
        const float sep = dist - A.radius - B.radius;
        if (sep <= accepted)
            append(out, make_contact(n, sep, add(cl.pa, mul(A.radius, n)), sub(cl.pb, mul(B.radius, n)),
                                     feature_key(MR_SHAPE_CAPSULE, capsule_feature(cl.ta)),
                                     feature_key(MR_SHAPE_CAPSULE, capsule_feature(cl.tb))));
    } else if (cls == MR_PAIR_SPHERE_BOX) {
        const bool sphere_a = A.type == MR_SHAPE_SPHERE;
        const Collider& sp = sphere_a ? A : B;
        const Collider& bx = sphere_a ? B : A;
        V3 n, bp;
        float sep;
        uint32_t feat;
        sphere_box(sp, bx, n, bp, sep, feat);
        if (sep <= accepted) {
            RawContact c = make_contact(n, sep, add(sp.center, mul(sp.radius, n)), bp, feature_key(MR_SHAPE_SPHERE, 0u),
                                        feature_key(MR_SHAPE_BOX, feat));
            if (!sphere_a) c = swapped(c);
            append(out, c);
        }
    } else if (cls == MR_PAIR_CAPSULE_BOX) {
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
    } else if (cls == MR_PAIR_BOX_BOX) {
        box_box(ia, ib, A, B, accepted, out);
    } else if (cls == MR_PAIR_SPHERE_PLANE || cls == MR_PAIR_CAPSULE_PLANE || cls == MR_PAIR_BOX_PLANE) {
        const bool plane_a = A.type == MR_SHAPE_PLANE;
        const Collider& pl = plane_a ? A : B;
        const Collider& fs = plane_a ? B : A;
        const uint32_t pf = feature_key(MR_SHAPE_PLANE, 0u);
        if (cls == MR_PAIR_SPHERE_PLANE) {
            const V3 surf = sub(fs.center, mul(fs.radius, pl.plane_normal));
            const float sep = dot(pl.plane_normal, sub(surf, pl.center));
            if (sep <= accepted)
                append_surface_contact(out, plane_a, pl.plane_normal, surf, sep, feature_key(MR_SHAPE_SPHERE, 0u), pf);
        } else if (cls == MR_PAIR_CAPSULE_PLANE) {
            for (uint32_t en = 0; en < 2u; ++en) {
                const V3 ep = en == 0u ? fs.p0 : fs.p1;
                const V3 surf = sub(ep, mul(fs.radius, pl.plane_normal));
                const float sep = dot(pl.plane_normal, sub(surf, pl.center));
                if (sep <= accepted)
                    append_surface_contact(out, plane_a, pl.plane_normal, surf, sep, feature_key(MR_SHAPE_CAPSULE, en),
                                           pf);
            }
        } else {
            for (uint32_t vi = 0; vi < 8u; ++vi) {
                const V3 wv = box_vertex(fs, vi);
                const float sep = dot(pl.plane_normal, sub(wv, pl.center));
                if (sep <= accepted)
                    append_surface_contact(out, plane_a, pl.plane_normal, wv, sep, feature_key(MR_SHAPE_BOX, vi), pf);
            }
        }
    } else if (cls == MR_PAIR_HEIGHTFIELD) {
        const bool hf_a = A.type == MR_SHAPE_HEIGHTFIELD;
        const Collider& fs = hf_a ? B : A;
        const uint32_t nq = hf_query_count(fs.type);
        for (uint32_t q = 0; q < nq; ++q) {
            const mr::HfHit h = hf_first != 0xffffffffu ? e.hfq[hf_first + q]
                                                        : heightfield_query(e, hf_query_point(fs, q), hf_query_radius(fs));
            if (h.d <= accepted)
                append_surface_contact(out, hf_a, V3{h.n[0], h.n[1], h.n[2]}, V3{h.p[0], h.p[1], h.p[2]}, h.d,
                                       feature_key(fs.type, q), h.feature);
        }
    } else if (kGeo && cls == MR_PAIR_CONVEX) {
        if (A.type == MR_SHAPE_PLANE || B.type == MR_SHAPE_PLANE) {
            const bool plane_a = A.type == MR_SHAPE_PLANE;
            convex_plane_contacts(make_core(e, plane_a ? ib : ia, plane_a ? B : A), plane_a ? A : B, plane_a, accepted,
                                  out);
        } else {
            convex_contacts(make_core(e, ia, A), make_core(e, ib, B), 0u, 0u, accepted, out);
        }
    } else if (kGeo && cls == MR_PAIR_MESH) {
        const bool mesh_a = A.type == MR_SHAPE_TRIANGLE_MESH;
        mesh_contacts(e, mesh_a ? ia : ib, mesh_a ? A : B, make_core(e, mesh_a ? ib : ia, mesh_a ? B : A),
                      mesh_a ? B : A, mesh_a, accepted, out);
    }
}


// ------------------------------------------------ continuous collision
//
// The hybrid mode splits a microstep at the earliest time of impact (TOI) of
// the CCD-enabled compiled pairs (mr_world_resolve_ccd and the event select /
// finalize kernels, Collision.metal:6707-6950, 7157-7330, 7657-7900). Over the
// remaining time T of the microstep every body moves rigidly: links with the
// COM and angular velocity of the current kinematics, dynamic scene bodies
// with their velocity plus T * gravity, kinematic bodies with theirs, static
// bodies not at all; rotations follow exp(t w) at a constant rate. A pair whose
// swept bounds overlap and whose colliders start farther apart than their
// contact distance (the sum of the contact offsets; within it the segment's
// speculative contacts hold the pair at any normal speed) gets a TOI: closed
// form for two spheres (centres moving linearly),
// conservative advancement otherwise, stepping t by distance / (closing-speed
// bound), at least the minimum advance, and reporting the first t at which the
// pair is within the tolerance. A step never closes the gap past contact, so
// the pair is touching there without penetrating (up to speed x minimum
// advance). The closing-speed bound is |v_a - v_b| + |w_a| r_a + |w_b| r_b, r
// the largest distance from a body's COM to a point of its shape (exact for
// the constant-rate rotation); against a static plane only the normal
// component of the velocity counts. A pair that does not resolve within the iteration
// cap is safe when both bodies move less than their contact offset plus the
// tolerance (the speculative contacts cover it), and fails the step otherwise.

struct CcdMotion {
    V3 p0;
    Q4 r0;
    V3 v, w;
};

MRD_INL CcdMotion ccd_motion(const Env& e, uint32_t b, float T) {
    const V3 zero{0.0f, 0.0f, 0.0f};
    CcdMotion m;
    if (b < e.nlinks) {
        m.p0 = e.pos[b];
        m.r0 = e.rot[b];
        m.v = e.vl[b];
        m.w = e.w[b];
        return m;
    }
    const MRBodyStateC& s = e.scene[b - e.nlinks];
    m.p0 = f3(s.position);
    m.r0 = q4(s.orientation);
    const uint32_t mt = e.topo->motion[b];
    if (mt == MR_MOTION_DYNAMIC) {
        m.v = add(f3(s.linear_velocity_inv_mass), mul(T, e.gravity));
        m.w = f3(s.angular_velocity);
    } else if (mt == MR_MOTION_KINEMATIC) {
        m.v = f3(s.linear_velocity_inv_mass);
        m.w = f3(s.angular_velocity);
    } else {
        m.v = zero;
        m.w = zero;
    }
    return m;
}

template <bool kGeo>
MRD_INL bool ccd_collider(const Env& e, uint32_t s, const CcdMotion& m, float t, Collider& c) {
    const V3 p = add(m.p0, mul(t, m.v));
    Q4 r;
    if (!normalized_quat(qmul(rotation_increment(mul(t, m.w)), m.r0), r)) return false;
    return project_collider_pose<kGeo>(e, s, p, r, c);
}

// Largest distance from the body's COM to a point of shape s; planes and
// heightfields are unbounded.
MRD_INL float ccd_reach(const MRShapeC& sh) {
    const float offset = length(f3(sh.local_position));
    if (sh.shape_type == MR_SHAPE_SPHERE) return offset + sh.dimensions.x;
    if (sh.shape_type == MR_SHAPE_CAPSULE) return offset + sh.dimensions.x + sh.dimensions.y;
    if (sh.shape_type == MR_SHAPE_BOX) return offset + length(f3(sh.dimensions));
    return __builtin_inff();
}

// Distance bound of a body's shape point over the remaining time T.
MRD_INL float ccd_motion_bound(const CcdMotion& m, const MRShapeC& sh, float T) {
    const float wl = length(m.w), vl = length(m.v);
    if (!(wl > 0.0f) && !(vl > 0.0f)) return 0.0f;
    const float reach = wl > 0.0f ? ccd_reach(sh) : 0.0f;
    return T * (vl + wl * reach);
}

// Separation of two colliders at any distance (negative when they overlap):
// the analytic pair functions without the contact-distance cut-off. For two
// boxes the largest separation over the 15 separating axes, a lower bound of
// their distance. Heightfields have no such query.
MRD_INL float box_box_separation(const Collider& A, const Collider& B) {
    const V3 d = sub(B.center, A.center);
    float best = -__builtin_inff();
    for (int i = 0; i < 3; ++i) {
        const V3 ax = box_axis(A, i);
        best = fmaxd(best, fabs_(dot(d, ax)) - comp(A.half_extents, i) - box_projection_radius(B, ax));
    }
    for (int j = 0; j < 3; ++j) {
        const V3 ax = box_axis(B, j);
        best = fmaxd(best, fabs_(dot(d, ax)) - box_projection_radius(A, ax) - comp(B.half_extents, j));
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const V3 c = cross(box_axis(A, i), box_axis(B, j));
            const float l = length(c);
            if (!(l > 1.0e-6f)) continue;
            const V3 ax = mul(1.0f / l, c);
            best = fmaxd(best, fabs_(dot(d, ax)) - box_projection_radius(A, ax) - box_projection_radius(B, ax));
        }
    return best;
}

MRD_INL bool ccd_separation(uint32_t ia, uint32_t ib, uint32_t cls, const Collider& A, const Collider& B, float& d) {
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
            d = __builtin_inff();
            for (uint32_t vi = 0; vi < 8u; ++vi) d = fmind(d, dot(pl.plane_normal, sub(box_vertex(fs, vi), pl.center)));
        }
    } else {
        return false;
    }
    return isfin(d);
}

enum : uint32_t { kCcdNone = 0u, kCcdImpact = 1u, kCcdUnsafe = 2u };

// TOI query of compiled pair k over the remaining time T. Returns kCcdImpact
// with toi (seconds), kCcdUnsafe for an unresolved pair the speculative
// contacts do not cover, kCcdNone otherwise (no CCD flag, apart over the whole
// interval, or touching already: the discrete contacts handle that).
template <bool kGeo>
MRD_INL uint32_t ccd_pair(const Env& e, uint32_t k, float T, float& toi) {
    const MRWorldStepConfigC& cfg = e.a->cfg;
    const MRCompiledPairC& pr = e.pairs[k];
    const MRShapeC& sa = e.shapes[pr.collider_a];
    const MRShapeC& sb = e.shapes[pr.collider_b];
    if (((sa.flags | sb.flags) & MR_SHAPE_FLAG_ENABLE_CCD) == 0u ||
        ((sa.flags | sb.flags) & MR_SHAPE_FLAG_SIMULATION_DISABLED) != 0u)
        return kCcdNone;
    const CcdMotion ma = ccd_motion(e, sa.body, T), mb = ccd_motion(e, sb.body, T);
    const float tol = cfg.ccd_distance_tolerance;
    const float motion_a = ccd_motion_bound(ma, sa, T), motion_b = ccd_motion_bound(mb, sb, T);
    const bool speculative_safe = motion_a <= sa.contact_rest_radius.x + tol && motion_b <= sb.contact_rest_radius.x + tol;
    Collider a0, b0, a1, b1;
    if (!ccd_collider<kGeo>(e, pr.collider_a, ma, 0.0f, a0) || !ccd_collider<kGeo>(e, pr.collider_b, mb, 0.0f, b0) ||
        !ccd_collider<kGeo>(e, pr.collider_a, ma, T, a1) || !ccd_collider<kGeo>(e, pr.collider_b, mb, T, b1))
        return speculative_safe ? kCcdNone : kCcdUnsafe;
    // swept bounds: the union of the start and end bounds
    Collider sa_sw = a0, sb_sw = b0;
    sa_sw.lower = vmin(a0.lower, a1.lower);
    sa_sw.upper = vmax(a0.upper, a1.upper);
    sb_sw.lower = vmin(b0.lower, b1.lower);
    sb_sw.upper = vmax(b0.upper, b1.upper);
    if (!pair_may_overlap(sa_sw, sb_sw)) return kCcdNone;
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
            const float alpha = (-bb - sqrt_rn(disc)) / (2.0f * aa);
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
        total = T * (fabs_(dot(fm.v, n)) + wl * reach);
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
        if (!ccd_collider<kGeo>(e, pr.collider_a, ma, alpha * T, a) || !ccd_collider<kGeo>(e, pr.collider_b, mb, alpha * T, b) ||
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

// --------------------------------------------------------------- manifold

MRD_INL unsigned long long cand_feature(const Candidate& c) {
    return ((unsigned long long)c.point.feature_and_life[0] << 32u) | (unsigned long long)c.point.feature_and_life[1];
}

MRD_INL bool cand_less(const Candidate& l, const Candidate& r) {
    if (l.separation != r.separation) return l.separation < r.separation;
    return cand_feature(l) < cand_feature(r);
}

MRD_INL V3 stable_tangent(V3 n) {
    const V3 a = vabs(n);
    V3 ref;
    if (a.x <= a.y && a.x <= a.z) ref = V3{1.0f, 0.0f, 0.0f};
    else if (a.y <= a.z) ref = V3{0.0f, 1.0f, 0.0f};
    else ref = V3{0.0f, 0.0f, 1.0f};
    return normalize(cross(ref, n));
}

// Candidate order as an index permutation packed in nibbles (count <= 12), so
// sorting and selection move 4-bit indices instead of 84-byte records.
MRD_INL uint32_t perm_get(unsigned long long p, uint32_t i) { return (uint32_t)(p >> (4u * i)) & 0xfu; }
MRD_INL unsigned long long perm_set(unsigned long long p, uint32_t i, uint32_t v) {
    return (p & ~(0xfull << (4u * i))) | ((unsigned long long)v << (4u * i));
}

// Deterministic reduction to at most four points (stable insertion sort by
// separation then feature, greedy spread selection, final sort by feature);
// returns the order of the kept candidates.
MRD_INL unsigned long long reduce_candidates(const Candidate* c, uint32_t& count, V3 n) {
    unsigned long long perm = 0ull;
    for (uint32_t i = 0; i < count; ++i) perm = perm_set(perm, i, i);
    for (uint32_t i = 1; i < count; ++i) {
        const uint32_t val = perm_get(perm, i);
        uint32_t d = i;
        while (d > 0u && cand_less(c[val], c[perm_get(perm, d - 1u)])) {
            perm = perm_set(perm, d, perm_get(perm, d - 1u));
            --d;
        }
        perm = perm_set(perm, d, val);
    }
    if (count > MR_MANIFOLD_POINT_CAPACITY) {
        uint32_t selected = 1u;  // bit i: sorted position i taken
        uint32_t chosen = 0u;    // nibble k: sorted position of the k-th pick
        uint32_t nch = 1u;
        for (uint32_t stage = 1; stage < MR_MANIFOLD_POINT_CAPACITY; ++stage) {
            uint32_t best = MR_INVALID_INDEX;
            float best_score = -1.0f;
            unsigned long long best_feat = ~0ull;
            const V3 w0 = c[perm_get(perm, chosen & 0xfu)].world_point;
            for (uint32_t i = 0; i < count; ++i) {
                if ((selected >> i) & 1u) continue;
                const Candidate& ci = c[perm_get(perm, i)];
                float score;
                if (stage == 1u) {
                    const V3 d = sub(ci.world_point, w0);
                    const V3 t = sub(d, mul(dot(d, n), n));
                    score = dot(t, t);
                } else if (stage == 2u) {
                    const V3 aa = sub(c[perm_get(perm, (chosen >> 4) & 0xfu)].world_point, w0);
                    const V3 bb = sub(ci.world_point, w0);
                    score = fabs_(dot(cross(aa, bb), n));
                } else {
                    score = __builtin_inff();
                    for (uint32_t q = 0; q < nch; ++q) {
                        const V3 d = sub(ci.world_point, c[perm_get(perm, (chosen >> (4u * q)) & 0xfu)].world_point);
                        const V3 t = sub(d, mul(dot(d, n), n));
                        score = fmind(score, dot(t, t));
                    }
                }
                const unsigned long long feat = cand_feature(ci);
                if (score > best_score + 1.0e-12f || (fabs_(score - best_score) <= 1.0e-12f && feat < best_feat)) {
                    best = i;
                    best_score = score;
                    best_feat = feat;
                }
            }
            if (best != MR_INVALID_INDEX) {
                chosen |= best << (4u * nch);
                ++nch;
                selected |= 1u << best;
            }
        }
        unsigned long long red = 0ull;
        for (uint32_t k = 0; k < nch; ++k) red = perm_set(red, k, perm_get(perm, (chosen >> (4u * k)) & 0xfu));
        perm = red;
        count = nch;
    }
    for (uint32_t i = 1; i < count; ++i) {
        const uint32_t val = perm_get(perm, i);
        const unsigned long long f = cand_feature(c[val]);
        uint32_t d = i;
        while (d > 0u && f < cand_feature(c[perm_get(perm, d - 1u)])) {
            perm = perm_set(perm, d, perm_get(perm, d - 1u));
            --d;
        }
        perm = perm_set(perm, d, val);
    }
    return perm;
}

MRD_INL mr_float4 anchor_from_world(V3 bpos, Q4 rot, V3 x) {
    const V3 l = qrot_inv(rot, sub(x, bpos));
    return mr_float4{l.x, l.y, l.z, 0.0f};
}

MRD_INL bool build_manifold(const Env& e, const MRCompiledPairC& pair, const RawBatch& raw,
                            const MRManifoldHeaderC* old_h, const MRManifoldPointC* old_p, MRManifoldHeaderC& out_h,
                            MRManifoldPointC* out_p, uint32_t& retained, uint32_t& fresh) {
    if (raw.count == 0u) return false;
    const MRShapeC& sa = e.shapes[pair.collider_a];
    const MRShapeC& sb = e.shapes[pair.collider_b];
    V3 pa_pos, pb_pos;
    Q4 pa_q, pb_q;
    body_pose(e, sa.body, pa_pos, pa_q);
    body_pose(e, sb.body, pb_pos, pb_q);
    Q4 ra, rb;
    if (!checked_quat(pa_q, ra) || !checked_quat(pb_q, rb)) return false;
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
    const MRWorldStepConfigC& cfg = e.a->cfg;
    Candidate cand[kMaxCandidates];
    uint32_t nc = 0;
    bool compatible = false;
    V3 transported{0.0f, 0.0f, 0.0f};
    float old_age = 0.0f;
    if (old_h != nullptr) {
        const MRManifoldHeaderC oh = *old_h;
        const bool gen = oh.generations_and_flags[0] == sa.slot_generation &&
            oh.generations_and_flags[1] == sb.slot_generation && oh.pair_and_count[3] <= MR_MANIFOLD_POINT_CAPACITY;
        if (gen) {
            const V3 on = normalize(qrot(ra, f3(oh.normal_and_age)));
            compatible = finite3(on) && dot(on, n) >= cfg.manifold_normal_cosine;
            if (compatible) {
                const V3 ot = qrot(ra, f3(oh.tangent_and_metric));
                transported = sub(ot, mul(dot(ot, n), n));
                old_age = oh.normal_and_age.w;
                for (uint32_t p = 0; p < oh.pair_and_count[3]; ++p) {
                    const MRManifoldPointC op = old_p[p];
                    const V3 pa = add(pa_pos, qrot(ra, f3(op.local_anchor_a)));
                    const V3 pb = add(pb_pos, qrot(rb, f3(op.local_anchor_b)));
                    const V3 d = sub(pb, pa);
                    const float sep = dot(d, n);
                    const V3 tan = sub(d, mul(sep, n));
                    const float drift = length(tan);
                    if (sep > cfg.manifold_breaking_separation || drift > cfg.manifold_breaking_tangential) continue;
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
    const float merge_sq = cfg.manifold_merge_distance * cfg.manifold_merge_distance;
    for (uint32_t k = 0; k < raw.count; ++k) {
        const RawContact& rc = raw.c[k];
        const mr_float4 la = anchor_from_world(pa_pos, ra, rc.point_a);
        const mr_float4 lb = anchor_from_world(pb_pos, rb, rc.point_b);
        uint32_t match = MR_INVALID_INDEX;
        for (uint32_t i = 0; i < nc; ++i) {
            if (cand[i].point.feature_and_life[0] == rc.feature_a && cand[i].point.feature_and_life[1] == rc.feature_b) {
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
            Candidate z;
            z.point.local_anchor_a = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
            z.point.local_anchor_b = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
            z.point.impulses = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
            z.point.feature_and_life[0] = 0u;
            z.point.feature_and_life[1] = 0u;
            z.point.feature_and_life[2] = 0u;
            z.point.feature_and_life[3] = 0u;
            z.world_point = V3{0.0f, 0.0f, 0.0f};
            z.separation = 0.0f;
            z.drift = 0.0f;
            cand[match] = z;
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
    const unsigned long long order = reduce_candidates(cand, nc, n);
    V3 t = transported;
    if (!compatible || !finite3(t) || dot(t, t) <= 1.0e-12f) t = stable_tangent(n);
    else t = normalize(t);
    out_h.pair_and_count[0] = e.env;
    out_h.pair_and_count[1] = pair.collider_a;
    out_h.pair_and_count[2] = pair.collider_b;
    out_h.pair_and_count[3] = nc;
    out_h.generations_and_flags[0] = sa.slot_generation;
    out_h.generations_and_flags[1] = sb.slot_generation;
    out_h.generations_and_flags[2] = 0u;
    out_h.generations_and_flags[3] = 0u;
    const V3 ln = qrot_inv(ra, n);
    out_h.normal_and_age = mr_float4{ln.x, ln.y, ln.z, compatible ? fmind(old_age + 1.0f, 65535.0f) : 0.0f};
    float metric = 0.0f;
    for (uint32_t i = 0; i < nc; ++i) {
        const Candidate& ci = cand[perm_get(order, i)];
        out_p[i] = ci.point;
        metric = fmaxd(metric, fmaxd(fmaxd(ci.separation, 0.0f), ci.drift));
    }
    for (uint32_t i = nc; i < MR_MANIFOLD_POINT_CAPACITY; ++i) {
        out_p[i].local_anchor_a = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
        out_p[i].local_anchor_b = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
        out_p[i].impulses = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
        out_p[i].feature_and_life[0] = 0u;
        out_p[i].feature_and_life[1] = 0u;
        out_p[i].feature_and_life[2] = 0u;
        out_p[i].feature_and_life[3] = 0u;
    }
    const V3 lt = qrot_inv(ra, t);
    out_h.tangent_and_metric = mr_float4{lt.x, lt.y, lt.z, metric};
    return true;
}

// ------------------------------------------------------------ cone solve

// Normal clamp (and the optional maximum normal impulse), then radial scaling
// of the tangents onto the elliptic cone (MetalWorldContact.metal:482-509).
MRD_INL void project_cone(float lam[3], float mu_u, float mu_v, float max_normal) {
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
        const float s = 1.0f / sqrt_rn(ns);
        lam[1] *= s;
        lam[2] *= s;
    }
}

MRD_INL bool invert3x3(const float m[3][3], float inv[9]) {
    float scale = 0.0f;
    for (int rr = 0; rr < 3; ++rr)
        for (int cc = 0; cc < 3; ++cc) {
            if (!isfin(m[rr][cc])) return false;
            scale = fmaxd(scale, fabs_(m[rr][cc]));
        }
    if (!(scale > kMatrixFloor)) return false;
    float g[3][3];
    for (int rr = 0; rr < 3; ++rr)
        for (int cc = 0; cc < 3; ++cc) {
            g[rr][cc] = 0.5f * (m[rr][cc] + m[cc][rr]) / scale;
            if (rr == cc) g[rr][cc] += kContactRegularization;
        }
    const float c00 = g[1][1] * g[2][2] - g[1][2] * g[2][1];
    const float c01 = g[1][2] * g[2][0] - g[1][0] * g[2][2];
    const float c02 = g[1][0] * g[2][1] - g[1][1] * g[2][0];
    const float det = g[0][0] * c00 + g[0][1] * c01 + g[0][2] * c02;
    if (!(det > kMatrixFloor) || !isfin(det)) return false;
    const float rcp = 1.0f / (det * scale);
    inv[0] = c00 * rcp;
    inv[1] = (g[0][2] * g[2][1] - g[0][1] * g[2][2]) * rcp;
    inv[2] = (g[0][1] * g[1][2] - g[0][2] * g[1][1]) * rcp;
    inv[3] = c01 * rcp;
    inv[4] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) * rcp;
    inv[5] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) * rcp;
    inv[6] = c02 * rcp;
    inv[7] = (g[0][1] * g[2][0] - g[0][0] * g[2][1]) * rcp;
    inv[8] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) * rcp;
    for (int k = 0; k < 9; ++k)
        if (!isfin(inv[k])) return false;
    return true;
}

MRD_INL V3 inv_inertia_mul(const MRBodyStateC& b, V3 x) {
    return V3{dot(f3(b.inverse_inertia_world_row0), x), dot(f3(b.inverse_inertia_world_row1), x),
              dot(f3(b.inverse_inertia_world_row2), x)};
}

MRD_INL V3 scene_response(const MRBodyStateC& b, V3 x, V3 d) {
    const V3 lever = sub(x, f3(b.position));
    const V3 dw = inv_inertia_mul(b, cross(lever, d));
    return add(mul(b.linear_velocity_inv_mass.w, d), cross(dw, lever));
}

MRD_INL V3 scene_point_velocity(const MRBodyStateC& b, V3 x) {
    return add(f3(b.linear_velocity_inv_mass), cross(f3(b.angular_velocity), sub(x, f3(b.position))));
}

MRD_INL void apply_scene_impulse(MRBodyStateC& b, V3 x, V3 imp) {
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

MRD_INL V3 hot_dir(const HotConstraint& h, int r) { return V3{h.dir[3 * r], h.dir[3 * r + 1], h.dir[3 * r + 2]}; }

// Point velocities of a constraint's scene endpoints at the contact point
// (zero for links and static bodies).
MRD_INL void scene_point_velocities(const Env& e, uint32_t kind, uint32_t ba, uint32_t bb, V3 x, V3& va, V3& vb) {
    const V3 zero{0.0f, 0.0f, 0.0f};
    va = (kind & mr::kHotDynamicA) != 0u ? scene_point_velocity(e.scene[ba - e.nlinks], x) : zero;
    vb = (kind & mr::kHotDynamicB) != 0u ? scene_point_velocity(e.scene[bb - e.nlinks], x) : zero;
}

// Relative velocity of one constraint row: the articulation part `art_dot`
// (J v, already folded across the DoF lanes) plus the point velocity of a
// dynamic body B minus that of a dynamic body A (static and link endpoints
// add nothing), in the oracle's order.
MRD_INL float rel_row(uint32_t kind, V3 d, V3 va, V3 vb, float art_dot) {
    float rel = art_dot;
    if ((kind & mr::kHotDynamicB) != 0u) rel = rel + dot(d, vb);
    if ((kind & mr::kHotDynamicA) != 0u) rel = rel - dot(d, va);
    return rel;
}

MRD_INL uint32_t uniform_u32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// Optional per-stage shader-clock counters: lane 0 adds each stage's elapsed
// cycles straight into the environment's global counters (no register array
// held across the kernel when profiling is off).
// ------------------------------------------------ relaxed block Jacobi
// (mrWorldWave32SolvePacket, MetalWorldContact.metal:6142-7213). Articulated
// dot products are serial in DoF order and sums over contacts run in
// constraint order; contacts of other islands add exact zeros, so these are
// the island sums.

// Velocity change at point xc of scene body b from a unit impulse along d at
// point xj (the scene part of typedNormalCrossContactResponse, :6740-6784).
MRD_INL V3 scene_cross_response(const MRBodyStateC& b, V3 xc, V3 xj, V3 d) {
    const V3 dw = inv_inertia_mul(b, cross(sub(xj, f3(b.position)), d));
    return add(mul(b.linear_velocity_inv_mass.w, d), cross(dw, sub(xc, f3(b.position))));
}

// sum_k x[k] * y[k] added in index order (the order the oracle uses), with
// the loads of eight terms issued before their products are added, so that
// their memory latencies overlap instead of following one another.
MRD_INL float serial_dot(const float* x, const float* y, uint32_t n) {
    float a = 0.0f;
    uint32_t k = 0;
    for (; k + 8u <= n; k += 8u) {
        float xs[8], ys[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            xs[i] = x[k + i];
            ys[i] = y[k + i];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) a = a + xs[i] * ys[i];
    }
    for (; k < n; ++k) a = a + x[k] * y[k];
    return a;
}

// W_nn(c, j): normal velocity at contact c per unit normal impulse of contact j.
MRD_INL float jacobi_normal_cross(const Env& e, uint32_t c, uint32_t j, uint32_t mnv) {
    const uint32_t pc = e.cpair[c], pj = e.cpair[j];
    float w = 0.0f;
    if (((pc >> 16) & mr::kHotArticulated) != 0u && ((pj >> 16) & mr::kHotArticulated) != 0u) {
        w = serial_dot(e.jrow + (size_t)c * 3u * mnv, e.resp + (size_t)j * 3u * mnv, e.nv);
    }
    const uint32_t cb[2] = {pc & 0xffu, (pc >> 8) & 0xffu}, jb[2] = {pj & 0xffu, (pj >> 8) & 0xffu};
    for (int x = 0; x < 2; ++x) {
        if (!is_dynamic_scene(e, cb[x])) continue;
        for (int y = 0; y < 2; ++y) {
            if (jb[y] != cb[x]) continue;
            const HotConstraint& hc = e.hot[c];
            const HotConstraint& hj = e.hot[j];
            const float t = dot(hot_dir(hc, 0), scene_cross_response(e.scene[cb[x] - e.nlinks],
                                                                     V3{hc.point[0], hc.point[1], hc.point[2]},
                                                                     V3{hj.point[0], hj.point[1], hj.point[2]},
                                                                     hot_dir(hj, 0)));
            w = (x == y) ? w + t : w - t;
        }
    }
    return w;
}

// All impulse deltas of e.jdelta at once: a DoF lane sums its responses in
// constraint order and applies the sum when any response touches it; a scene
// lane accumulates linear and angular impulse over its contacts and applies
// them once (:6858-6990).
MRD_INL void jacobi_apply(Env& e, uint32_t lane, uint32_t ncons, uint32_t mnv, bool dof_lane, float& vj) {
    const float* jd = e.jdelta;
    if (dof_lane) {
        float vd = 0.0f;
        bool part = false;
        for (uint32_t c = 0; c < ncons; ++c) {
            if (((e.cpair[c] >> 16) & mr::kHotArticulated) == 0u) continue;
            const float* rb = e.resp + (size_t)c * 3u * mnv + lane;
            const float r0 = rb[0], r1 = rb[mnv], r2 = rb[2u * mnv];
            part = part || r0 != 0.0f || r1 != 0.0f || r2 != 0.0f;
            vd = vd + ((r0 * jd[3u * c] + r1 * jd[3u * c + 1u]) + r2 * jd[3u * c + 2u]);
        }
        if (part) vj = vj + vd;
    }
    for (uint32_t s = lane; s < e.nscene; s += kLanes) {
        const uint32_t b = e.nlinks + s;
        if (!is_dynamic_scene(e, b)) continue;
        MRBodyStateC st = e.scene[s];
        V3 lin{0.0f, 0.0f, 0.0f}, ang{0.0f, 0.0f, 0.0f};
        bool touched = false;
        for (uint32_t c = 0; c < ncons; ++c) {
            const uint32_t pr = e.cpair[c];
            if ((pr & 0xffu) != b && ((pr >> 8) & 0xffu) != b) continue;
            touched = true;
            const HotConstraint& hk = e.hot[c];
            const float* d = jd + 3u * c;
            const V3 imp = add(add(mul(d[0], hot_dir(hk, 0)), mul(d[1], hot_dir(hk, 1))), mul(d[2], hot_dir(hk, 2)));
            const V3 lever = sub(V3{hk.point[0], hk.point[1], hk.point[2]}, f3(st.position));
            if (hk.body_a == b) {
                lin = sub(lin, imp);
                ang = add(ang, cross(lever, neg(imp)));
            }
            if (hk.body_b == b) {
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
        e.scene[s] = st;
    }
}

struct StageClock {
    unsigned long long* out;
    unsigned long long last;
    MRD_INL void init(unsigned long long* o, uint32_t lane) {
        out = lane == 0u ? o : nullptr;
        last = out != nullptr ? (unsigned long long)clock64() : 0ull;
    }
    MRD_INL void mark(uint32_t stage) {
        if (out == nullptr) return;
        const unsigned long long now = (unsigned long long)clock64();
        out[stage] += now - last;
        last = now;
    }
};

} // namespace

// ------------------------------------------------ learned warm start

static_assert(mr::kMlpTileLdsBytes == mr::kMlpLdsBytes, "MLP tile LDS size");
// Normalized features beyond this many training standard deviations keep the
// plain warm start.
constexpr float kMlpFeatureBound = 8.0f;

// Features of constraint c at the start of its cone solve: normal relative
// velocity (J v_free, the evaluation's staged row), penetration depth, the
// Euclidean norms of the normal and tangent Jacobian rows (articulated DoFs
// plus the twist coordinates (dir, r x dir) of each dynamic scene endpoint)
// and the projected previous-step impulse the plain path warm-starts from.
MRD_INL void contact_features(const Env& e, uint32_t c, uint32_t mnv, float f[mr::kMlpIn]) {
    const HotConstraint& hk = e.hot[c];
    const ColdConstraint& ck = e.cold[c];
    const uint32_t kind = e.cpair[c] >> 16;
    f[0] = hk.stage[0];
    f[1] = -ck.separation;
    const V3 x{hk.point[0], hk.point[1], hk.point[2]};
    for (int r = 0; r < 3; ++r) {
        float s2 = 0.0f;
        if ((kind & mr::kHotArticulated) != 0u) {
            const float* row = e.jrow + ((size_t)c * 3u + (uint32_t)r) * mnv;
            for (uint32_t j = 0; j < e.nv; ++j) s2 += row[j] * row[j];
        }
        const V3 d = hot_dir(hk, r);
        if ((kind & mr::kHotDynamicA) != 0u) {
            const V3 t = cross(sub(x, f3(e.scene[hk.body_a - e.nlinks].position)), d);
            s2 += dot(d, d) + dot(t, t);
        }
        if ((kind & mr::kHotDynamicB) != 0u) {
            const V3 t = cross(sub(x, f3(e.scene[hk.body_b - e.nlinks].position)), d);
            s2 += dot(d, d) + dot(t, t);
        }
        f[2 + r] = sqrt_rn(s2);
    }
    f[5] = e.lam[3u * c];
    f[6] = e.lam[3u * c + 1u];
    f[7] = e.lam[3u * c + 2u];
}

MRD_INL mr::MlpView mlp_view(const MRMlpWeightsC& w) {
    mr::MlpView m;
    m.w1 = reinterpret_cast<const __bf16*>(w.w1);
    m.b1 = w.b1;
    m.w2 = reinterpret_cast<const __bf16*>(w.w2);
    m.b2 = w.b2;
    m.w3 = reinterpret_cast<const __bf16*>(w.w3);
    m.b3 = w.b3;
    m.in_shift = w.in_shift;
    m.in_scale = w.in_scale;
    m.out_scale = w.out_scale;
    return m;
}

// Features of every constraint, 16 at a time: written to `samples` (the
// environment's capture rows, or null) and, with `infer`, run through the
// MLP, which predicts the correction to the previous-step impulse:
// lambda = lambda_prev + out_scale * y_hat, projected onto the cone. The
// prediction replaces the plain warm start (kept in e.mlp_lam, as are the
// scene bodies, for the fallback) when some component of the correction
// exceeds the weights' min_correction, unless a normalized feature lies
// beyond kMlpFeatureBound (outside what the network was trained on) or the
// result is not finite; otherwise the warm start stays untouched, bit for bit.
// Returns whether any warm start was replaced. Uniform in `infer`; all 64
// lanes must be active.
MRD_INL bool mlp_warm_start(const Env& e, uint32_t lane, uint32_t ncons, uint32_t mnv, bool infer, float* samples) {
    const mr::MlpLds s = mr::mlp_lds(e.mlp_lds);
    const mr::MlpView m = mlp_view(e.a->mlp);
    if (infer) {
        for (uint32_t i = lane; i < 3u * ncons; i += kLanes) e.mlp_lam[i] = e.lam[i];
        for (uint32_t b = lane; b < e.nscene; b += kLanes) e.mlp_scene[b] = e.scene[b];
    }
    bool replaced = false;  // this lane replaced a warm start
    for (uint32_t t0 = 0; t0 < ncons; t0 += mr::kMlpRows) {
        bool in_range = false;  // lanes 0-15: their row's features are all within bounds
        __syncthreads();
        if (lane < mr::kMlpRows) {
            const uint32_t c = t0 + lane;
            float f[mr::kMlpIn] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (c < ncons) {
                contact_features(e, c, mnv, f);
                if (samples != nullptr) {
                    float* row = samples + (size_t)c * mr::kMlpSampleFloats;
                    for (uint32_t k = 0; k < mr::kMlpIn; ++k) row[k] = f[k];
                    // the warm start the solve begins from, unless the network replaces it below
                    for (uint32_t k = 0; k < 3u; ++k) row[12u + k] = e.lam[3u * c + k];
                    row[15] = 0.0f;
                    for (uint32_t k = 16u; k < 20u; ++k) row[k] = 0.0f;
                }
            }
            if (infer) {
                in_range = true;
                for (uint32_t k = 0; k < mr::kMlpIn; ++k) {
                    const float xn = c < ncons ? (f[k] - m.in_shift[k]) * m.in_scale[k] : 0.0f;
                    in_range = in_range && fabs_(xn) <= kMlpFeatureBound;
                    s.x[lane * mr::kMlpIn + k] = in_range ? xn : 0.0f;
                }
            }
        }
        __syncthreads();
        if (!infer) continue;
        float y[4];
        mr::mlp_tile_forward(m, s, lane, y, nullptr, 0u, mr::kMlpRows);
        __syncthreads();
        const uint32_t col = lane & 15u;
        if (col < mr::kMlpOut) {
            const float scale = m.out_scale[col];
            for (uint32_t r = 0; r < 4u; ++r) s.x[((lane >> 4) * 4u + r) * mr::kMlpIn + col] = y[r] * scale;
        }
        __syncthreads();
        if (samples != nullptr && lane < mr::kMlpRows && t0 + lane < ncons && in_range) {
            float* row = samples + (size_t)(t0 + lane) * mr::kMlpSampleFloats;
            for (uint32_t k = 0; k < 3u; ++k) row[16u + k] = s.x[lane * mr::kMlpIn + k];
            row[19] = 1.0f;
        }
        if (lane < mr::kMlpRows && t0 + lane < ncons && in_range) {
            const uint32_t c = t0 + lane;
            float lam[3];
            bool changed = false;
            const float gate = e.a->mlp.min_correction;
            for (uint32_t k = 0; k < 3u; ++k) changed = changed || fabs_(s.x[lane * mr::kMlpIn + k]) > gate;
            for (uint32_t k = 0; k < 3u; ++k) {
                const float prev = e.lam[3u * c + k], d = s.x[lane * mr::kMlpIn + k];
                lam[k] = changed ? prev + d : prev;
            }
            const HotConstraint& hk = e.hot[c];
            if (changed) project_cone(lam, hk.mu_u, hk.mu_v, e.a->cfg.maximum_normal_impulse);
            if (changed && isfin(lam[0]) && isfin(lam[1]) && isfin(lam[2])) {
                e.lam[3u * c] = lam[0];
                e.lam[3u * c + 1u] = lam[1];
                e.lam[3u * c + 2u] = lam[2];
                replaced = true;
                if (samples != nullptr) {
                    float* row = samples + (size_t)c * mr::kMlpSampleFloats;
                    for (uint32_t k = 0; k < 3u; ++k) row[12u + k] = lam[k];
                    row[15] = 1.0f;
                }
            }
        }
    }
    __syncthreads();
    return __ballot(replaced) != 0ull;
}

// --------------------------------------------------------------- kernel

// One control step of one environment (one wavefront per workgroup). kMlp
// selects the instance with the learned warm start and its sample capture;
// the plain instance carries none of that code.
// `step` selects the control step of a multi-step launch: the command, reset
// mask and status records of step `step` ([steps][E] each); 0 elsewhere.
template <bool kMlp, bool kCcd = false, bool kGeo = false>
__device__ __forceinline__ void control_step(const KernelArgs& args, uint32_t env,
                                                                 uint32_t step = 0u) {
    extern __shared__ __align__(16) unsigned char lds_base[];
    __shared__ unsigned long long s_fail;
    __shared__ uint32_t s_count[8];
    __shared__ float s_fmax[2];
    __shared__ float s_root[7];
    __shared__ MRContactStatusC s_cs;
    __shared__ unsigned long long s_t0;  // start of this control step on the constant-rate clock (step timing)

    const uint32_t lane = threadIdx.x;
    if (env >= args.environment_count) return;
    const KernelArgs& A = args;

    Env e;
    e.a = &A;
    e.env = env;
    e.hd = &A.models[A.environment_model[env]];
    const MRModelHeaderC& hd = *e.hd;
    e.bodies = A.bodies + hd.body_offset;
    e.dofs = A.dofs + hd.dof_offset;
    e.acts = A.actuators != nullptr ? A.actuators + hd.dof_offset : nullptr;
    e.shapes = A.shapes + hd.shape_offset;
    e.mats = A.materials + hd.material_offset;
    e.pairs = A.pairs + hd.pair_offset;
    e.nlinks = hd.link_count;
    e.nbodies = hd.body_count;
    e.nv = hd.nv;
    e.nq = hd.nq;
    e.nshapes = hd.shape_count;
    e.npairs = hd.pair_count;
    e.nscene = hd.scene_body_count;
    e.levels = hd.level_count;
    e.root_type = hd.root_type;
    e.h = A.h;
    e.gravity = V3{A.gravity[0], A.gravity[1], A.gravity[2]};
    const mr::LdsLayout& L = A.lds;
    e.topo = lds_at<mr::TopoCache>(lds_base, L.topo);
    e.lconst = lds_at<mr::LinkConst>(lds_base, L.lconst);
    e.arm = lds_at<float>(lds_base, L.arm);
    e.q = lds_at<float>(lds_base, L.q);
    e.v = lds_at<float>(lds_base, L.v);
    e.vcand = lds_at<float>(lds_base, L.vcand);
    e.effort = lds_at<float>(lds_base, L.effort);
    e.accel = lds_at<float>(lds_base, L.accel);
    e.cmd = lds_at<float>(lds_base, L.cmd);
    e.pos = lds_at<V3>(lds_base, L.pos);
    e.rot = lds_at<Q4>(lds_base, L.rot);
    e.w = lds_at<V3>(lds_base, L.w);
    e.vl = lds_at<V3>(lds_base, L.vl);
    e.sa = lds_at<V3>(lds_base, L.sa);
    e.sl = lds_at<V3>(lds_base, L.sl);
    e.ca = lds_at<V3>(lds_base, L.ca);
    e.cl = lds_at<V3>(lds_base, L.cl);
    e.r = lds_at<V3>(lds_base, L.r);
    e.I = lds_at<SpI>(lds_base, L.I);
    e.pt = lds_at<V3>(lds_base, L.pt);
    e.pb = lds_at<V3>(lds_base, L.pb);
    e.Ut = lds_at<V3>(lds_base, L.Ut);
    e.Ub = lds_at<V3>(lds_base, L.Ub);
    e.D = lds_at<float>(lds_base, L.D);
    e.u = lds_at<float>(lds_base, L.u);
    e.alpha = lds_at<V3>(lds_base, L.alpha);
    e.acc = lds_at<V3>(lds_base, L.acc);
    e.L = lds_at<float>(lds_base, L.L);
    e.scene = lds_at<MRBodyStateC>(lds_base, L.scene);
    e.col = lds_at<Collider>(lds_base, L.col);
    e.active = lds_at<uint32_t>(lds_base, L.active);
    e.npts = lds_at<uint32_t>(lds_base, L.npts);
    e.moff = lds_at<uint32_t>(lds_base, L.moff);
    e.okey = lds_at<uint32_t>(lds_base, L.okey);
    e.hfitem = lds_at<uint32_t>(lds_base, L.hfitem);
    e.hfq = lds_at<mr::HfHit>(lds_base, L.hfq);
    e.lam = lds_at<float>(lds_base, L.lam);
    e.cpair = lds_at<uint32_t>(lds_base, L.cpair);
    e.rslot = lds_at<float>(lds_base, L.rslot);
    e.rstack = lds_at<float>(lds_base, L.rstack);
    e.isl = lds_at<uint8_t>(lds_base, L.isl);
    e.resid = lds_at<float>(lds_base, L.resid);
    e.jrelax = lds_at<float>(lds_base, L.jrelax);
    e.jdelta = lds_at<float>(lds_base, L.jdelta);
    if constexpr (kMlp) {
        e.mlp_lds = lds_base + L.mlp;
        e.mlp_lam = lds_at<float>(lds_base, L.mlp_lam);
        e.mlp_scene = lds_at<MRBodyStateC>(lds_base, L.mlp_scene);
    }
    const uint32_t pcap = A.pair_capacity, mcap = A.manifold_capacity, ccap = A.constraint_capacity;
    e.stage_h = A.stage_h + (size_t)env * pcap;
    e.stage_p = A.stage_p + (size_t)env * pcap * MR_MANIFOLD_POINT_CAPACITY;
    e.jrow = A.jrow + (size_t)env * ccap * 3u * A.row_stride;
    e.resp = A.resp + (size_t)env * ccap * 3u * A.row_stride;
    e.cold = A.cold + (size_t)env * ccap;
    e.hot = A.hot + (size_t)env * ccap;

    // ---- topology cache (one lane per body / DoF; 64 lanes cover the limits)
    {
        mr::TopoCache& T = *e.topo;
        uint8_t par = 0xffu, vi = 0xffu, qi = 0xffu, dep = 0u, mot = 0u, jt = 0u;
        if (lane < e.nbodies) {
            const MRBodyC& b = e.bodies[lane];
            mot = (uint8_t)b.motion_type;
            if (lane < e.nlinks) {
                par = b.parent == MR_INVALID_INDEX ? (uint8_t)0xffu : (uint8_t)b.parent;
                vi = b.v_index == MR_INVALID_INDEX ? (uint8_t)0xffu : (uint8_t)b.v_index;
                qi = b.q_index == MR_INVALID_INDEX ? (uint8_t)0xffu : (uint8_t)b.q_index;
                dep = (uint8_t)b.depth;
                jt = (uint8_t)b.joint_type;
            }
        }
        T.qidx[lane] = qi;
        T.jtype[lane] = jt;
        T.parent[lane] = par;
        T.vidx[lane] = vi;
        T.depth[lane] = dep;
        T.motion[lane] = mot;
        T.dof_body[lane] = lane < e.nv ? (uint8_t)e.dofs[lane].body : (uint8_t)0xffu;
        if (lane < e.nlinks) {
            const MRBodyC& b = e.bodies[lane];
            mr::LinkConst lc;
            const Q4 cc = qconj(q4(b.child_rotation));
            const V3 ax = normalize(f3(b.axis));
            lc.prot[0] = b.parent_rotation.x; lc.prot[1] = b.parent_rotation.y;
            lc.prot[2] = b.parent_rotation.z; lc.prot[3] = b.parent_rotation.w;
            lc.crot_conj[0] = cc.x; lc.crot_conj[1] = cc.y; lc.crot_conj[2] = cc.z; lc.crot_conj[3] = cc.w;
            lc.axis[0] = ax.x; lc.axis[1] = ax.y; lc.axis[2] = ax.z;
            lc.panchor[0] = b.parent_anchor.x; lc.panchor[1] = b.parent_anchor.y; lc.panchor[2] = b.parent_anchor.z;
            lc.canchor[0] = b.child_anchor.x; lc.canchor[1] = b.child_anchor.y; lc.canchor[2] = b.child_anchor.z;
            e.lconst[lane] = lc;
        }
        if (lane < e.nv) e.arm[lane] = effective_armature(e, lane);
        __syncthreads();
        // model validation guarantees parent < child, so both walks end
        unsigned long long anc = 0ull, ch = 0ull;
        if (lane < e.nlinks) {
            for (uint32_t b = lane; b != 0u && b != 0xffu; b = T.parent[b]) anc |= 1ull << b;
            for (uint32_t c = lane + 1u; c < e.nlinks; ++c)
                if (T.parent[c] == lane) ch |= 1ull << c;
        }
        T.anc[lane] = anc;
        T.children[lane] = ch;
        T.slot[lane] = A.response_slots[(size_t)A.environment_model[env] * MR_MAX_BODIES + lane];
    }

    const uint32_t nv = e.nv, nq = e.nq, nl = e.nlinks, nsc = e.nscene;
    float* gq = A.q + (size_t)env * A.max_nq;
    float* gv = A.v + (size_t)env * A.max_nv;
    MRBodyStateC* gs = A.scene + (size_t)env * A.max_scene;
    const uint32_t cur = A.mcurrent[env];
    const size_t E = A.environment_count;

    // ---- prepare: reset, load state (the HBM state is the checkpoint)
    const size_t step_rec = (size_t)step * E + env;  // this step's record of this environment
    if (A.step_us != nullptr && lane == 0u) s_t0 = __builtin_amdgcn_s_memrealtime();
    const bool reset = A.reset_mask != nullptr && A.reset_mask[step_rec] != 0u;
    for (uint32_t i = lane; i < nq; i += kLanes) {
        const float val = reset ? A.reset_q[(size_t)env * A.max_nq + i] : gq[i];
        e.q[i] = val;
        if (reset) gq[i] = val;
    }
    for (uint32_t i = lane; i < nv; i += kLanes) {
        const float val = reset ? A.reset_v[(size_t)env * A.max_nv + i] : gv[i];
        e.v[i] = val;
        if (reset) gv[i] = val;
        e.cmd[i] = A.command[step_rec * A.max_nv + i];
    }
    for (uint32_t s = lane; s < nsc; s += kLanes) {
        const MRBodyStateC val = reset ? A.reset_scene[(size_t)env * A.max_scene + s] : gs[s];
        e.scene[s] = val;
        if (reset) gs[s] = val;
    }
    if (lane == 0u) {
        if (reset) A.mcount[(size_t)cur * E + env] = 0u;
        s_fail = ~0ull;
    }
    __syncthreads();

    StageClock clk;
    clk.init(A.stage_cycles != nullptr ? A.stage_cycles + (size_t)env * mr::kStageCount : nullptr, lane);
    uint32_t src_buf = cur;
    uint32_t dst_buf = (cur + 1u) % 3u;
    const uint32_t alt_buf = (cur + 2u) % 3u;
    uint32_t mcount_src = A.mcount[(size_t)cur * E + env];
    if (mcount_src > mcap) mcount_src = mcap;

    // Contact status lives in LDS (uniform fields are written with identical
    // values by every lane; lane-0-only fields by lane 0).
    MRContactStatusC& cs = s_cs;
    if (lane == 0u) {
        unsigned char* p = reinterpret_cast<unsigned char*>(&cs);
        for (uint32_t k = 0; k < sizeof(cs); ++k) p[k] = 0;
        cs.first_failing_pair = MR_INVALID_INDEX;
        cs.first_failing_constraint = MR_INVALID_INDEX;
    }
    __syncthreads();
    uint32_t successful = 0u;
    float min_piv = __builtin_inff(), max_piv = 0.0f, max_acc = 0.0f;
    uint32_t failing_substep = MR_INVALID_INDEX;
    float h = e.h;
    const MRWorldStepConfigC& cfg = A.cfg;
    const bool floating = e.root_type == MR_ROOT_FLOATING;
    uint32_t impact_segments = 0u;  // segments that ended at a time of impact (hybrid CCD)

    // A microstep is one segment, or (hybrid CCD) a sequence of segments ending
    // at successive times of impact; each segment runs every stage below with
    // h = its duration, and the microstep counts once its last segment commits.
    float remaining = A.h;
    uint32_t segment = 0u;
    for (uint32_t substep = 0; substep < cfg.physics_substeps;) {
        if constexpr (kCcd) {
            h = remaining;
            e.h = h;
        }
        bool impact = false;
        // ---------------------------------------------------------- drives
        // Implicit drives refresh from the current state every microstep; an
        // effort command is clamped once, at the control step's start velocity,
        // and held (MetalWorld.metal:184-270 prepare, 280-366 refresh). They
        // depend on h, so the hybrid mode refreshes them once the segment's
        // duration is known, below.
        if constexpr (!kCcd) {
            if (substep == 0u || cfg.actuation_mode == MR_ACTUATION_IMPLICIT_POSITION_DRIVE)
                for (uint32_t j = lane; j < nv; j += kLanes) e.effort[j] = drive_effort(e, j);
            __syncthreads();
        }
        // ------------------------------------------------------ kinematics
        if (lane == 0u) {
            if (floating) {
                Q4 rq;
                if (!checked_quat(Q4{e.q[3], e.q[4], e.q[5], e.q[6]}, rq)) {
                    report_failure(&s_fail, 0u, MR_STEP_NONFINITE_INPUT, 3u);
                    rq = Q4{0.0f, 0.0f, 0.0f, 1.0f};
                }
                e.pos[0] = V3{e.q[0], e.q[1], e.q[2]};
                e.rot[0] = rq;
                e.vl[0] = V3{e.v[0], e.v[1], e.v[2]};
                e.w[0] = V3{e.v[3], e.v[4], e.v[5]};
            } else {
                e.pos[0] = V3{0.0f, 0.0f, 0.0f};
                e.rot[0] = Q4{0.0f, 0.0f, 0.0f, 1.0f};
                e.vl[0] = V3{0.0f, 0.0f, 0.0f};
                e.w[0] = V3{0.0f, 0.0f, 0.0f};
            }
            const V3 zero{0.0f, 0.0f, 0.0f};
            e.r[0] = zero;
            e.sa[0] = zero;
            e.sl[0] = zero;
            e.ca[0] = zero;
            e.cl[0] = zero;
        }
        __syncthreads();
        for (uint32_t d = 1; d < e.levels; ++d) {
            for (uint32_t i = 1u + lane; i < nl; i += kLanes)
                if (e.topo->depth[i] == d && !fk_link(e, i))
                    report_failure(&s_fail, 1u + i, MR_STEP_NONFINITE_RESULT, i);
            __syncthreads();
        }
        clk.mark(mr::kStageKinematics);
        if (s_fail != ~0ull) break;
        // ----------------------------------------------- time of impact
        if constexpr (kCcd) {
            uint32_t my_bits = ~0u, my_pair = ~0u, my_unsafe = ~0u;
            for (uint32_t k = lane; k < e.npairs; k += kLanes) {
                float toi = 0.0f;
                const uint32_t r = ccd_pair<kGeo>(e, k, remaining, toi);
                if (r == kCcdImpact) {
                    const uint32_t bits = __float_as_uint(fmaxd(toi, 0.0f));
                    if (bits < my_bits) {
                        my_bits = bits;
                        my_pair = k;
                    }
                } else if (r == kCcdUnsafe && my_unsafe == ~0u) {
                    my_unsafe = k;
                }
            }
            // earliest (TOI bits, pair) over the wave: non-negative floats order as their bits
            const uint32_t best_bits = wave_min_u32(my_bits);
            const uint32_t best_pair = wave_min_u32(my_bits == best_bits ? my_pair : ~0u);
            const uint32_t unsafe = wave_min_u32(my_unsafe);
            if (unsafe != ~0u) report_failure(&s_fail, unsafe, MR_STEP_DID_NOT_CONVERGE, unsafe);
            if (best_bits != ~0u) {
                const float tau = fmaxd(__uint_as_float(best_bits), cfg.ccd_time_tolerance);
                if (tau < remaining - cfg.ccd_time_tolerance) {
                    if (segment >= cfg.max_ccd_events)
                        report_failure(&s_fail, best_pair, MR_STEP_CCD_EVENT_BUDGET_EXHAUSTED, best_pair);
                    impact = true;
                    h = tau;
                    e.h = h;
                }
            }
            __syncthreads();
            if (s_fail != ~0ull) break;
        }
        const bool final_sub = substep + 1u == cfg.physics_substeps && !impact;
        const bool capture = cfg.capture_evidence != 0u && substep == cfg.capture_substep && segment == 0u &&
            A.evidence_pairs != nullptr;
        if constexpr (kCcd) {
            if ((substep == 0u && segment == 0u) || cfg.actuation_mode == MR_ACTUATION_IMPLICIT_POSITION_DRIVE)
                for (uint32_t j = lane; j < nv; j += kLanes) e.effort[j] = drive_effort(e, j);
            __syncthreads();
        }
        // -------------------------------------------------------------- ABA
        for (uint32_t i = lane; i < nl; i += kLanes) aba_init_link(e, i);
        __syncthreads();
        if (lane == 0u && floating) {
            SpI& I = e.I[0];
            I.C[0] += e.arm[0];
            I.C[3] += e.arm[1];
            I.C[5] += e.arm[2];
            I.A[0] += e.arm[3];
            I.A[3] += e.arm[4];
            I.A[5] += e.arm[5];
            e.pt[0] = sub(e.pt[0], V3{e.effort[3], e.effort[4], e.effort[5]});
            e.pb[0] = sub(e.pb[0], V3{e.effort[0], e.effort[1], e.effort[2]});
        }
        __syncthreads();
        float lane_min_piv = __builtin_inff(), lane_max_piv = 0.0f;
        for (uint32_t d = e.levels - 1u; d >= 1u; --d) {
            for (uint32_t i = 1u + lane; i < nl; i += kLanes) {
                if (e.topo->depth[i] != d) continue;
                float piv;
                if (!aba_project_link(e, i, piv))
                    report_failure(&s_fail, 0xffffu - i, MR_STEP_FACTORIZATION_FAILED, e.bodies[i].v_index);
                if (piv >= 0.0f) {
                    lane_min_piv = fmind(lane_min_piv, piv);
                    lane_max_piv = fmaxd(lane_max_piv, piv);
                }
            }
            __syncthreads();
            for (uint32_t p = lane; p < nl; p += kLanes)
                if (e.topo->depth[p] == d - 1u) aba_accumulate(e, p);
            __syncthreads();
        }
        if (s_fail != ~0ull) break;
        lane_min_piv = -wave_max(-lane_min_piv);
        lane_max_piv = wave_max(lane_max_piv);
        if (lane == 0u) {
            if (floating) {
                uint32_t fail_row = 0;
                float mn = lane_min_piv, mx = lane_max_piv;
                if (!root_factor(e, mn, mx, fail_row)) {
                    report_failure(&s_fail, 0x10000u, MR_STEP_FACTORIZATION_FAILED, fail_row);
                } else {
                    const float rhs[6] = {-e.pt[0].x, -e.pt[0].y, -e.pt[0].z, -e.pb[0].x, -e.pb[0].y, -e.pb[0].z};
                    float x[6];
                    root_solve(e.L, rhs, x);
                    e.alpha[0] = V3{x[0], x[1], x[2]};
                    e.acc[0] = V3{x[3], x[4], x[5]};
                    e.accel[0] = x[3]; e.accel[1] = x[4]; e.accel[2] = x[5];
                    e.accel[3] = x[0]; e.accel[4] = x[1]; e.accel[5] = x[2];
                }
                lane_min_piv = mn;
                lane_max_piv = mx;
            } else {
                e.alpha[0] = V3{0.0f, 0.0f, 0.0f};
                e.acc[0] = V3{0.0f, 0.0f, 0.0f};
            }
            s_fmax[0] = lane_min_piv;
            s_fmax[1] = lane_max_piv;
        }
        __syncthreads();
        if (s_fail != ~0ull) break;
        for (uint32_t d = 1; d < e.levels; ++d) {
            for (uint32_t i = 1u + lane; i < nl; i += kLanes) {
                if (e.topo->depth[i] != d) continue;
                const uint32_t parent = link_parent(e, i), vi = link_vidx(e, i);
                const V3 ap = e.alpha[parent];
                const V3 lp = add(e.acc[parent], cross(ap, e.r[i]));
                V3 ai = add(ap, e.ca[i]);
                V3 li = add(lp, e.cl[i]);
                if (vi != MR_INVALID_INDEX) {
                    const float proj = dot(e.Ut[i], ap) + dot(e.Ub[i], lp);
                    const float qdd = (e.u[i] - proj) / e.D[i];
                    e.accel[vi] = qdd;
                    ai = add(ai, mul(qdd, e.sa[i]));
                    li = add(li, mul(qdd, e.sl[i]));
                }
                e.alpha[i] = ai;
                e.acc[i] = li;
            }
            __syncthreads();
        }
        for (uint32_t j = lane; j < nv; j += kLanes) {
            const float a = e.accel[j];
            const float nvv = e.v[j] + h * a;
            if (!isfin(a) || !isfin(nvv)) report_failure(&s_fail, j, MR_STEP_NONFINITE_RESULT, j);
            e.vcand[j] = nvv;
        }
        __syncthreads();
        clk.mark(mr::kStageABA);
        if (s_fail != ~0ull) break;
        // ------------------------------------------- scene prediction
        for (uint32_t s = lane; s < nsc; s += kLanes) {
            const uint32_t b = nl + s;
            const MRBodyC& bp = e.bodies[b];
            MRBodyStateC st = e.scene[s];
            Q4 rq;
            if (!normalized_quat(q4(st.orientation), rq) || !finite3(f3(st.position)) ||
                !finite3(f3(st.linear_velocity_inv_mass)) || !finite3(f3(st.angular_velocity))) {
                report_failure(&s_fail, b, MR_STEP_NONFINITE_INPUT, b);
                continue;
            }
            st.position.w = 1.0f;
            st.orientation = mr_float4{rq.x, rq.y, rq.z, rq.w};
            st.linear_velocity_inv_mass.w = bp.motion_type == MR_MOTION_DYNAMIC ? bp.mass.y : 0.0f;
            st.angular_velocity.w = 0.0f;
            st.flags_indices[0] = bp.motion_type;
            st.flags_indices[1] = MR_INVALID_INDEX;
            st.flags_indices[2] = b;
            world_inverse_inertia(bp, rq, st);
            if (bp.motion_type == MR_MOTION_DYNAMIC) {
                // velocity damping is a per-microstep factor: its first segment applies it
                const float dl = !kCcd || segment == 0u ? bp.damping_scale.x : 1.0f;
                const float da = !kCcd || segment == 0u ? bp.damping_scale.y : 1.0f;
                const V3 lv = add(mul(dl, f3(st.linear_velocity_inv_mass)), mul(h, e.gravity));
                const V3 av = mul(da, f3(st.angular_velocity));
                st.linear_velocity_inv_mass.x = lv.x;
                st.linear_velocity_inv_mass.y = lv.y;
                st.linear_velocity_inv_mass.z = lv.z;
                st.angular_velocity.x = av.x;
                st.angular_velocity.y = av.y;
                st.angular_velocity.z = av.z;
            } else {
                st.linear_velocity_inv_mass.x = 0.0f;
                st.linear_velocity_inv_mass.y = 0.0f;
                st.linear_velocity_inv_mass.z = 0.0f;
                st.angular_velocity.x = 0.0f;
                st.angular_velocity.y = 0.0f;
                st.angular_velocity.z = 0.0f;
            }
            e.scene[s] = st;
        }
        __syncthreads();
        if (s_fail != ~0ull) break;
        // ------------------------------------------- collider projection
        for (uint32_t s = lane; s < e.nshapes; s += kLanes) {
            Collider c;
            if (!project_collider<kGeo>(e, s, c)) report_failure(&s_fail, s, MR_STEP_NONFINITE_INPUT, s);
            e.col[s] = c;
        }
        __syncthreads();
        clk.mark(mr::kStageProjection);
        if (s_fail != ~0ull) break;
        // ----------------------------------------------- broadphase flags
        uint32_t nact = 0;
        for (uint32_t base = 0; base < e.npairs; base += kLanes) {
            const uint32_t p = base + lane;
            bool flag = false;
            if (p < e.npairs) {
                const MRCompiledPairC pr = e.pairs[p];
                flag = pair_may_overlap(e.col[pr.collider_a], e.col[pr.collider_b]);
            }
            const unsigned long long bal = __ballot(flag);
            const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            if (flag) {
                const uint32_t slot = nact + before;
                if (slot < pcap) e.active[slot] = p;
            }
            nact += (uint32_t)__popcll(bal);
        }
        __syncthreads();
        if (capture) {
            const uint32_t nrec = nact < pcap ? nact : pcap;
            for (uint32_t k = lane; k < nrec; k += kLanes) {
                const MRCompiledPairC pr = e.pairs[e.active[k]];
                MRPairRecordC rec;
                rec.environment = env;
                rec.collider_a = pr.collider_a;
                rec.collider_b = pr.collider_b;
                rec.pair_class = pr.pair_class;
                A.evidence_pairs[(size_t)env * pcap + k] = rec;
            }
            if (lane == 0u && A.evidence_counts != nullptr) A.evidence_counts[env] = nact;
        }
        cs.required_pairs = nact;
        cs.active_pairs = nact;
        if (cs.max_active_pairs < nact) cs.max_active_pairs = nact;
        if (nact > pcap) {
            if (lane == 0u) report_failure(&s_fail, 0u, MR_STEP_PAIR_CAPACITY_OVERFLOW, 0u);
            // first overflowing compiled pair is the (pcap+1)-th flagged pair
            uint32_t seen = 0, idx = 0;
            for (uint32_t base = 0; base < e.npairs; base += kLanes) {
                const uint32_t p = base + lane;
                bool flag = false;
                if (p < e.npairs) {
                    const MRCompiledPairC pr = e.pairs[p];
                    flag = pair_may_overlap(e.col[pr.collider_a], e.col[pr.collider_b]);
                }
                const unsigned long long bal = __ballot(flag);
                const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                if (flag && seen + before == pcap) idx = p + 1u;
                seen += (uint32_t)__popcll(bal);
            }
            idx = wave_max_u32(idx);
            if (lane == 0u) cs.first_failing_pair = idx - 1u;
            __syncthreads();
            if (lane == 0u) s_fail = ((unsigned long long)MR_STEP_PAIR_CAPACITY_OVERFLOW << 24) | ((idx - 1u) & 0xffffffu);
            __syncthreads();
            break;
        }
        clk.mark(mr::kStageBroadphase);
        // ---------------------------------- narrowphase + manifold refresh
        const MRManifoldHeaderC* old_h = A.mh + ((size_t)src_buf * E + env) * mcap;
        const MRManifoldPointC* old_p = A.mp + ((size_t)src_buf * E + env) * mcap * MR_MANIFOLD_POINT_CAPACITY;
        uint32_t my_raw = 0, my_ret = 0, my_fresh = 0;
        float my_pen = 0.0f;
        // collider keys of the previous manifolds, staged once for the lookups
        for (uint32_t m = lane; m < mcount_src; m += kLanes) {
            const uint32_t ka = old_h[m].pair_and_count[1], kb = old_h[m].pair_and_count[2];
            e.okey[m] = (ka < 0x10000u && kb < 0x10000u) ? (ka | (kb << 16)) : 0xffffffffu;
        }
        // Heightfield queries of all active pairs, one lane per query point
        // (a sphere has one, a capsule three, a box eight), into e.hfq; moff
        // holds each pair's first query until the manifold prefix below. A pair
        // whose queries do not fit the table evaluates them in its own lane.
        // Query offsets grow with the pair index, so the pairs that fit own
        // exactly the slots [0, used); the slots of a pair that straddles the
        // table end are never written and never read.
        {
            uint32_t nhf = 0, used = 0;
            for (uint32_t base = 0; base < nact; base += kLanes) {
                const uint32_t k = base + lane;
                uint32_t nq = 0;
                if (k < nact) {
                    const MRCompiledPairC pr = e.pairs[e.active[k]];
                    if (pr.pair_class == MR_PAIR_HEIGHTFIELD) {
                        const uint32_t fsi = e.col[pr.collider_a].type == MR_SHAPE_HEIGHTFIELD ? pr.collider_b
                                                                                              : pr.collider_a;
                        nq = hf_query_count(e.col[fsi].type);
                    }
                }
                uint32_t tot;
                const uint32_t before = wave_prefix_sum(nq, lane, tot);
                uint32_t end = 0;
                if (k < nact) {
                    const uint32_t first = nhf + before;
                    const bool fits = nq > 0u && first + nq <= mr::kHfSlots;
                    e.moff[k] = fits ? first : 0xffffffffu;
                    if (fits) {
                        for (uint32_t q = 0; q < nq; ++q) e.hfitem[first + q] = k | (q << 16);
                        end = first + nq;
                    }
                }
                const uint32_t chunk_end = wave_max_u32(end);
                used = chunk_end > used ? chunk_end : used;
                nhf += tot;
            }
            __syncthreads();
            const uint32_t items = used;
            for (uint32_t i = lane; i < items; i += kLanes) {
                const uint32_t k = e.hfitem[i] & 0xffffu, q = e.hfitem[i] >> 16;
                const MRCompiledPairC pr = e.pairs[e.active[k]];
                const uint32_t fsi = e.col[pr.collider_a].type == MR_SHAPE_HEIGHTFIELD ? pr.collider_b : pr.collider_a;
                const Collider& fs = e.col[fsi];
                e.hfq[i] = heightfield_query(e, hf_query_point(fs, q), hf_query_radius(fs));
            }
        }
        __syncthreads();
        for (uint32_t k = lane; k < nact; k += kLanes) {
            const MRCompiledPairC pr = e.pairs[e.active[k]];
            const Collider ca_ = e.col[pr.collider_a];
            const Collider cb_ = e.col[pr.collider_b];
            RawBatch raw;
            generate_contacts<kGeo>(e, pr.collider_a, pr.collider_b, pr.pair_class, ca_, cb_, raw, e.moff[k]);
            bool bad = false;
            for (uint32_t c = 0; c < raw.count; ++c)
                if (!finite3(raw.c[c].normal) || !isfin(raw.c[c].separation) || !finite3(raw.c[c].point_a) ||
                    !finite3(raw.c[c].point_b))
                    bad = true;
            uint32_t points = 0;
            if (bad) {
                report_failure(&s_fail, k, MR_STEP_NONFINITE_RESULT, e.active[k]);
            } else if (raw.count > 0u) {
                my_raw += raw.count;
                const float rest = e.shapes[pr.collider_a].contact_rest_radius.y + e.shapes[pr.collider_b].contact_rest_radius.y;
                for (uint32_t c = 0; c < raw.count; ++c) my_pen = fmaxd(my_pen, fmaxd(rest - raw.c[c].separation, 0.0f));
                const MRManifoldHeaderC* oh = nullptr;
                const MRManifoldPointC* op = nullptr;
                const uint32_t key = pr.collider_a | (pr.collider_b << 16);
                for (uint32_t m = 0; m < mcount_src; ++m) {
                    if (e.okey[m] == key) {
                        oh = &old_h[m];
                        op = &old_p[(size_t)m * MR_MANIFOLD_POINT_CAPACITY];
                        break;
                    }
                }
                MRManifoldHeaderC mh;
                MRManifoldPointC mp[MR_MANIFOLD_POINT_CAPACITY];
                if (!build_manifold(e, pr, raw, oh, op, mh, mp, my_ret, my_fresh)) {
                    report_failure(&s_fail, k, MR_STEP_NONFINITE_RESULT, e.active[k]);
                } else {
                    points = mh.pair_and_count[3];
                    e.stage_h[k] = mh;
                    for (uint32_t p = 0; p < MR_MANIFOLD_POINT_CAPACITY; ++p)
                        e.stage_p[(size_t)k * MR_MANIFOLD_POINT_CAPACITY + p] = mp[p];
                }
            }
            e.npts[k] = points;
        }
        const uint32_t raw_total = wave_sum_u32(my_raw);
        const uint32_t ret_total = wave_sum_u32(my_ret);
        const uint32_t fresh_total = wave_sum_u32(my_fresh);
        const float pen = wave_max(my_pen);
        __syncthreads();
        clk.mark(mr::kStageNarrowphase);
        if (s_fail != ~0ull) break;
        // prefix over manifolds (count>0) and constraint offsets
        uint32_t nman = 0, ncons = 0;
        for (uint32_t base = 0; base < nact; base += kLanes) {
            const uint32_t k = base + lane;
            const uint32_t pts = k < nact ? e.npts[k] : 0u;
            const unsigned long long bal = __ballot(pts > 0u);
            const uint32_t mbefore = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            uint32_t tot;
            const uint32_t cbefore = wave_prefix_sum(pts, lane, tot);
            if (k < nact) e.moff[k] = pts > 0u ? ((nman + mbefore) | ((ncons + cbefore) << 16u)) : 0xffffffffu;
            nman += (uint32_t)__popcll(bal);
            ncons += tot;
        }
        cs.raw_contacts = raw_total;
        cs.required_manifolds = nman;
        cs.manifolds = nman;
        cs.retained_points = ret_total;
        cs.new_points = fresh_total;
        cs.diagnostics.x = pen;
        if (nman > mcap) {
            if (lane == 0u) s_fail = ((unsigned long long)MR_STEP_MANIFOLD_CAPACITY_OVERFLOW << 24) | (nman & 0xffffffu);
            __syncthreads();
            break;
        }
        cs.required_constraints = ncons;
        if (ncons > ccap) {
            if (lane == 0u) s_fail = ((unsigned long long)MR_STEP_CONSTRAINT_CAPACITY_OVERFLOW << 24) | (ncons & 0xffffffu);
            __syncthreads();
            break;
        }
        cs.constraints = ncons;
        if (cs.max_constraints < ncons) cs.max_constraints = ncons;
        __syncthreads();
        // ---------------------------------------- compact + emission
        MRManifoldHeaderC* new_h = A.mh + ((size_t)dst_buf * E + env) * mcap;
        MRManifoldPointC* new_p = A.mp + ((size_t)dst_buf * E + env) * mcap * MR_MANIFOLD_POINT_CAPACITY;
        for (uint32_t k = lane; k < nact; k += kLanes) {
            const uint32_t mo = e.moff[k];
            if (mo == 0xffffffffu) continue;
            const uint32_t m = mo & 0xffffu, c0 = mo >> 16u;
            const MRManifoldHeaderC mh = e.stage_h[k];
            new_h[m] = mh;
            for (uint32_t p = 0; p < MR_MANIFOLD_POINT_CAPACITY; ++p)
                new_p[(size_t)m * MR_MANIFOLD_POINT_CAPACITY + p] = e.stage_p[(size_t)k * MR_MANIFOLD_POINT_CAPACITY + p];
            const MRShapeC& sa = e.shapes[mh.pair_and_count[1]];
            const MRShapeC& sb = e.shapes[mh.pair_and_count[2]];
            V3 pa_pos, pb_pos;
            Q4 pa_q, pb_q;
            body_pose(e, sa.body, pa_pos, pa_q);
            body_pose(e, sb.body, pb_pos, pb_q);
            Q4 ra, rb;
            checked_quat(pa_q, ra);
            checked_quat(pb_q, rb);
            const V3 n = normalize(qrot(ra, f3(mh.normal_and_age)));
            V3 t = qrot(ra, f3(mh.tangent_and_metric));
            t = sub(t, mul(dot(t, n), n));
            t = dot(t, t) > 1.0e-12f ? normalize(t) : stable_tangent(n);
            const V3 bt = cross(n, t);
            const MRMaterialC& ma = e.mats[sa.material];
            const MRMaterialC& mb = e.mats[sb.material];
            const float rest = sa.contact_rest_radius.y + sb.contact_rest_radius.y;
            for (uint32_t p = 0; p < mh.pair_and_count[3]; ++p) {
                const MRManifoldPointC mp = e.stage_p[(size_t)k * MR_MANIFOLD_POINT_CAPACITY + p];
                const uint32_t c = c0 + p;
                HotConstraint hk;
                ColdConstraint ck;
                hk.body_a = sa.body;
                hk.body_b = sb.body;
                hk.flags = mp.feature_and_life[2] == 0u ? MR_CONSTRAINT_FLAG_NEW_IMPACT : MR_CONSTRAINT_FLAG_WARM_STARTED;
                hk.kind = ((is_link(e, sa.body) || is_link(e, sb.body)) ? mr::kHotArticulated : 0u) |
                          (is_dynamic_scene(e, sa.body) ? mr::kHotDynamicA : 0u) |
                          (is_dynamic_scene(e, sb.body) ? mr::kHotDynamicB : 0u);
                const V3 pa = add(pa_pos, qrot(ra, f3(mp.local_anchor_a)));
                const V3 pb = add(pb_pos, qrot(rb, f3(mp.local_anchor_b)));
                const V3 pw = mul(0.5f, add(pa, pb));
                hk.point[0] = pw.x; hk.point[1] = pw.y; hk.point[2] = pw.z;
                hk.dir[0] = n.x; hk.dir[1] = n.y; hk.dir[2] = n.z;
                hk.dir[3] = t.x; hk.dir[4] = t.y; hk.dir[5] = t.z;
                hk.dir[6] = bt.x; hk.dir[7] = bt.y; hk.dir[8] = bt.z;
                ck.pa[0] = pa.x; ck.pa[1] = pa.y; ck.pa[2] = pa.z;
                ck.pb[0] = pb.x; ck.pb[1] = pb.y; ck.pb[2] = pb.z;
                ck.separation = dot(sub(pb, pa), n) - rest;
                ck.manifold = m;
                ck.point = p;
                ck.flags = hk.flags;
                const float fa[4] = {ma.friction.x, ma.friction.y, ma.friction.z, ma.friction.w};
                const float fb[4] = {mb.friction.x, mb.friction.y, mb.friction.z, mb.friction.w};
                for (int q = 0; q < 4; ++q) ck.friction[q] = sqrt_rn(fmaxd(0.0f, fa[q] * fb[q]));
                ck.restitution = fmaxd(ma.response.x, mb.response.x);
                ck.restitution_threshold = fmaxd(ma.response.y, mb.response.y);
                ck.compliance = ma.response.z + mb.response.z;
                ck.dissipation = ma.response.w + mb.response.w;
                const bool warm = cfg.warm_start != 0u;
                float wi[3] = {mp.impulses.x, mp.impulses.y, mp.impulses.z};
                if constexpr (kCcd) {
                    // impulses scale with the step: a segment warm-starts from the last
                    // segment's impulses times the ratio of the two durations
                    if (mp.impulses.w > 0.0f) {
                        const float s = h / mp.impulses.w;
                        for (int r = 0; r < 3; ++r) wi[r] = wi[r] * s;
                    }
                }
                e.lam[3u * c + 0u] = warm ? wi[0] : 0.0f;
                e.lam[3u * c + 1u] = warm ? wi[1] : 0.0f;
                e.lam[3u * c + 2u] = warm ? wi[2] : 0.0f;
                e.cpair[c] = sa.body | (sb.body << 8) | (hk.kind << 16);
                e.hot[c] = hk;
                e.cold[c] = ck;
            }
        }
        __syncthreads();
        clk.mark(mr::kStageEmission);
        // --------------------------- Jacobian rows + inverse-ABA responses
        for (uint32_t item = lane; item < 3u * ncons; item += kLanes) {
            const uint32_t c = item / 3u;
            const int row = (int)(item - 3u * c);
            const HotConstraint& hk = e.hot[c];
            if ((hk.kind & mr::kHotArticulated) == 0u) continue;
            const ColdConstraint& ck = e.cold[c];
            const bool la = is_link(e, hk.body_a), lb = is_link(e, hk.body_b);
            const V3 pa{ck.pa[0], ck.pa[1], ck.pa[2]};
            const V3 pb{ck.pb[0], ck.pb[1], ck.pb[2]};
            const V3 d = hot_dir(hk, row);
            float* jr = e.jrow + ((size_t)c * 3u + (size_t)row) * A.row_stride;
            for (uint32_t j = 0; j < nv; ++j) {
                const V3 cb = lb ? point_jacobian_column(e, hk.body_b, pb, j) : V3{0.0f, 0.0f, 0.0f};
                const V3 ca = la ? point_jacobian_column(e, hk.body_a, pa, j) : V3{0.0f, 0.0f, 0.0f};
                jr[j] = dot(d, sub(cb, ca));
            }
            float* rs = e.resp + ((size_t)c * 3u + (size_t)row) * A.row_stride;
            inverse_aba(e, lane, la ? hk.body_a : MR_INVALID_INDEX, pa, neg(d), lb ? hk.body_b : MR_INVALID_INDEX, pb, d, rs);
        }
        __syncthreads();
        clk.mark(mr::kStageResponses);
        // ------------------------------------------ IR evaluation + blocks
        // Lanes own DoFs for every articulation dot product (canonical fold);
        // per-constraint scalars are then evaluated lane-parallel over
        // constraints. target/winv carry rel and J M^-1 J^T between the two
        // passes.
        const uint32_t mnv = A.row_stride;  // >= max(1, max_nv)
        const bool dof_lane = lane < nv;
        float vj = dof_lane ? e.vcand[lane] : 0.0f;
        // Jacobian / response rows of constraint c for this DoF lane; the
        // loops below fetch constraint c + 1 while working on c. Rows of
        // constraints without an articulation endpoint are never written and
        // their (unused) values are ignored.
        // Branch-free on purpose: every lane loads (lanes past the stride read the
        // row's last element, ignored), so the wait counters can leave the
        // prefetch of the next constraint in flight.
        const uint32_t row_lane = lane < mnv ? lane : mnv - 1u;
        auto load_rows = [&](uint32_t c, float* j3, float* r3) {
            const float* jb = e.jrow + (size_t)c * 3u * mnv + row_lane;
            const float* rb = e.resp + (size_t)c * 3u * mnv + row_lane;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                j3[r] = jb[(size_t)r * mnv];
                r3[r] = rb[(size_t)r * mnv];
            }
        };
        // Pass 1, every lane, articulated constraints only: the DoF-lane folds
        // of J v (rows 0-2) and of the 3x3 block J M^-1 J^T, staged in the
        // record. Constraints between scene bodies need no fold and are left
        // to the constraint-lane pass below, which also adds every constraint's
        // scene endpoint terms.
        float nj[3] = {0.0f, 0.0f, 0.0f}, nr[3] = {0.0f, 0.0f, 0.0f};
        if (ncons > 0u) load_rows(0u, nj, nr);
        for (uint32_t c = 0; c < ncons; ++c) {
            const float jr[3] = {nj[0], nj[1], nj[2]};
            const float rr[3] = {nr[0], nr[1], nr[2]};
            if (c + 1u < ncons) load_rows(c + 1u, nj, nr);
            if ((uniform_u32(e.cpair[c] >> 16) & mr::kHotArticulated) == 0u) continue;
            float f[12];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                f[r] = dof_lane ? jr[r] * vj : 0.0f;
#pragma unroll
                for (int q = 0; q < 3; ++q) f[3 + 3 * r + q] = dof_lane ? jr[r] * rr[q] : 0.0f;
            }
            wave_fold_n(f);
            if (lane == 0u) {
                HotConstraint& hw = e.hot[c];
                hw.stage[0] = f[0];
                hw.stage[1] = f[1];
                hw.stage[2] = f[2];
#pragma unroll
                for (int k = 0; k < 9; ++k) hw.winv[k] = f[3 + k];
            }
        }
        __syncthreads();
        const float slop = cfg.contact_slop;
        const float maxdep = cfg.max_depenetration_velocity;
        for (uint32_t c = lane; c < ncons; c += kLanes) {
            HotConstraint hk = e.hot[c];
            const ColdConstraint ck = e.cold[c];
            if ((hk.kind & mr::kHotArticulated) == 0u) {
                for (int r = 0; r < 3; ++r) hk.stage[r] = 0.0f;
                for (int k = 0; k < 9; ++k) hk.winv[k] = 0.0f;
            }
            if ((hk.kind & (mr::kHotDynamicA | mr::kHotDynamicB)) != 0u) {
                V3 va, vb;
                scene_point_velocities(e, hk.kind, hk.body_a, hk.body_b, V3{hk.point[0], hk.point[1], hk.point[2]}, va, vb);
                for (int r = 0; r < 3; ++r) hk.stage[r] = rel_row(hk.kind, hot_dir(hk, r), va, vb, hk.stage[r]);
            }
            const float rel[3] = {hk.stage[0], hk.stage[1], hk.stage[2]};
            const float slip = sqrt_rn(rel[1] * rel[1] + rel[2] * rel[2]);
            const bool stat = slip <= 1.0e-3f;
            hk.mu_u = stat ? ck.friction[0] : ck.friction[2];
            hk.mu_v = stat ? ck.friction[1] : ck.friction[3];
            const float src_target = ck.separation > 0.0f ? -ck.separation / h : 0.0f;
            const float tau = 2.0f * h;
            const float err = fmind(ck.separation + slop, 0.0f);
            const float ratio = h / tau;
            const float den = 1.0f + 2.0f * 1.0f * ratio + ratio * ratio;
            float stab = -h * err / (tau * tau * den);
            stab = clampd(fmaxd(stab, 0.0f), 0.0f, maxdep);
            if ((hk.flags & MR_CONSTRAINT_FLAG_NEW_IMPACT) != 0u) {
                const float incoming = rel[0] - src_target;
                if (incoming < -ck.restitution_threshold) stab = fmaxd(stab, -ck.restitution * incoming);
            }
            hk.target0 = src_target + stab;
            hk.reg0 = fmaxd(ck.compliance / (h * h) + ck.dissipation / h, 0.0f);
            bool ok = true;
            if (!isfin(hk.target0) || !isfin(hk.reg0)) {
                report_failure(&s_fail, c, MR_STEP_NONFINITE_RESULT, c);
                ok = false;
            }
            if (ok) {
                const V3 x{hk.point[0], hk.point[1], hk.point[2]};
                const bool sa_dyn = (hk.kind & mr::kHotDynamicA) != 0u, sb_dyn = (hk.kind & mr::kHotDynamicB) != 0u;
                float W[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) {
                        float val = hk.winv[3 * r + q];
                        if (sa_dyn) val += dot(hot_dir(hk, r), scene_response(e.scene[hk.body_a - nl], x, hot_dir(hk, q)));
                        if (sb_dyn) val += dot(hot_dir(hk, r), scene_response(e.scene[hk.body_b - nl], x, hot_dir(hk, q)));
                        if (r == q) val += r == 0 ? hk.reg0 : 0.0f;  // tangent rows carry none
                        W[r][q] = val;
                    }
                if (!invert3x3(W, hk.winv)) report_failure(&s_fail, c, MR_STEP_FACTORIZATION_FAILED, c);
            }
            float lam[3] = {e.lam[3u * c], e.lam[3u * c + 1u], e.lam[3u * c + 2u]};
            project_cone(lam, hk.mu_u, hk.mu_v, e.a->cfg.maximum_normal_impulse);
            e.lam[3u * c] = lam[0];
            e.lam[3u * c + 1u] = lam[1];
            e.lam[3u * c + 2u] = lam[2];
            e.hot[c] = hk;
        }
        __syncthreads();
        clk.mark(mr::kStageEvaluation);
        if (s_fail != ~0ull) break;
        // -------------------------------------- learned warm start (MLP)
        const uint32_t base_iters = cfg.velocity_iterations + (final_sub ? cfg.final_velocity_iterations : 0u);
        const bool adaptive = cfg.residual_tolerance > 0.0f;
        const uint32_t max_iters = adaptive ? cfg.max_iterations : base_iters;
        float sweep_delta = 0.0f;
        uint32_t iters = 0, sweeps_run = 0;
        bool solve_failed = false;
        const SolveBlock* solve = reinterpret_cast<const SolveBlock*>(e.hot);
        const uint32_t hot_words = (uint32_t)(sizeof(HotConstraint) / sizeof(SolveBlock));
        bool mlp_now = false;
        float* capture_rows = nullptr;
        const float vj_free = vj;
        if constexpr (kMlp) {
            mlp_now = A.mlp_enabled != 0u && ncons > 0u;
            if (A.mlp_samples != nullptr && substep == cfg.capture_substep)
                capture_rows = A.mlp_samples + (size_t)env * ccap * mr::kMlpSampleFloats;
            if (mlp_now || capture_rows != nullptr)
                mlp_now = mlp_warm_start(e, lane, ncons, mnv, mlp_now, capture_rows) && mlp_now;
        }
        // The plain path runs the attempt below once. When the MLP replaced any
        // warm start, a microstep whose sweeps miss the residual tolerance is
        // solved again from the plain warm start and the same free velocity.
        for (uint32_t attempt = 0;; ++attempt) {
            sweep_delta = 0.0f;
            iters = 0;
            if (cfg.solver == MR_SOLVER_RELAXED_JACOBI) {
                // ---------------------------------- relaxed block Jacobi
                // Lanes own contacts for the proposals, DoFs and scene bodies for
                // the velocity update (jacobi_apply).
                // damping per contact: diag / sum_j |W_nn(c, j)| clamped to [1/32, 1] (:6722-6798)
                for (uint32_t c = lane; c < ncons; c += kLanes) {
                    float sum = 0.0f, diag = 0.0f;
                    for (uint32_t j = 0; j < ncons; ++j) {
                        const float w = fabs_(jacobi_normal_cross(e, c, j, mnv));
                        sum = sum + w;
                        if (j == c) diag = w + e.hot[c].reg0;
                    }
                    e.jrelax[c] = sum > fmaxd(diag, kMatrixFloor) ? clampd(diag / sum, 1.0f / 32.0f, 1.0f) : 1.0f;
                }
                // warm start: the projected warm impulses are the first deltas (:6402-6470)
                for (uint32_t i = lane; i < 3u * ncons; i += kLanes) e.jdelta[i] = e.lam[i];
                __syncthreads();
                jacobi_apply(e, lane, ncons, mnv, dof_lane, vj);
                __syncthreads();
                clk.mark(mr::kStageWarmStart);
                float largest = 0.0f;
                for (uint32_t it = 0; it < max_iters && ncons > 0u; ++it) {
                    if (dof_lane) e.vcand[lane] = vj;
                    __syncthreads();
                    float my_delta = 0.0f;
                    uint32_t my_fail = MR_INVALID_INDEX;
                    for (uint32_t c = lane; c < ncons; c += kLanes) {
                        const SolveBlock sb = solve[(size_t)c * hot_words];
                        const uint32_t kind = e.cpair[c] >> 16;
                        const HotConstraint& hk = e.hot[c];
                        float rel[3] = {0.0f, 0.0f, 0.0f};
                        if ((kind & mr::kHotArticulated) != 0u) {
                            const float* jb = e.jrow + (size_t)c * 3u * mnv;
                            for (int r = 0; r < 3; ++r) rel[r] = serial_dot(jb + (size_t)r * mnv, e.vcand, nv);
                        }
                        if ((kind & (mr::kHotDynamicA | mr::kHotDynamicB)) != 0u) {
                            V3 va, vb;
                            scene_point_velocities(e, kind, hk.body_a, hk.body_b, V3{hk.point[0], hk.point[1], hk.point[2]},
                                                   va, vb);
                            for (int r = 0; r < 3; ++r) rel[r] = rel_row(kind, hot_dir(hk, r), va, vb, rel[r]);
                        }
                        const float lam[3] = {e.lam[3u * c], e.lam[3u * c + 1u], e.lam[3u * c + 2u]};
                        const float tgt[3] = {sb.target0, 0.0f, 0.0f}, reg[3] = {sb.reg0, 0.0f, 0.0f};
                        float rhs[3];
                        for (int r = 0; r < 3; ++r) rhs[r] = (tgt[r] - rel[r]) - reg[r] * lam[r];
                        float prop[3] = {lam[0], lam[1], lam[2]};
                        for (int r = 0; r < 3; ++r)
                            for (int q = 0; q < 3; ++q) prop[r] += sb.winv[3 * r + q] * rhs[q];
                        project_cone(prop, sb.mu_u, sb.mu_v, e.a->cfg.maximum_normal_impulse);
                        const float relax = e.jrelax[c];
                        float d[3], cand[3];
                        for (int r = 0; r < 3; ++r) {
                            d[r] = relax * (prop[r] - lam[r]);
                            cand[r] = lam[r] + d[r];
                        }
                        if (!isfin(cand[0]) || !isfin(cand[1]) || !isfin(cand[2]) || !isfin(d[0]) || !isfin(d[1]) ||
                            !isfin(d[2])) {
                            my_fail = c < my_fail ? c : my_fail;
                            continue;
                        }
                        for (int r = 0; r < 3; ++r) {
                            e.lam[3u * c + r] = cand[r];
                            e.jdelta[3u * c + r] = d[r];
                        }
                        my_delta = fmaxd(my_delta, fmaxd(fabs_(d[0]), fmaxd(fabs_(d[1]), fabs_(d[2]))));
                    }
                    const uint32_t first_fail = wave_min_u32(my_fail);
                    if (first_fail != MR_INVALID_INDEX) {
                        if (lane == 0u) s_fail = ((unsigned long long)MR_STEP_NONFINITE_RESULT << 24) | (first_fail & 0xffffffu);
                        solve_failed = true;
                        break;
                    }
                    __syncthreads();
                    jacobi_apply(e, lane, ncons, mnv, dof_lane, vj);
                    __syncthreads();
                    const float it_delta = wave_max(my_delta);
                    largest = fmaxd(largest, it_delta);
                    ++iters;
                    if (adaptive && it_delta <= cfg.residual_tolerance) break;
                }
                sweep_delta = largest;
            } else {
            // ---------------------------------------------------- warm start
            // DoF lanes accumulate the warm-start impulses in constraint order;
            // scene lanes apply theirs in the same order.
            if (dof_lane) {
                for (uint32_t c = 0; c < ncons; ++c) {
                    if (((e.cpair[c] >> 16) & mr::kHotArticulated) == 0u) continue;
                    const float* rb = e.resp + (size_t)c * 3u * mnv + lane;
                    const float* lc = e.lam + 3u * c;
                    vj += (rb[0] * lc[0] + rb[mnv] * lc[1]) + rb[2u * mnv] * lc[2];
                }
            }
            for (uint32_t s = lane; s < nsc; s += kLanes) {
                const uint32_t b = nl + s;
                if (!is_dynamic_scene(e, b)) continue;
                MRBodyStateC st = e.scene[s];
                for (uint32_t c = 0; c < ncons; ++c) {
                    const uint32_t pr = e.cpair[c];
                    if ((pr & 0xffu) != b && ((pr >> 8) & 0xffu) != b) continue;
                    const HotConstraint& hk = e.hot[c];
                    const float* lc = e.lam + 3u * c;
                    const V3 imp = add(add(mul(lc[0], hot_dir(hk, 0)), mul(lc[1], hot_dir(hk, 1))), mul(lc[2], hot_dir(hk, 2)));
                    const V3 x{hk.point[0], hk.point[1], hk.point[2]};
                    if (hk.body_a == b) apply_scene_impulse(st, x, neg(imp));
                    if (hk.body_b == b) apply_scene_impulse(st, x, imp);
                }
                e.scene[s] = st;
            }
            __syncthreads();
            clk.mark(mr::kStageWarmStart);
            // ------------------------------------------ ordered block sweeps
            // Every lane runs the constraint loop; the block update is uniform,
            // DoF lanes keep the candidate velocity in a register, and the next
            // constraint's Jacobian / response rows are prefetched. The loop is
            // compiled twice: a substep whose contacts all lack a dynamic scene
            // endpoint (an articulation on static geometry) runs it without any
            // scene work; the others also prefetch the next contact's geometric
            // half (bodies, directions, point), spread over lanes 0-15 of one
            // register, word k in lane k, and read back as scalars when used.
            bool scene_lane = false;
            for (uint32_t c = lane; c < ncons; c += kLanes)
                scene_lane |= ((e.cpair[c] >> 16) & (mr::kHotDynamicA | mr::kHotDynamicB)) != 0u;
            const bool any_scene = wave_any(scene_lane);
            auto ordered_sweeps = [&](auto scene_tag) {
                constexpr bool kScene = decltype(scene_tag)::value;
                const uint32_t* geo_words = reinterpret_cast<const uint32_t*>(e.hot) + sizeof(SolveBlock) / 4u;
                const uint32_t hot_u32 = (uint32_t)(sizeof(HotConstraint) / 4u);
                auto load_geo = [&](uint32_t c) -> uint32_t {
                    return lane < sizeof(GeoBlock) / 4u ? geo_words[(size_t)c * hot_u32 + lane] : 0u;
                };
                uint32_t gnext = 0u;
                if (ncons > 0u) {
                    load_rows(0u, nj, nr);
                    if constexpr (kScene) gnext = load_geo(0u);
                }
                for (uint32_t it = 0; it < max_iters && ncons > 0u; ++it) {
                    sweep_delta = 0.0f;
                    for (uint32_t c = 0; c < ncons; ++c) {
                        const float jr[3] = {nj[0], nj[1], nj[2]};
                        const float rr[3] = {nr[0], nr[1], nr[2]};
                        // loaded here, not a contact ahead: the fold below covers
                        // its latency, and a loop-carried copy costs 16 registers
                        const SolveBlock sb = solve[(size_t)c * hot_words];
                        [[maybe_unused]] const uint32_t gb = gnext;
                        const uint32_t cn = c + 1u < ncons ? c + 1u : 0u;
                        load_rows(cn, nj, nr);
                        if constexpr (kScene) gnext = load_geo(cn);
                        const float lam0 = e.lam[3u * c], lam1 = e.lam[3u * c + 1u], lam2 = e.lam[3u * c + 2u];
                        const uint32_t kind = uniform_u32(e.cpair[c] >> 16);
                        const bool art = (kind & mr::kHotArticulated) != 0u;
                        const bool sa_dyn = (kind & mr::kHotDynamicA) != 0u, sb_dyn = (kind & mr::kHotDynamicB) != 0u;
                        float rel[3] = {0.0f, 0.0f, 0.0f};
                        if (art) {
        #pragma unroll
                            for (int r = 0; r < 3; ++r) rel[r] = dof_lane ? jr[r] * vj : 0.0f;
                            wave_fold_n(rel);
                        }
                        V3 x{0.0f, 0.0f, 0.0f};
                        uint32_t ba = 0u, bb = 0u;
                        V3 dirs[3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
                        if (kScene && (sa_dyn || sb_dyn)) {
                            auto word = [&](uint32_t k) { return (uint32_t)__builtin_amdgcn_readlane((int)gb, (int)k); };
                            auto wordf = [&](uint32_t k) { return __uint_as_float(word(k)); };
                            constexpr uint32_t kDir = offsetof(GeoBlock, dir) / 4u, kPoint = offsetof(GeoBlock, point) / 4u;
                            ba = word(offsetof(GeoBlock, body_a) / 4u);
                            bb = word(offsetof(GeoBlock, body_b) / 4u);
                            x = V3{wordf(kPoint), wordf(kPoint + 1u), wordf(kPoint + 2u)};
                            V3 va, vb;
                            scene_point_velocities(e, kind, ba, bb, x, va, vb);
        #pragma unroll
                            for (int r = 0; r < 3; ++r) {
                                dirs[r] = V3{wordf(kDir + 3u * r), wordf(kDir + 3u * r + 1u), wordf(kDir + 3u * r + 2u)};
                                rel[r] = rel_row(kind, dirs[r], va, vb, rel[r]);
                            }
                        }
                        const float lamc[3] = {lam0, lam1, lam2};
                        // the tangent rows' stored target and regularisation are +0
                        const float tgt[3] = {sb.target0, 0.0f, 0.0f}, reg[3] = {sb.reg0, 0.0f, 0.0f};
                        float rhs[3];
                        for (int r = 0; r < 3; ++r) rhs[r] = (tgt[r] - rel[r]) - reg[r] * lamc[r];
                        float cand[3] = {lam0, lam1, lam2};
                        for (int r = 0; r < 3; ++r)
                            for (int q = 0; q < 3; ++q) cand[r] += sb.winv[3 * r + q] * rhs[q];
                        project_cone(cand, sb.mu_u, sb.mu_v, e.a->cfg.maximum_normal_impulse);
                        const float d0 = cand[0] - lam0, d1 = cand[1] - lam1, d2 = cand[2] - lam2;
                        if (wave_any(!isfin(cand[0]) | !isfin(cand[1]) | !isfin(cand[2]) | !isfin(d0) | !isfin(d1) | !isfin(d2))) {
                            if (lane == 0u) s_fail = ((unsigned long long)MR_STEP_NONFINITE_RESULT << 24) | (c & 0xffffffu);
                            solve_failed = true;
                            break;
                        }
                        sweep_delta = fmaxd(sweep_delta, fmaxd(fabs_(d0), fmaxd(fabs_(d1), fabs_(d2))));
                        if (art && dof_lane) vj += (rr[0] * d0 + rr[1] * d1) + rr[2] * d2;
                        // The workgroup is one wavefront and its LDS operations
                        // complete in issue order, so lane 0's writes below follow
                        // every lane's reads above and precede the next constraint's
                        // reads; the wave barrier only stops the compiler from
                        // moving memory operations across these points.
                        __builtin_amdgcn_wave_barrier();
                        if (lane == 0u) {
                            if (kScene && (sa_dyn || sb_dyn)) {
                                const V3 imp = add(add(mul(d0, dirs[0]), mul(d1, dirs[1])), mul(d2, dirs[2]));
                                if (sa_dyn) apply_scene_impulse(e.scene[ba - nl], x, neg(imp));
                                if (sb_dyn) apply_scene_impulse(e.scene[bb - nl], x, imp);
                            }
                            e.lam[3u * c] = cand[0];
                            e.lam[3u * c + 1u] = cand[1];
                            e.lam[3u * c + 2u] = cand[2];
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                    if (solve_failed) break;
                    ++iters;
                    if (adaptive && sweep_delta <= cfg.residual_tolerance) break;
                    __syncthreads();
                }
            };
            if (any_scene) ordered_sweeps(BoolTag<true>{});
            else ordered_sweeps(BoolTag<false>{});
            }
            __syncthreads();
            clk.mark(mr::kStageSweeps);
            sweeps_run += iters;
            if constexpr (kMlp) {
                if (mlp_now && attempt == 0u && adaptive && !solve_failed) {
                    const bool missed = !(sweep_delta <= cfg.residual_tolerance);
                    bool diverged = false;
                    if (!missed && A.mlp.max_normal_residual > 0.0f) {
                        // The learned solve reached the tolerance: its normal
                        // residuals, computed as the residual pass below does,
                        // must also stay within max_normal_residual.
                        float* res_n = e.resid;
                        for (uint32_t c = 0; c < ncons; ++c) {
                            if ((uniform_u32(e.cpair[c] >> 16) & mr::kHotArticulated) == 0u) continue;
                            const float j0 = e.jrow[(size_t)c * 3u * mnv + row_lane];
                            const float a0 = wave_fold(dof_lane ? j0 * vj : 0.0f);
                            if (lane == 0u) res_n[c] = a0;
                        }
                        __syncthreads();
                        bool over = false;
                        for (uint32_t c = lane; c < ncons; c += kLanes) {
                            const HotConstraint hk = e.hot[c];
                            float rel0 = (hk.kind & mr::kHotArticulated) != 0u ? res_n[c] : 0.0f;
                            if ((hk.kind & (mr::kHotDynamicA | mr::kHotDynamicB)) != 0u) {
                                V3 va, vb;
                                scene_point_velocities(e, hk.kind, hk.body_a, hk.body_b,
                                                       V3{hk.point[0], hk.point[1], hk.point[2]}, va, vb);
                                rel0 = rel_row(hk.kind, hot_dir(hk, 0), va, vb, rel0);
                            }
                            const float l0 = e.lam[3u * c];
                            const float eq = (rel0 - hk.target0) + hk.reg0 * l0;
                            const float nres = l0 > kConeEpsilon ? fabs_(eq) : fmaxd(-eq, 0.0f);
                            over = over || !(nres <= A.mlp.max_normal_residual);
                        }
                        diverged = __ballot(over) != 0ull;
                        __syncthreads();
                    }
                    if (missed || diverged) {
                        vj = vj_free;
                        for (uint32_t i = lane; i < 3u * ncons; i += kLanes) e.lam[i] = e.mlp_lam[i];
                        for (uint32_t b = lane; b < nsc; b += kLanes) e.scene[b] = e.mlp_scene[b];
                        if (lane == 0u) {
                            cs.diagnostics.z += 1.0f;
                            if (diverged) cs.diagnostics.w += 1.0f;
                        }
                        __syncthreads();
                        continue;
                    }
                }
            }
            break;
        }
        if constexpr (kMlp) {
            // capture targets: the impulse the solve converged to
            if (capture_rows != nullptr && !solve_failed) {
                for (uint32_t c = lane; c < ccap; c += kLanes) {
                    float* row = capture_rows + (size_t)c * mr::kMlpSampleFloats;
                    if (c < ncons) {
                        row[8] = e.lam[3u * c];
                        row[9] = e.lam[3u * c + 1u];
                        row[10] = e.lam[3u * c + 2u];
                    }
                    row[11] = c < ncons ? 1.0f : 0.0f;
                }
            }
        }
        if (s_fail != ~0ull) break;
        // ------------------------------------- residuals + impulse store
        // Pass 1, every lane, articulated constraints only: the DoF-lane fold
        // of the normal row's J v.
        float* res_n = e.resid;                 // [ccap]: articulated J v, then the normal residual
        float* res_c = e.resid + ccap;          // [ccap]: cone violation
        float nj0 = ncons > 0u ? e.jrow[row_lane] : 0.0f;
        for (uint32_t c = 0; c < ncons; ++c) {
            const float j0 = nj0;
            if (c + 1u < ncons) nj0 = e.jrow[(size_t)(c + 1u) * 3u * mnv + row_lane];
            if ((uniform_u32(e.cpair[c] >> 16) & mr::kHotArticulated) == 0u) continue;
            const float a0 = wave_fold(dof_lane ? j0 * vj : 0.0f);
            if (lane == 0u) res_n[c] = a0;
        }
        __syncthreads();
        // Pass 2, constraint lanes: the dynamic scene endpoints' terms of the
        // normal-row relative velocity, residual terms and the impulse store.
        for (uint32_t c = lane; c < ncons; c += kLanes) {
            const HotConstraint hk = e.hot[c];
            float rel0 = (hk.kind & mr::kHotArticulated) != 0u ? res_n[c] : 0.0f;
            if ((hk.kind & (mr::kHotDynamicA | mr::kHotDynamicB)) != 0u) {
                V3 va, vb;
                scene_point_velocities(e, hk.kind, hk.body_a, hk.body_b, V3{hk.point[0], hk.point[1], hk.point[2]}, va, vb);
                rel0 = rel_row(hk.kind, hot_dir(hk, 0), va, vb, rel0);
            }
            const float lamr[3] = {e.lam[3u * c], e.lam[3u * c + 1u], e.lam[3u * c + 2u]};
            const float eq = (rel0 - hk.target0) + hk.reg0 * lamr[0];
            const float nres = lamr[0] > kConeEpsilon ? fabs_(eq) : fmaxd(-eq, 0.0f);
            const float lu = hk.mu_u * lamr[0], lv = hk.mu_v * lamr[0];
            float viol;
            if (lu > 0.0f && lv > 0.0f)
                viol = fmaxd(sqrt_rn((lamr[1] * lamr[1]) / (lu * lu) + (lamr[2] * lamr[2]) / (lv * lv)) - 1.0f, 0.0f);
            else
                viol = sqrt_rn(lamr[1] * lamr[1] + lamr[2] * lamr[2]);
            res_n[c] = nres;
            res_c[c] = viol;
            const ColdConstraint& ck = e.cold[c];
            new_p[(size_t)ck.manifold * MR_MANIFOLD_POINT_CAPACITY + ck.point].impulses =
                mr_float4{lamr[0], lamr[1], lamr[2], kCcd ? h : 0.0f};
        }
        __syncthreads();
        // Pass 3, uniform: maxima in constraint order (the oracle's order,
        // including its handling of a non-finite term).
        float max_normal = 0.0f, max_cone = 0.0f;
        for (uint32_t c = 0; c < ncons; ++c) {
            max_normal = fmaxd(max_normal, res_n[c]);
            max_cone = fmaxd(max_cone, res_c[c]);
        }
        if (dof_lane) e.vcand[lane] = vj;
        __syncthreads();
        if (!isfin(sweep_delta) || !isfin(max_normal) || !isfin(max_cone)) {
            if (lane == 0u) s_fail = ((unsigned long long)MR_STEP_NONFINITE_RESULT << 24) | 0xffffffu;
            __syncthreads();
            break;
        }
        cs.solver_iterations = ncons > 0u ? iters : 0u;
        cs.residuals = mr_float4{ncons > 0u ? sweep_delta : 0.0f, max_normal, max_cone,
                                 cs.residuals.w + (float)(ncons > 0u ? sweeps_run : 0u)};
        if (lane == 0u) {
            uint8_t* parent = e.isl;                      // LDS, dead response scratch
            uint8_t* used = e.isl + (MR_MAX_BODIES + 1u);
            const uint32_t nodes = 1u + nsc;
            for (uint32_t i = 0; i < nodes; ++i) {
                parent[i] = (uint8_t)i;
                used[i] = 0u;
            }
            for (uint32_t c = 0; c < ncons; ++c) {
                const uint32_t ba = e.cpair[c] & 0xffu, bb = (e.cpair[c] >> 8) & 0xffu;
                const uint32_t na = is_link(e, ba) ? 0u : (is_dynamic_scene(e, ba) ? 1u + (ba - nl) : MR_INVALID_INDEX);
                const uint32_t nb = is_link(e, bb) ? 0u : (is_dynamic_scene(e, bb) ? 1u + (bb - nl) : MR_INVALID_INDEX);
                if (na != MR_INVALID_INDEX) used[na] = true;
                if (nb != MR_INVALID_INDEX) used[nb] = true;
                if (na != MR_INVALID_INDEX && nb != MR_INVALID_INDEX) {
                    uint32_t ra = na, rb = nb;
                    while (parent[ra] != ra) ra = parent[ra];
                    while (parent[rb] != rb) rb = parent[rb];
                    if (ra != rb) {
                        if (ra < rb) parent[rb] = ra;
                        else parent[ra] = rb;
                    }
                }
            }
            uint32_t islands = 0;
            for (uint32_t i = 0; i < nodes; ++i) {
                uint32_t rr = i;
                while (parent[rr] != rr) rr = parent[rr];
                if (used[i] && rr == i) ++islands;
            }
            s_count[0] = islands;
        }
        // ------------------------------------------------- joint limits
        for (uint32_t j = lane; j < nv; j += kLanes) {
            const MRDofC& dof = e.dofs[j];
            float vel = e.vcand[j];
            if ((dof.flags & MR_DOF_FLAG_VELOCITY_LIMIT) != 0u) vel = clampd(vel, -dof.limits.z, dof.limits.z);
            if ((dof.flags & MR_DOF_FLAG_POSITION_LIMIT) != 0u && dof.q_index != MR_INVALID_INDEX) {
                const float pos = e.q[dof.q_index];
                vel = clampd(vel, (dof.limits.x - pos) / h, (dof.limits.y - pos) / h);
            }
            if (!isfin(vel)) report_failure(&s_fail, j, MR_STEP_NONFINITE_RESULT, j);
            e.vcand[j] = vel;
        }
        __syncthreads();
        cs.islands = s_count[0];
        clk.mark(mr::kStageResiduals);
        if (s_fail != ~0ull) break;
        // ---------------------------------------------------- integration
        // New coordinates are staged in registers / s_root and written after
        // a barrier, because other lanes still read q in this phase.
        float qnew = 0.0f;
        uint32_t qslot = MR_INVALID_INDEX;
        if (lane == 0u && floating) {
            const Q4 inc = rotation_increment(V3{h * e.vcand[3], h * e.vcand[4], h * e.vcand[5]});
            Q4 o;
            if (!normalized_quat(qmul(inc, Q4{e.q[3], e.q[4], e.q[5], e.q[6]}), o)) {
                report_failure(&s_fail, 0u, MR_STEP_NONFINITE_RESULT, 3u);
                o = Q4{e.q[3], e.q[4], e.q[5], e.q[6]};
            }
            for (uint32_t i = 0; i < 3u; ++i) {
                s_root[i] = e.q[i] + h * e.vcand[i];
                if (!isfin(s_root[i])) report_failure(&s_fail, 1u + i, MR_STEP_NONFINITE_RESULT, i);
            }
            s_root[3] = o.x;
            s_root[4] = o.y;
            s_root[5] = o.z;
            s_root[6] = o.w;
        }
        for (uint32_t i = 1u + lane; i < nl; i += kLanes) {
            const MRBodyC& b = e.bodies[i];
            if (b.q_index != MR_INVALID_INDEX) {
                qslot = b.q_index;
                qnew = e.q[b.q_index] + h * e.vcand[b.v_index];
                if (!isfin(qnew)) report_failure(&s_fail, 1u + b.q_index, MR_STEP_NONFINITE_RESULT, b.q_index);
            }
        }
        for (uint32_t s = lane; s < nsc; s += kLanes) {
            const uint32_t b = nl + s;
            const MRBodyC& bp = e.bodies[b];
            if (bp.motion_type != MR_MOTION_DYNAMIC) continue;
            MRBodyStateC st = e.scene[s];
            V3 lin = f3(st.linear_velocity_inv_mass);
            V3 ang = f3(st.angular_velocity);
            const float ll = length(lin), al = length(ang);
            if (bp.damping.z > 0.0f && ll > bp.damping.z) lin = mul(bp.damping.z / ll, lin);
            if (bp.damping.w > 0.0f && al > bp.damping.w) ang = mul(bp.damping.w / al, ang);
            st.linear_velocity_inv_mass.x = lin.x;
            st.linear_velocity_inv_mass.y = lin.y;
            st.linear_velocity_inv_mass.z = lin.z;
            st.angular_velocity.x = ang.x;
            st.angular_velocity.y = ang.y;
            st.angular_velocity.z = ang.z;
            const V3 np = add(f3(st.position), mul(h, lin));
            st.position = mr_float4{np.x, np.y, np.z, 1.0f};
            const Q4 inc = rotation_increment(mul(h, ang));
            Q4 o;
            if (!normalized_quat(qmul(inc, q4(st.orientation)), o) || !finite3(np)) {
                report_failure(&s_fail, 0x10000u + b, MR_STEP_NONFINITE_RESULT, b);
                continue;
            }
            st.orientation = mr_float4{o.x, o.y, o.z, o.w};
            world_inverse_inertia(bp, o, st);
            e.scene[s] = st;
        }
        __syncthreads();
        if (s_fail != ~0ull) break;
        if (qslot != MR_INVALID_INDEX) e.q[qslot] = qnew;
        if (lane == 0u && floating)
            for (uint32_t i = 0; i < 7u; ++i) e.q[i] = s_root[i];
        for (uint32_t j = lane; j < nv; j += kLanes) e.v[j] = e.vcand[j];
        __syncthreads();
        if (s_fail != ~0ull) break;
        // --------------------------------------------- microstep commit
        if (lane == 0u) A.mcount[(size_t)dst_buf * E + env] = nman;
        mcount_src = nman;
        src_buf = dst_buf;
        dst_buf = dst_buf == (cur + 1u) % 3u ? alt_buf : (cur + 1u) % 3u;
        if (!impact) successful = substep + 1u;
        min_piv = fmind(min_piv, s_fmax[0]);
        max_piv = fmaxd(max_piv, s_fmax[1]);
        float my_acc = 0.0f;
        for (uint32_t j = lane; j < nv; j += kLanes) my_acc = fmaxd(my_acc, fabs_(e.accel[j]));
        max_acc = fmaxd(max_acc, wave_max(my_acc));
        __syncthreads();
        clk.mark(mr::kStageIntegration);
        if constexpr (kCcd) {
            if (impact) {
                remaining -= h;
                ++segment;
                ++impact_segments;  // counted once the segment commits
                continue;
            }
            remaining = A.h;
            segment = 0u;
        }
        ++substep;
    }
    if constexpr (kCcd) e.h = A.h;

    // ------------------------------------------- transactional commit
    __syncthreads();
    const unsigned long long fk = s_fail;
    const bool failed = fk != ~0ull;
    const uint32_t code = failed ? (uint32_t)((fk >> 24) & 0xffull) : MR_STEP_SUCCESS;
    uint32_t findex = failed ? (uint32_t)(fk & 0xffffffull) : MR_INVALID_INDEX;
    if (findex == 0xffffffu) findex = MR_INVALID_INDEX;
    if (failed) failing_substep = successful;
    if (!failed) {
        for (uint32_t i = lane; i < nq; i += kLanes) gq[i] = e.q[i];
        for (uint32_t j = lane; j < nv; j += kLanes) gv[j] = e.v[j];
        for (uint32_t s = lane; s < nsc; s += kLanes) gs[s] = e.scene[s];
        if (A.acceleration != nullptr)
            for (uint32_t j = lane; j < nv; j += kLanes) A.acceleration[(size_t)env * A.max_nv + j] = e.accel[j];
    } else if (A.acceleration != nullptr) {
        for (uint32_t j = lane; j < nv; j += kLanes) A.acceleration[(size_t)env * A.max_nv + j] = 0.0f;
    }
    if (lane == 0u) {
        if (!failed && successful > 0u) A.mcurrent[env] = src_buf;
        MRWorldStatusC ws;
        ws.code = code;
        ws.environment = env;
        ws.control_step = A.control_step + step;
        ws.successful_substeps = successful;
        ws.aba_code = 0u;
        ws.failing_substep = failed ? failing_substep : MR_INVALID_INDEX;
        ws.failing_index = findex;
        ws.flags = 0u;
        if (!failed) {
            float qerr = 0.0f;
            if (floating) {
                const float n2 = e.q[3] * e.q[3] + e.q[4] * e.q[4] + e.q[5] * e.q[5] + e.q[6] * e.q[6];
                qerr = fabs_(sqrt_rn(n2) - 1.0f);
            }
            ws.diagnostics = mr_float4{isfin(min_piv) ? min_piv : 0.0f, max_piv, max_acc, qerr};
        } else {
            ws.diagnostics = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        A.status[step_rec] = ws;
        if constexpr (kCcd) cs.diagnostics.w = (float)impact_segments;
        cs.code = code;
        if (failed) {
            if (code == MR_STEP_PAIR_CAPACITY_OVERFLOW) cs.first_failing_pair = findex;
            else cs.first_failing_constraint = findex;
        }
        A.cstatus[step_rec] = cs;
        // wall-clock duration of this control step of this world (the GPU's constant-rate clock)
        if (A.step_us != nullptr)
            A.step_us[step_rec] = A.us_per_tick * (float)(__builtin_amdgcn_s_memrealtime() - s_t0);
    }
}

// Two wavefronts per SIMD: the register budget is capped at 256 so that two
// worlds share each SIMD and hide each other's memory and LDS latency (the
// per-world stages are latency-bound; one wave per SIMD left it idle).
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
mr_control_step_kernel(KernelArgs args) {
    control_step<false>(args, blockIdx.x);
}

// Persistent work-queue instance: a fixed grid of workgroups, each taking the
// next environment of args.queue (in the queue's order) until the queue is
// drained, so the order (heaviest first, say) and the number of resident
// workgroups are the caller's choice. The head is zeroed by
// mr_queue_reset_kernel before every launch.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
mr_control_step_queue_kernel(KernelArgs args) {
    __shared__ uint32_t s_next;
    for (;;) {
        if (threadIdx.x == 0u) s_next = atomicAdd(args.queue_head, 1u);
        __syncthreads();
        const uint32_t idx = s_next;
        __syncthreads();
        if (idx >= args.queue_count) break;
        control_step<false>(args, args.queue[idx]);
        __syncthreads();
    }
}

// Persistent horizon instance: the work queue of mr_control_step_queue_kernel,
// where each entry is a world advanced through all args.horizon control steps
// back to back (step t's command, reset mask and status records at [t][E]).
// A world's next control step starts as soon as its previous one ends, so
// the launch is not held to the slowest world of every step; each world's
// steps run in order on one workgroup, so its results are those of one
// launch per control step.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
mr_control_horizon_queue_kernel(KernelArgs args) {
    __shared__ uint32_t s_next;
    for (;;) {
        if (threadIdx.x == 0u) s_next = atomicAdd(args.queue_head, 1u);
        __syncthreads();
        const uint32_t idx = s_next;
        __syncthreads();
        if (idx >= args.queue_count) break;
        const uint32_t env = args.queue[idx];
        for (uint32_t t = 0; t < args.horizon; ++t) {
            control_step<false>(args, env, t);
            __syncthreads();
        }
    }
}

extern "C" __global__ void mr_queue_reset_kernel(uint32_t* head) {
    if (threadIdx.x == 0u && blockIdx.x == 0u) *head = 0u;
}

// Work-queue order for the next control step: the environments of `status`
// (one control step's contact statuses) sorted by the cone sweeps that step
// took (residuals.w), most first, by a counting sort over 1024 sweep bins in
// one workgroup. Only the schedule depends on it: which workgroup takes which
// world never changes a world's results. Environments with equal counts land
// in any order (atomic placement).
//
// With steps > 1 (after a persistent horizon launch, status holding [steps][E]
// records) the key is each world's mean sweeps per control step over the
// horizon, so the worlds that cost most over the whole horizon start first
// (longest-processing-time-first order for the next horizon).
extern "C" __global__ void __launch_bounds__(1024)
mr_queue_order_kernel(const MRContactStatusC* status, uint32_t environments, uint32_t steps, uint32_t* order) {
    __shared__ uint32_t bins[1024];
    __shared__ uint32_t wave_base[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    bins[t] = 0u;
    __syncthreads();
    auto key_of = [&](uint32_t e) {
        float sum = 0.0f;
        for (uint32_t s = 0; s < steps; ++s) sum += status[(size_t)s * environments + e].residuals.w;
        const float w = fminf(fmaxf(sum / (float)steps, 0.0f), 1023.0f);  // NaN -> 0
        return 1023u - (uint32_t)w;
    };
    for (uint32_t e = t; e < environments; e += 1024u) atomicAdd(&bins[key_of(e)], 1u);
    __syncthreads();
    // exclusive prefix over the bins: wave scans, then the wave totals
    const uint32_t v = bins[t];
    uint32_t x = v;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63u) wave_base[wave] = x;
    __syncthreads();
    if (t == 0u) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < 16u; ++w) {
            const uint32_t c = wave_base[w];
            wave_base[w] = run;
            run += c;
        }
    }
    __syncthreads();
    bins[t] = (x - v) + wave_base[wave];
    __syncthreads();
    for (uint32_t e = t; e < environments; e += 1024u) order[atomicAdd(&bins[key_of(e)], 1u)] = e;
}

// The learned-warm-start instance: the MLP tile keeps its accumulators and
// operands in registers next to the solver state, so it runs one wavefront
// per SIMD with the full register file.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
mr_control_step_mlp_kernel(KernelArgs args) {
    control_step<true>(args, blockIdx.x);
}

// The hybrid CCD mode: one workgroup per environment, microsteps split at
// times of impact. A separate instance, so the speculative kernels carry none
// of its registers. Two wavefronts per SIMD: at one, the compiler moves state
// into AGPRs instead of spilling, and the chain7+boxes hybrid workload ran
// 19 % slower (502k against 597k env-steps/s at 4096 worlds).
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
mr_control_step_ccd_kernel(KernelArgs args) {
    control_step<false, true>(args, blockIdx.x);
}

// Worlds with convex hulls or triangle meshes: one workgroup per
// environment, with the GJK / MPR / patch and BVH narrowphase (speculative
// and hybrid CCD modes). Separate instances, so the other kernels carry none
// of its private arrays.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
mr_control_step_geo_kernel(KernelArgs args) {
    control_step<false, false, true>(args, blockIdx.x);
}

extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
mr_control_step_ccd_geo_kernel(KernelArgs args) {
    control_step<false, true, true>(args, blockIdx.x);
}

// Advance summary: one thread clears it, then a grid-stride pass over the
// [steps][environments] status records folds it with per-wave reductions and
// one atomic per wave (kernels rather than a memset, so the sequence is also
// valid inside a captured graph). With a subset work-queue schedule only the
// listed environments were stepped, so only their records are folded:
// `order` lists them (null for every environment).
extern "C" __global__ void mr_summary_clear_kernel(mr::AdvanceSummaryDev* s) {
    if (threadIdx.x == 0u && blockIdx.x == 0u) {
        s->first_failure = ~0ull;
        s->failed = 0u;
        s->max_active_pairs = 0u;
        s->max_constraints = 0u;
        s->pad = 0u;
    }
}

extern "C" __global__ void __launch_bounds__(256)
mr_summary_reduce_kernel(const MRWorldStatusC* status, const MRContactStatusC* cstatus, uint64_t steps,
                         uint32_t environments, const uint32_t* order, uint32_t order_count,
                         mr::AdvanceSummaryDev* s) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t failed = 0u, pairs = 0u, cons = 0u;
    unsigned long long first = ~0ull;
    const uint32_t stepped = order != nullptr ? order_count : environments;
    const uint64_t count = steps * stepped;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t step = i / stepped;
        const uint64_t env = order != nullptr ? order[i % stepped] : i % stepped;
        const uint64_t rec = step * environments + env;
        const uint32_t code = status[rec].code;
        if (code != MR_STEP_SUCCESS) {
            ++failed;
            const unsigned long long key = (step << 40) | (env << 8) | (unsigned long long)(code & 0xffu);
            first = key < first ? key : first;
        }
        const MRContactStatusC& c = cstatus[rec];
        pairs = c.max_active_pairs > pairs ? c.max_active_pairs : pairs;
        cons = c.max_constraints > cons ? c.max_constraints : cons;
    }
    failed = wave_sum_u32(failed);
    pairs = wave_max_u32(pairs);
    cons = wave_max_u32(cons);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = ((unsigned long long)__shfl_xor((int)(first >> 32), d, 64) << 32) |
                                     (unsigned long long)(uint32_t)__shfl_xor((int)(first & 0xffffffffull), d, 64);
        first = o < first ? o : first;
    }
    if (lane == 0u) {
        if (failed != 0u) atomicAdd(&s->failed, failed);
        atomicMax(&s->max_active_pairs, pairs);
        atomicMax(&s->max_constraints, cons);
        if (first != ~0ull) atomicMin(&s->first_failure, first);
    }
}

namespace mr {

hipError_t launch_advance_summary(const MRWorldStatusC* status, const MRContactStatusC* cstatus, uint64_t steps,
                                  uint32_t environments, const uint32_t* order, uint32_t order_count,
                                  AdvanceSummaryDev* summary, hipStream_t stream) {
    hipLaunchKernelGGL(mr_summary_clear_kernel, dim3(1), dim3(64), 0, stream, summary);
    const uint64_t count = steps * (order != nullptr ? order_count : environments);
    const uint64_t blocks = (count + 255u) / 256u;
    const uint32_t grid = (uint32_t)(blocks < 1024u ? (blocks > 0u ? blocks : 1u) : 1024u);
    hipLaunchKernelGGL(mr_summary_reduce_kernel, dim3(grid), dim3(256), 0, stream, status, cstatus, steps,
                       environments, order, order_count, summary);
    return hipGetLastError();
}

static_assert(sizeof(SpI) == 84, "SpI layout");
static_assert(sizeof(Collider) == 124, "Collider layout");
static_assert(sizeof(HotConstraint) == 128, "HotConstraint layout");
static_assert(sizeof(SolveBlock) == 64 && sizeof(HotConstraint) % sizeof(SolveBlock) == 0, "SolveBlock layout");
static_assert(sizeof(GeoBlock) == 64 && offsetof(HotConstraint, body_a) == sizeof(SolveBlock), "GeoBlock layout");

hipError_t launch_control_step(const KernelArgs& args, uint32_t lds_bytes, hipStream_t stream, bool mlp,
                               uint32_t queue_grid) {
    if (queue_grid > 0u && args.horizon > 0u) {
        hipLaunchKernelGGL(mr_queue_reset_kernel, dim3(1), dim3(64), 0, stream, args.queue_head);
        hipLaunchKernelGGL(mr_control_horizon_queue_kernel, dim3(queue_grid), dim3(64), lds_bytes, stream, args);
        return hipGetLastError();
    }
    if (queue_grid > 0u) {
        hipLaunchKernelGGL(mr_queue_reset_kernel, dim3(1), dim3(64), 0, stream, args.queue_head);
        hipLaunchKernelGGL(mr_control_step_queue_kernel, dim3(queue_grid), dim3(64), lds_bytes, stream, args);
        return hipGetLastError();
    }
    if (mlp)
        hipLaunchKernelGGL(mr_control_step_mlp_kernel, dim3(args.environment_count), dim3(64), lds_bytes, stream, args);
    else if (args.geometry != 0u && args.cfg.ccd_mode == MR_CCD_HYBRID)
        hipLaunchKernelGGL(mr_control_step_ccd_geo_kernel, dim3(args.environment_count), dim3(64), lds_bytes, stream,
                           args);
    else if (args.geometry != 0u)
        hipLaunchKernelGGL(mr_control_step_geo_kernel, dim3(args.environment_count), dim3(64), lds_bytes, stream, args);
    else if (args.cfg.ccd_mode == MR_CCD_HYBRID)
        hipLaunchKernelGGL(mr_control_step_ccd_kernel, dim3(args.environment_count), dim3(64), lds_bytes, stream, args);
    else
        hipLaunchKernelGGL(mr_control_step_kernel, dim3(args.environment_count), dim3(64), lds_bytes, stream, args);
    return hipGetLastError();
}

// Scratch of the allocation-free entry point (mr_world_enqueue_device_static):
// one file-scope device array, which every GPU the library runs on holds at
// its own address until the process ends. Host code never passes the array's
// name to a kernel (on the host it is a stand-in, not a device address); it
// passes the address static_scratch() looked up for the current device.
static __device__ __align__(256) unsigned char scratch[MR_WORLD_STATIC_SCRATCH_BYTES];

// Device address of `scratch` on the current device: looked up once per
// device with hipGetSymbolAddress, then served from the cache.
hipError_t static_scratch(void** out, size_t* bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex guard;
    static void* address[kMaxDevices] = {};
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(guard);
    if (address[device] == nullptr) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (ptr == nullptr) return hipErrorInvalidSymbol;
        address[device] = ptr;
    }
    *out = address[device];
    *bytes = MR_WORLD_STATIC_SCRATCH_BYTES;
    return hipSuccess;
}

// Spins one thread until the constant-rate clock (s_memrealtime) has advanced
// `ticks`, or `max_iterations` reads have passed, and reports the ticks it saw;
// timed with events on the host, it gives the clock's rate.
extern "C" __global__ void mr_clock_spin_kernel(unsigned long long ticks, unsigned long long max_iterations,
                                                unsigned long long* seen) {
    if (threadIdx.x != 0u || blockIdx.x != 0u) return;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    unsigned long long t = t0;
    for (unsigned long long i = 0; i < max_iterations && t - t0 < ticks; ++i) t = __builtin_amdgcn_s_memrealtime();
    *seen = t - t0;
}

// Microseconds per tick of s_memrealtime on the current device, measured once
// per device: a spin of 2^20 ticks (a few milliseconds at the rates GPUs use)
// timed with events.
hipError_t clock_us_per_tick(float* out) {
    constexpr int kMaxDevices = 64;
    static std::mutex guard;
    static float cached[kMaxDevices] = {};
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(guard);
    if (cached[device] == 0.0f) {
        unsigned long long* seen = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        e = hipMalloc(&seen, sizeof(unsigned long long));
        if (e == hipSuccess) e = hipEventCreate(&a);
        if (e == hipSuccess) e = hipEventCreate(&b);
        float ms = 0.0f;
        unsigned long long ticks = 0ull;
        if (e == hipSuccess) {
            (void)hipEventRecord(a, nullptr);
            hipLaunchKernelGGL(mr_clock_spin_kernel, dim3(1), dim3(64), 0, nullptr, 1ull << 20, 1ull << 26, seen);
            (void)hipEventRecord(b, nullptr);
            e = hipEventSynchronize(b);
        }
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess) e = hipMemcpy(&ticks, seen, sizeof(ticks), hipMemcpyDeviceToHost);
        if (a != nullptr) (void)hipEventDestroy(a);
        if (b != nullptr) (void)hipEventDestroy(b);
        if (seen != nullptr) (void)hipFree(seen);
        if (e != hipSuccess) return e;
        if (ticks == 0ull || !(ms > 0.0f)) return hipErrorUnknown;
        cached[device] = 1000.0f * ms / (float)ticks;
    }
    *out = cached[device];
    return hipSuccess;
}

hipError_t launch_queue_order(const MRContactStatusC* status, uint32_t environments, uint32_t steps, uint32_t* order,
                              hipStream_t stream) {
    hipLaunchKernelGGL(mr_queue_order_kernel, dim3(1), dim3(1024), 0, stream, status, environments, steps, order);
    return hipGetLastError();
}

// Workgroups of the plain or work-queue control-step kernel one CU holds.
hipError_t control_step_occupancy(uint32_t lds_bytes, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                        reinterpret_cast<const void*>(&mr_control_step_queue_kernel),
                                                        64, lds_bytes);
}

// The dynamic-LDS limit is an attribute of the kernel function, shared by
// every world in the process: it only ever grows to the largest world's need.
hipError_t configure_control_step(uint32_t lds_bytes, uint32_t mlp_lds_bytes) {
    static std::mutex guard;
    static uint32_t plain_max = 0u, mlp_max = 0u;
    std::lock_guard<std::mutex> lock(guard);
    if (lds_bytes > plain_max) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mr_control_step_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mr_control_step_queue_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mr_control_horizon_queue_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
        for (const void* k : {reinterpret_cast<const void*>(&mr_control_step_ccd_kernel),
                              reinterpret_cast<const void*>(&mr_control_step_geo_kernel),
                              reinterpret_cast<const void*>(&mr_control_step_ccd_geo_kernel)}) {
            e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
            if (e != hipSuccess) return e;
        }
        plain_max = lds_bytes;
    }
    if (mlp_lds_bytes > mlp_max) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mr_control_step_mlp_kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlp_lds_bytes);
        if (e != hipSuccess) return e;
        mlp_max = mlp_lds_bytes;
    }
    return hipSuccess;
}

} // namespace mr
