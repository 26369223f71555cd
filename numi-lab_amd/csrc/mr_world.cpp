// Host executor: device arenas, LDS layout, per-control-step launches and
// status publication for the fused temporalCone kernel.
#include "mr_world.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace mr {

hipError_t launch_control_step(const KernelArgs& args, uint32_t lds_bytes, hipStream_t stream, bool mlp,
                               uint32_t queue_grid);
hipError_t control_step_occupancy(uint32_t lds_bytes, int* blocks_per_cu);
hipError_t launch_queue_order(const MRContactStatusC* status, uint32_t environments, uint32_t steps, uint32_t* order,
                              hipStream_t stream);
hipError_t configure_control_step(uint32_t lds_bytes, uint32_t mlp_lds_bytes);
hipError_t launch_advance_summary(const MRWorldStatusC* status, const MRContactStatusC* cstatus, uint64_t steps,
                                  uint32_t environments, const uint32_t* order, uint32_t order_count,
                                  AdvanceSummaryDev* summary, hipStream_t stream);
hipError_t static_scratch(void** out, size_t* bytes);
hipError_t clock_us_per_tick(float* out);

namespace {
constexpr uint32_t kSpIBytes = 84u;
constexpr uint32_t kColliderBytes = 124u;
constexpr uint32_t kMaxLdsBytes = 160u * 1024u;

template <typename T>
void upload(DeviceBuffer& b, const std::vector<T>& v) {
    const size_t n = std::max<size_t>(v.size() * sizeof(T), 16u);
    b.reserve(n);
    if (!v.empty()) hip_check(hipMemcpy(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
}
} // namespace

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

void DeviceBuffer::reserve(size_t n) {
    if (n <= bytes) return;
    release();
    hip_check(hipMalloc(&ptr, n), "hipMalloc");
    bytes = n;
}

void DeviceBuffer::release() {
    if (ptr != nullptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

World::World(const ModelData& model, const MRWorldStepConfigC& config) : model_(model), cfg_(config) {
    if (cfg_.physics_substeps == 0u || cfg_.physics_substeps > MR_MAX_PHYSICS_SUBSTEPS)
        throw std::runtime_error("physics_substeps must be in [1, 64]");
    if (!(cfg_.control_timestep > 0.0f)) throw std::runtime_error("control_timestep must be positive");
    if (cfg_.velocity_iterations == 0u && cfg_.residual_tolerance <= 0.0f)
        throw std::runtime_error("velocity_iterations must be positive");
    if (cfg_.residual_tolerance > 0.0f && cfg_.max_iterations == 0u)
        throw std::runtime_error("max_iterations must be positive with a residual tolerance");
    if (cfg_.solver != MR_SOLVER_ORDERED_SWEEPS && cfg_.solver != MR_SOLVER_RELAXED_JACOBI)
        throw std::runtime_error("unknown solver mode");
    if (!(cfg_.maximum_normal_impulse >= 0.0f) || !std::isfinite(cfg_.maximum_normal_impulse))
        throw std::runtime_error("maximum_normal_impulse must be finite and non-negative (0: no cap)");
    if (cfg_.ccd_mode != MR_CCD_SPECULATIVE && cfg_.ccd_mode != MR_CCD_HYBRID) throw std::runtime_error("unknown ccd_mode");
    if (cfg_.ccd_mode == MR_CCD_HYBRID) {
        if (cfg_.max_ccd_events == 0u || cfg_.max_ccd_events > MR_CCD_MAX_EVENTS)
            throw std::runtime_error("max_ccd_events must be in [1, 16] in the hybrid CCD mode");
        if (cfg_.max_ccd_iterations == 0u || cfg_.max_ccd_iterations > 256u)
            throw std::runtime_error("max_ccd_iterations must be in [1, 256] in the hybrid CCD mode");
        for (float x : {cfg_.ccd_distance_tolerance, cfg_.ccd_time_tolerance, cfg_.ccd_minimum_advance})
            if (!(x > 0.0f) || !std::isfinite(x))
                throw std::runtime_error("CCD tolerances and the minimum advance must be finite and positive");
    }
    if (cfg_.pair_capacity == 0u || cfg_.pair_capacity > 4096u) throw std::runtime_error("pair_capacity in [1, 4096]");
    if (cfg_.manifold_capacity == 0u || cfg_.manifold_capacity > 4096u)
        throw std::runtime_error("manifold_capacity in [1, 4096]");
    if (cfg_.constraint_capacity == 0u || cfg_.constraint_capacity > 4096u)
        throw std::runtime_error("constraint_capacity in [1, 4096]");
    if (cfg_.capture_evidence != 0u && cfg_.capture_substep >= cfg_.physics_substeps)
        throw std::runtime_error("capture_substep must be below physics_substeps");
    finalize_model(model_, cfg_.control_timestep / (float)cfg_.physics_substeps);
    for (const MRShapeC& sh : model_.shapes)
        if (sh.shape_type == MR_SHAPE_CONVEX || sh.shape_type == MR_SHAPE_TRIANGLE_MESH)
            throw std::runtime_error(
                "convex and mesh shapes: the device narrowphase is switched off (its kernel instance took a memory "
                "aperture violation on the GPU whose cause is not yet shown); the CPU oracle implements them");
    env_ = model_.environment_count();
    hip_check(hipGetDevice(&device_), "hipGetDevice");

    compute_lds_layout();
    hip_check(configure_control_step(lds_.total, lds_mlp_.total), "hipFuncSetAttribute");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&ev_begin_), "hipEventCreate");
    hip_check(hipEventCreate(&ev_end_), "hipEventCreate");
    upload(d_models_, model_.models);
    upload(d_env_model_, model_.environment_model);
    {
        // outward response slots per model (the plan depends on the tree only)
        std::vector<uint8_t> plan(model_.models.size() * MR_MAX_BODIES, 0xffu);
        for (size_t m = 0; m < model_.models.size(); ++m) {
            const MRModelHeaderC& h = model_.models[m];
            const MRBodyC* b = model_.bodies.data() + h.body_offset;
            if (h.link_count > 0u)
                plan_response_slots([&](uint32_t i) { return b[i].parent; }, h.link_count, &plan[m * MR_MAX_BODIES]);
        }
        upload(d_slot_plan_, plan);
    }
    upload(d_bodies_, model_.bodies);
    upload(d_dofs_, model_.dofs);
    upload(d_actuators_, model_.actuators);
    upload(d_shapes_, model_.shapes);
    upload(d_materials_, model_.materials);
    upload(d_pairs_, model_.pairs);
    upload(d_heights_, model_.heights);
    upload(d_geometries_, model_.geometries);
    upload(d_geometry_vertices_, model_.geometry_vertices);
    upload(d_triangles_, model_.triangles);
    upload(d_bvh_nodes_, model_.bvh_nodes);
    for (const MRShapeC& sh : model_.shapes)
        geometry_ = geometry_ || sh.shape_type == MR_SHAPE_CONVEX || sh.shape_type == MR_SHAPE_TRIANGLE_MESH;
    const size_t E = env_;
    d_q_.reserve(E * model_.max_nq * sizeof(float) + 16u);
    d_v_.reserve(E * model_.max_nv * sizeof(float) + 16u);
    d_scene_.reserve(E * std::max(1u, model_.max_scene_bodies) * sizeof(MRBodyStateC));
    d_reset_q_.reserve(d_q_.bytes);
    d_reset_v_.reserve(d_v_.bytes);
    d_reset_scene_.reserve(d_scene_.bytes);
    const size_t mcap = cfg_.manifold_capacity, pcap = cfg_.pair_capacity, ccap = cfg_.constraint_capacity;
    d_mh_.reserve(3u * E * mcap * sizeof(MRManifoldHeaderC));
    d_mp_.reserve(3u * E * mcap * MR_MANIFOLD_POINT_CAPACITY * sizeof(MRManifoldPointC));
    d_mcount_.reserve(3u * E * sizeof(uint32_t));
    d_mcurrent_.reserve(E * sizeof(uint32_t));
    hip_check(hipMemset(d_mcount_.ptr, 0, 3u * E * sizeof(uint32_t)), "hipMemset");
    hip_check(hipMemset(d_mcurrent_.ptr, 0, E * sizeof(uint32_t)), "hipMemset");
    d_stage_h_.reserve(E * pcap * sizeof(MRManifoldHeaderC));
    d_stage_p_.reserve(E * pcap * MR_MANIFOLD_POINT_CAPACITY * sizeof(MRManifoldPointC));
    const size_t row_stride = std::max(1u, model_.max_nv);
    d_jrow_.reserve(E * ccap * 3u * row_stride * sizeof(float));
    d_resp_.reserve(E * ccap * 3u * row_stride * sizeof(float));
    d_cold_.reserve(E * ccap * sizeof(ColdConstraint));
    d_hot_.reserve(E * ccap * sizeof(HotConstraint));
    d_accel_.reserve(E * model_.max_nv * sizeof(float) + 16u);
    hip_check(hipMemset(d_accel_.ptr, 0, d_accel_.bytes), "hipMemset");
    d_ev_pairs_.reserve(E * pcap * sizeof(MRPairRecordC));
    d_ev_counts_.reserve(E * sizeof(uint32_t));
    hip_check(hipMemset(d_ev_counts_.ptr, 0, E * sizeof(uint32_t)), "hipMemset");
    if (!model_.initial_q.empty()) {
        set_state(model_.initial_q.data(), model_.initial_v.data(), model_.initial_scene.data());
        set_reset_state(model_.initial_q.data(), model_.initial_v.data(), model_.initial_scene.data());
    }
}

World::~World() {
    if (stream_ != nullptr) (void)hipStreamSynchronize(stream_);
    drop_graph();
    if (stream_ != nullptr) {
        (void)hipStreamDestroy(stream_);
    }
    if (ev_begin_ != nullptr) (void)hipEventDestroy(ev_begin_);
    if (ev_end_ != nullptr) (void)hipEventDestroy(ev_end_);
    for (hipEvent_t ev : step_events_) (void)hipEventDestroy(ev);
}

void World::set_step_timing(bool enabled) {
    // per-world step times are stamped on the constant-rate clock, whose rate is measured once per device
    if (enabled && us_per_tick_ == 0.0f) hip_check(clock_us_per_tick(&us_per_tick_), "clock rate");
    step_timing_ = enabled;
}

void World::set_stage_profiling(bool enabled) {
    stage_profiling_ = enabled;
    if (enabled) d_stage_cycles_.reserve((size_t)env_ * kStageCount * sizeof(unsigned long long));
}

void World::read_stage_cycles(unsigned long long* out) const {
    if (!stage_profiling_) throw std::runtime_error("stage profiling is not enabled");
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(out, d_stage_cycles_.ptr, (size_t)env_ * kStageCount * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost),
              "read stage cycles");
}

void World::read_step_milliseconds(double* out, uint32_t steps) const {
    if (last_horizon_) throw std::runtime_error("the last advance ran as one horizon launch: no per-launch times");
    if (!step_timing_ || steps > last_steps_ || step_events_.size() < (size_t)steps + 1u)
        throw std::runtime_error("step timing was not enabled for the last advance");
    hip_check(hipStreamSynchronize(stream_), "sync");
    for (uint32_t s = 0; s < steps; ++s) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, step_events_[s], step_events_[s + 1u]), "event elapsed");
        out[s] = ms;
    }
}

// Per-lane response scratch every model needs: outward slots (same plan as
// the kernel) and the inward joint-term stack, bounded by the largest number
// of DoF links on the union of two root paths.
void World::response_scratch_bounds(uint32_t& slots, uint32_t& stack) const {
    for (const MRModelHeaderC& h : model_.models) {
        const MRBodyC* b = model_.bodies.data() + h.body_offset;
        const uint32_t n = h.link_count;
        uint8_t plan[MR_MAX_BODIES];
        slots = std::max(slots, plan_response_slots([&](uint32_t i) { return b[i].parent; }, n, plan));
        unsigned long long anc[MR_MAX_BODIES] = {};
        unsigned long long dof_links = 0ull;
        for (uint32_t i = 1; i < n; ++i) {
            anc[i] = anc[b[i].parent] | (1ull << i);
            if (b[i].v_index != MR_INVALID_INDEX) dof_links |= 1ull << i;
        }
        for (uint32_t a = 0; a < n; ++a)
            for (uint32_t c = a; c < n; ++c)
                stack = std::max(stack, (uint32_t)__builtin_popcountll((anc[a] | anc[c]) & dof_links));
    }
}

void World::compute_lds_layout() {
    uint32_t links = 1u, shapes = 1u;
    for (const MRModelHeaderC& h : model_.models) {
        links = std::max(links, h.link_count);
        shapes = std::max(shapes, h.shape_count);
    }
    const uint32_t nq = std::max(1u, model_.max_nq), nv = std::max(1u, model_.max_nv);
    const uint32_t nscene = std::max(1u, model_.max_scene_bodies);
    uint32_t off = 0;
    auto take = [&](uint32_t bytes, uint32_t align) {
        off = (off + align - 1u) / align * align;
        const uint32_t at = off;
        off += bytes;
        return at;
    };
    LdsLayout& L = lds_;
    // arrays live for the whole microstep
    L.topo = take((uint32_t)sizeof(TopoCache), 16u);
    L.lconst = take((uint32_t)sizeof(LinkConst) * links, 4u);
    L.arm = take(4u * nv, 4u);
    L.q = take(4u * nq, 4u);
    L.v = take(4u * nv, 4u);
    L.vcand = take(4u * nv, 4u);
    L.effort = take(4u * nv, 4u);
    L.accel = take(4u * nv, 4u);
    L.cmd = take(4u * nv, 4u);
    L.pos = take(12u * links, 4u);
    L.rot = take(16u * links, 4u);
    L.sa = take(12u * links, 4u);
    L.sl = take(12u * links, 4u);
    L.r = take(12u * links, 4u);
    L.Ut = take(12u * links, 4u);
    L.Ub = take(12u * links, 4u);
    L.D = take(4u * links, 4u);
    L.L = take(4u * 36u, 4u);
    L.scene = take((uint32_t)sizeof(MRBodyStateC) * nscene, 16u);
    L.lam = take(12u * cfg_.constraint_capacity, 4u);
    L.cpair = take(4u * cfg_.constraint_capacity, 4u);
    // Three phase-local regions share one union: the kinematics / ABA
    // temporaries (dead once the accelerations exist), the collision arrays
    // (dead after emission) and the response scratch.
    const uint32_t union_start = take(0u, 16u);
    L.w = take(12u * links, 4u);
    L.vl = take(12u * links, 4u);
    L.ca = take(12u * links, 4u);
    L.cl = take(12u * links, 4u);
    L.I = take(kSpIBytes * links, 4u);
    L.pt = take(12u * links, 4u);
    L.pb = take(12u * links, 4u);
    L.u = take(4u * links, 4u);
    L.alpha = take(12u * links, 4u);
    L.acc = take(12u * links, 4u);
    uint32_t union_end = off;
    off = union_start;
    L.col = take(kColliderBytes * shapes, 4u);
    L.active = take(4u * cfg_.pair_capacity, 4u);
    L.npts = take(4u * cfg_.pair_capacity, 4u);
    L.moff = take(4u * cfg_.pair_capacity, 4u);
    L.okey = take(4u * cfg_.manifold_capacity, 4u);
    // the heightfield query table exists only where some model has a heightfield
    bool heightfield = false;
    for (const MRShapeC& sh : model_.shapes) heightfield = heightfield || sh.shape_type == MR_SHAPE_HEIGHTFIELD;
    const uint32_t hf_slots = heightfield ? kHfSlots : 0u;
    L.hfitem = take(4u * hf_slots, 4u);
    L.hfq = take((uint32_t)sizeof(HfHit) * hf_slots, 16u);
    union_end = std::max(union_end, off);
    off = union_start;
    uint32_t slots = 1u, stack = 1u;
    response_scratch_bounds(slots, stack);
    L.rslot = take(4u * 6u * 64u * slots, 16u);
    L.rstack = take(4u * 64u * stack, 16u);
    union_end = std::max(union_end, off);
    off = union_start;
    L.isl = take(2u * (MR_MAX_BODIES + 1u), 4u);
    L.resid = take(8u * cfg_.constraint_capacity, 4u);
    const bool jacobi = cfg_.solver == MR_SOLVER_RELAXED_JACOBI;
    L.jrelax = take(jacobi ? 4u * cfg_.constraint_capacity : 0u, 4u);
    L.jdelta = take(jacobi ? 12u * cfg_.constraint_capacity : 0u, 4u);
    off = std::max(union_end, off);
    L.total = (off + 15u) / 16u * 16u;
    if (L.total > kMaxLdsBytes)
        throw std::runtime_error("world needs " + std::to_string(L.total) +
                                 " bytes of LDS per environment; lower pair/constraint capacity");
    // the learned-warm-start instance appends its regions
    LdsLayout& M = lds_mlp_;
    M = L;
    off = L.total;
    M.mlp = take(kMlpTileLdsBytes, 16u);
    M.mlp_lam = take(12u * cfg_.constraint_capacity, 16u);
    M.mlp_scene = take((uint32_t)sizeof(MRBodyStateC) * nscene, 16u);
    M.total = (off + 15u) / 16u * 16u;
    if (M.total > kMaxLdsBytes) M.total = 0u;  // the MLP instance is unavailable for this world
}

namespace {
template <size_t N>
size_t sum_bytes(const DeviceBuffer* const (&bufs)[N]) {
    size_t total = 0;
    for (const DeviceBuffer* b : bufs) total += b->bytes;
    return total;
}
} // namespace

size_t World::immutable_bytes() const {
    const DeviceBuffer* const all[] = {&d_models_, &d_env_model_, &d_slot_plan_, &d_bodies_, &d_dofs_,
                                       &d_actuators_, &d_shapes_, &d_materials_, &d_pairs_, &d_heights_,
                                       &d_geometries_, &d_geometry_vertices_, &d_triangles_, &d_bvh_nodes_};
    return sum_bytes(all);
}

size_t World::persistent_bytes() const {
    const DeviceBuffer* const all[] = {&d_q_, &d_v_, &d_scene_, &d_mh_, &d_mp_, &d_mcount_, &d_mcurrent_,
                                       &d_reset_q_, &d_reset_v_, &d_reset_scene_};
    return sum_bytes(all);
}

size_t World::transient_bytes() const {
    const DeviceBuffer* const all[] = {&d_commands_, &d_masks_, &d_status_, &d_cstatus_, &d_accel_, &d_ev_pairs_,
                                       &d_ev_counts_, &d_stage_h_, &d_stage_p_, &d_jrow_, &d_resp_, &d_cold_,
                                       &d_hot_, &d_summary_, &d_samples_, &d_stage_cycles_, &d_queue_,
                                       &d_queue_head_};
    return sum_bytes(all);
}

size_t World::retained_bytes() const { return immutable_bytes() + persistent_bytes() + transient_bytes(); }

void World::set_state(const float* q, const float* v, const MRBodyStateC* scene) {
    hip_check(hipStreamSynchronize(stream_), "sync");
    const size_t E = env_;
    hip_check(hipMemcpy(d_q_.ptr, q, E * model_.max_nq * sizeof(float), hipMemcpyHostToDevice), "set_state q");
    hip_check(hipMemcpy(d_v_.ptr, v, E * model_.max_nv * sizeof(float), hipMemcpyHostToDevice), "set_state v");
    if (model_.max_scene_bodies > 0u)
        hip_check(hipMemcpy(d_scene_.ptr, scene, E * model_.max_scene_bodies * sizeof(MRBodyStateC),
                            hipMemcpyHostToDevice),
                  "set_state scene");
    hip_check(hipMemset(d_mcount_.ptr, 0, 3u * E * sizeof(uint32_t)), "hipMemset");
    hip_check(hipMemset(d_mcurrent_.ptr, 0, E * sizeof(uint32_t)), "hipMemset");
}

void World::set_reset_state(const float* q, const float* v, const MRBodyStateC* scene) {
    hip_check(hipStreamSynchronize(stream_), "sync");
    const size_t E = env_;
    hip_check(hipMemcpy(d_reset_q_.ptr, q, E * model_.max_nq * sizeof(float), hipMemcpyHostToDevice), "reset q");
    hip_check(hipMemcpy(d_reset_v_.ptr, v, E * model_.max_nv * sizeof(float), hipMemcpyHostToDevice), "reset v");
    if (model_.max_scene_bodies > 0u)
        hip_check(hipMemcpy(d_reset_scene_.ptr, scene, E * model_.max_scene_bodies * sizeof(MRBodyStateC),
                            hipMemcpyHostToDevice),
                  "reset scene");
}

void World::set_manifolds(const MRManifoldHeaderC* h, const MRManifoldPointC* p, const uint32_t* count) {
    hip_check(hipStreamSynchronize(stream_), "sync");
    const size_t E = env_, mcap = cfg_.manifold_capacity;
    for (size_t e = 0; e < E; ++e)
        if (count[e] > mcap) throw std::runtime_error("manifold count above capacity");
    hip_check(hipMemset(d_mcurrent_.ptr, 0, E * sizeof(uint32_t)), "hipMemset");
    hip_check(hipMemcpy(d_mh_.ptr, h, E * mcap * sizeof(MRManifoldHeaderC), hipMemcpyHostToDevice), "set mh");
    hip_check(hipMemcpy(d_mp_.ptr, p, E * mcap * MR_MANIFOLD_POINT_CAPACITY * sizeof(MRManifoldPointC),
                        hipMemcpyHostToDevice),
              "set mp");
    hip_check(hipMemcpy(d_mcount_.ptr, count, E * sizeof(uint32_t), hipMemcpyHostToDevice), "set mcount");
}

void World::get_state(float* q, float* v, MRBodyStateC* scene) const {
    hip_check(hipStreamSynchronize(stream_), "sync");
    const size_t E = env_;
    if (q) hip_check(hipMemcpy(q, d_q_.ptr, E * model_.max_nq * sizeof(float), hipMemcpyDeviceToHost), "get q");
    if (v) hip_check(hipMemcpy(v, d_v_.ptr, E * model_.max_nv * sizeof(float), hipMemcpyDeviceToHost), "get v");
    if (scene && model_.max_scene_bodies > 0u)
        hip_check(hipMemcpy(scene, d_scene_.ptr, E * model_.max_scene_bodies * sizeof(MRBodyStateC),
                            hipMemcpyDeviceToHost),
                  "get scene");
}

void World::get_manifolds(MRManifoldHeaderC* h, MRManifoldPointC* p, uint32_t* count) const {
    hip_check(hipStreamSynchronize(stream_), "sync");
    const size_t E = env_, mcap = cfg_.manifold_capacity;
    std::vector<uint32_t> cur(E), counts(3u * E);
    hip_check(hipMemcpy(cur.data(), d_mcurrent_.ptr, E * sizeof(uint32_t), hipMemcpyDeviceToHost), "mcurrent");
    hip_check(hipMemcpy(counts.data(), d_mcount_.ptr, 3u * E * sizeof(uint32_t), hipMemcpyDeviceToHost), "mcount");
    const auto* bh = static_cast<const MRManifoldHeaderC*>(d_mh_.ptr);
    const auto* bp = static_cast<const MRManifoldPointC*>(d_mp_.ptr);
    for (size_t e = 0; e < E; ++e) {
        const size_t b = cur[e];
        count[e] = counts[b * E + e];
        hip_check(hipMemcpy(h + e * mcap, bh + (b * E + e) * mcap, mcap * sizeof(MRManifoldHeaderC),
                            hipMemcpyDeviceToHost),
                  "get mh");
        hip_check(hipMemcpy(p + e * mcap * MR_MANIFOLD_POINT_CAPACITY,
                            bp + (b * E + e) * mcap * MR_MANIFOLD_POINT_CAPACITY,
                            mcap * MR_MANIFOLD_POINT_CAPACITY * sizeof(MRManifoldPointC), hipMemcpyDeviceToHost),
                  "get mp");
    }
}

World::DeviceView World::device_view() const {
    DeviceView d;
    d.q = static_cast<float*>(d_q_.ptr);
    d.v = static_cast<float*>(d_v_.ptr);
    d.reset_q = static_cast<const float*>(d_reset_q_.ptr);
    d.reset_v = static_cast<const float*>(d_reset_v_.ptr);
    d.dofs = static_cast<const MRDofC*>(d_dofs_.ptr);
    return d;
}

KernelArgs World::base_args(uint32_t step) const {
    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.models = static_cast<const MRModelHeaderC*>(d_models_.ptr);
    a.response_slots = static_cast<const uint8_t*>(d_slot_plan_.ptr);
    a.environment_model = static_cast<const uint32_t*>(d_env_model_.ptr);
    a.bodies = static_cast<const MRBodyC*>(d_bodies_.ptr);
    a.dofs = static_cast<const MRDofC*>(d_dofs_.ptr);
    a.actuators = model_.actuators.empty() ? nullptr : static_cast<const MRActuatorProfileC*>(d_actuators_.ptr);
    a.shapes = static_cast<const MRShapeC*>(d_shapes_.ptr);
    a.materials = static_cast<const MRMaterialC*>(d_materials_.ptr);
    a.pairs = static_cast<const MRCompiledPairC*>(d_pairs_.ptr);
    a.heights = static_cast<const float*>(d_heights_.ptr);
    a.geometries = static_cast<const MRGeometryC*>(d_geometries_.ptr);
    a.geometry_vertices = static_cast<const mr_float4*>(d_geometry_vertices_.ptr);
    a.triangles = static_cast<const MRTriangleC*>(d_triangles_.ptr);
    a.bvh_nodes = static_cast<const MRBvhNodeC*>(d_bvh_nodes_.ptr);
    a.geometry = geometry_ ? 1u : 0u;
    a.heightfield = model_.heightfield;
    a.gravity[0] = model_.gravity.x;
    a.gravity[1] = model_.gravity.y;
    a.gravity[2] = model_.gravity.z;
    a.cfg = cfg_;
    a.h = cfg_.control_timestep / (float)cfg_.physics_substeps;
    a.environment_count = env_;
    a.max_nq = model_.max_nq;
    a.max_nv = model_.max_nv;
    a.row_stride = std::max(1u, model_.max_nv);
    a.max_scene = model_.max_scene_bodies;
    a.control_step = step;
    a.pair_capacity = cfg_.pair_capacity;
    a.manifold_capacity = cfg_.manifold_capacity;
    a.constraint_capacity = cfg_.constraint_capacity;
    a.q = static_cast<float*>(d_q_.ptr);
    a.v = static_cast<float*>(d_v_.ptr);
    a.scene = static_cast<MRBodyStateC*>(d_scene_.ptr);
    a.mh = static_cast<MRManifoldHeaderC*>(d_mh_.ptr);
    a.mp = static_cast<MRManifoldPointC*>(d_mp_.ptr);
    a.mcount = static_cast<uint32_t*>(d_mcount_.ptr);
    a.mcurrent = static_cast<uint32_t*>(d_mcurrent_.ptr);
    a.reset_q = static_cast<const float*>(d_reset_q_.ptr);
    a.reset_v = static_cast<const float*>(d_reset_v_.ptr);
    a.reset_scene = static_cast<const MRBodyStateC*>(d_reset_scene_.ptr);
    a.acceleration = static_cast<float*>(d_accel_.ptr);
    a.evidence_pairs = cfg_.capture_evidence != 0u ? static_cast<MRPairRecordC*>(d_ev_pairs_.ptr) : nullptr;
    a.evidence_counts = cfg_.capture_evidence != 0u ? static_cast<uint32_t*>(d_ev_counts_.ptr) : nullptr;
    a.stage_h = static_cast<MRManifoldHeaderC*>(d_stage_h_.ptr);
    a.stage_p = static_cast<MRManifoldPointC*>(d_stage_p_.ptr);
    a.jrow = static_cast<float*>(d_jrow_.ptr);
    a.resp = static_cast<float*>(d_resp_.ptr);
    a.cold = static_cast<ColdConstraint*>(d_cold_.ptr);
    a.hot = static_cast<HotConstraint*>(d_hot_.ptr);
    a.stage_cycles = stage_profiling_ ? static_cast<unsigned long long*>(d_stage_cycles_.ptr) : nullptr;
    a.lds = mlp_instance() ? lds_mlp_ : lds_;
    a.mlp = mlp_;
    a.mlp_enabled = mlp_enabled_ ? 1u : 0u;
    a.mlp_samples = mlp_capture_ ? static_cast<float*>(d_samples_.ptr) : nullptr;
    a.queue = static_cast<const uint32_t*>(d_queue_.ptr);
    a.queue_count = queue_count_;
    a.queue_head = static_cast<uint32_t*>(d_queue_head_.ptr);
    a.us_per_tick = us_per_tick_;
    return a;
}

void World::set_warm_start_mlp(const MRMlpWeightsC* weights) {
    if (weights == nullptr) {
        mlp_enabled_ = false;
        mlp_ = MRMlpWeightsC{};
        return;
    }
    if (!(weights->w1 && weights->b1 && weights->w2 && weights->b2 && weights->w3 && weights->b3 &&
          weights->in_shift && weights->in_scale && weights->out_scale))
        throw std::runtime_error("warm-start MLP: every weight pointer must be set");
    if (cfg_.solver != MR_SOLVER_ORDERED_SWEEPS)
        throw std::runtime_error("warm-start MLP: the learned warm start runs with the ordered sweeps solver");
    if (cfg_.ccd_mode != MR_CCD_SPECULATIVE)
        throw std::runtime_error("warm-start MLP: not available in the hybrid CCD mode");
    if (geometry_) throw std::runtime_error("warm-start MLP: not available in worlds with convex or mesh shapes");
    if (lds_mlp_.total == 0u)
        throw std::runtime_error("warm-start MLP: the world's LDS plus the MLP tile exceeds 160 KiB");
    if (queue_grid_ > 0u)
        throw std::runtime_error("warm-start MLP: runs one workgroup per environment; clear the work-queue "
                                 "schedule first (mr_world_set_schedule with grid 0)");
    if (cfg_.residual_tolerance <= 0.0f)
        throw std::runtime_error("warm-start MLP: needs the adaptive solve (residual_tolerance > 0) for its fallback");
    mlp_ = *weights;
    mlp_enabled_ = true;
}

void World::set_mlp_capture(bool enabled) {
    if (enabled) {
        if (lds_mlp_.total == 0u)
            throw std::runtime_error("MLP sample capture: the world's LDS plus the MLP tile exceeds 160 KiB");
        if (cfg_.ccd_mode != MR_CCD_SPECULATIVE)
            throw std::runtime_error("MLP sample capture: not available in the hybrid CCD mode");
        if (geometry_)
            throw std::runtime_error("MLP sample capture: not available in worlds with convex or mesh shapes");
        if (queue_grid_ > 0u)
            throw std::runtime_error("MLP sample capture: runs one workgroup per environment; clear the work-queue "
                                     "schedule first (mr_world_set_schedule with grid 0)");
        d_samples_.reserve((size_t)env_ * cfg_.constraint_capacity * kMlpSampleFloats * sizeof(float));
        hip_check(hipStreamSynchronize(stream_), "sync");
        hip_check(hipMemset(d_samples_.ptr, 0, d_samples_.bytes), "hipMemset");
    }
    mlp_capture_ = enabled;
}

void World::read_mlp_samples(float* out) const {
    if (d_samples_.ptr == nullptr) throw std::runtime_error("MLP sample capture was never enabled");
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(out, d_samples_.ptr, (size_t)env_ * cfg_.constraint_capacity * kMlpSampleFloats *
                                                sizeof(float), hipMemcpyDeviceToHost),
              "hipMemcpy samples");
}

World::StepBuffers World::own_buffers() {
    StepBuffers b;
    b.stage_h = static_cast<MRManifoldHeaderC*>(d_stage_h_.ptr);
    b.stage_p = static_cast<MRManifoldPointC*>(d_stage_p_.ptr);
    b.jrow = static_cast<float*>(d_jrow_.ptr);
    b.resp = static_cast<float*>(d_resp_.ptr);
    b.cold = static_cast<ColdConstraint*>(d_cold_.ptr);
    b.hot = static_cast<HotConstraint*>(d_hot_.ptr);
    b.status = static_cast<MRWorldStatusC*>(d_status_.ptr);
    b.cstatus = static_cast<MRContactStatusC*>(d_cstatus_.ptr);
    b.summary = static_cast<AdvanceSummaryDev*>(d_summary_.ptr);
    return b;
}

void World::launch_steps(const float* cmd, const uint32_t* msk, uint32_t steps, bool events, const StepBuffers& buf,
                         bool horizon) {
    const size_t E = env_;
    const size_t per_step = E * model_.max_nv;
    // per-world step times only on eager launches (a captured graph keeps no pointer to this buffer)
    float* step_us = events && step_timing_ && d_step_us_.bytes >= (size_t)steps * E * sizeof(float)
        ? static_cast<float*>(d_step_us_.ptr) : nullptr;
    if (events) hip_check(hipEventRecord(ev_begin_, stream_), "event");
    if (horizon) {
        // one launch: every world runs all steps back to back
        KernelArgs a = base_args(0u);
        a.stage_h = buf.stage_h;
        a.stage_p = buf.stage_p;
        a.jrow = buf.jrow;
        a.resp = buf.resp;
        a.cold = buf.cold;
        a.hot = buf.hot;
        a.command = cmd;
        a.reset_mask = msk;
        a.status = buf.status;
        a.cstatus = buf.cstatus;
        a.step_us = step_us;
        a.horizon = steps;
        hip_check(launch_control_step(a, lds_.total, stream_, false, queue_grid_), "control horizon launch");
        if (reorder_)
            hip_check(launch_queue_order(buf.cstatus, (uint32_t)E, steps, static_cast<uint32_t*>(d_queue_.ptr), stream_),
                      "queue order launch");
    } else {
        if (events && step_timing_) hip_check(hipEventRecord(step_events_[0], stream_), "event");
        for (uint32_t s = 0; s < steps; ++s) {
            KernelArgs a = base_args(s);
            a.stage_h = buf.stage_h;
            a.stage_p = buf.stage_p;
            a.jrow = buf.jrow;
            a.resp = buf.resp;
            a.cold = buf.cold;
            a.hot = buf.hot;
            a.command = cmd + (size_t)s * per_step;
            a.reset_mask = msk != nullptr ? msk + (size_t)s * E : nullptr;
            a.status = buf.status + (size_t)s * E;
            a.cstatus = buf.cstatus + (size_t)s * E;
            a.step_us = step_us != nullptr ? step_us + (size_t)s * E : nullptr;
            hip_check(launch_control_step(a, mlp_instance() ? lds_mlp_.total : lds_.total, stream_, mlp_instance(),
                                          mlp_instance() ? 0u : queue_grid_),
                      "control step launch");
            if (reorder_ && queue_grid_ > 0u)
                hip_check(launch_queue_order(a.cstatus, (uint32_t)E, 1u, static_cast<uint32_t*>(d_queue_.ptr), stream_),
                          "queue order launch");
            if (events && step_timing_) hip_check(hipEventRecord(step_events_[s + 1u], stream_), "event");
        }
    }
    if (events) hip_check(hipEventRecord(ev_end_, stream_), "event");
    // a subset work queue steps only the listed environments: fold only theirs
    const bool subset = !mlp_instance() && queue_grid_ > 0u && queue_count_ < env_;
    hip_check(launch_advance_summary(buf.status, buf.cstatus, (uint64_t)steps, (uint32_t)E,
                                     subset ? static_cast<const uint32_t*>(d_queue_.ptr) : nullptr,
                                     subset ? queue_count_ : 0u, buf.summary, stream_),
              "advance summary");
    last_stepped_ = subset ? queue_count_ : env_;
    last_status_ = buf.status;
    last_cstatus_ = buf.cstatus;
    last_summary_ = buf.summary;
    last_horizon_ = horizon;
}

void World::set_schedule(const uint32_t* order, uint32_t count, uint32_t grid) {
    hip_check(hipStreamSynchronize(stream_), "sync");
    reorder_ = false;
    if (grid == 0u) {
        if (order != nullptr) throw std::runtime_error("set_schedule: an order needs the work-queue schedule (grid > 0)");
        queue_grid_ = 0u;
        queue_count_ = 0u;
        ++schedule_version_;
        return;
    }
    if (order == nullptr || count == 0u) throw std::runtime_error("set_schedule: the work queue needs an order");
    if (mlp_instance()) throw std::runtime_error("set_schedule: the learned warm start runs one workgroup per environment");
    if (cfg_.ccd_mode != MR_CCD_SPECULATIVE)
        throw std::runtime_error("set_schedule: the hybrid CCD mode runs one workgroup per environment");
    if (geometry_)
        throw std::runtime_error("set_schedule: worlds with convex or mesh shapes run one workgroup per environment");
    std::vector<uint8_t> seen(env_, 0u);
    for (uint32_t i = 0; i < count; ++i) {
        if (order[i] >= env_) throw std::runtime_error("set_schedule: environment index out of range");
        if (seen[order[i]]) throw std::runtime_error("set_schedule: an environment appears twice in the order");
        seen[order[i]] = 1u;
    }
    d_queue_.reserve((size_t)count * sizeof(uint32_t));
    hip_check(hipMemcpy(d_queue_.ptr, order, (size_t)count * sizeof(uint32_t), hipMemcpyHostToDevice), "queue");
    d_queue_head_.reserve(sizeof(uint32_t));
    queue_grid_ = grid;
    queue_count_ = count;
    ++schedule_version_;
}

void World::set_schedule_reorder(bool enabled) {
    hip_check(hipStreamSynchronize(stream_), "sync");
    if (enabled && (queue_grid_ == 0u || queue_count_ != env_))
        throw std::runtime_error("set_schedule_reorder: needs a work-queue schedule over every environment");
    reorder_ = enabled;
    ++schedule_version_;
}

uint32_t World::resident_workgroups() const {
    int per_cu = 0;
    hip_check(control_step_occupancy(lds_.total, &per_cu), "occupancy");
    int device = 0, cus = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
    return (uint32_t)(per_cu * cus);
}

void World::drop_graph() {
    if (graph_exec_ != nullptr) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_ != nullptr) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_commands_ = nullptr;
    graph_steps_ = 0u;
}

void World::enqueue_graph(const float* device_commands, uint32_t steps) {
    if (steps == 0u || device_commands == nullptr) throw std::runtime_error("enqueue_graph: commands and steps required");
    if (stage_profiling_) throw std::runtime_error("enqueue_graph: stage profiling is not captured in graphs");
    begin_submission();
    // the buffers the launches point at exist before the comparison and the capture
    hip_check(hipStreamSynchronize(stream_), "sync");
    d_status_.reserve((size_t)steps * env_ * sizeof(MRWorldStatusC));
    d_cstatus_.reserve((size_t)steps * env_ * sizeof(MRContactStatusC));
    d_summary_.reserve(sizeof(AdvanceSummaryDev));
    const KernelArgs now = base_args(0u);
    const bool reuse = graph_exec_ != nullptr && graph_commands_ == device_commands && graph_steps_ == steps &&
        std::memcmp(&now, &graph_args_, sizeof(KernelArgs)) == 0 && graph_status_ == d_status_.ptr &&
        graph_cstatus_ == d_cstatus_.ptr && graph_summary_ == d_summary_.ptr &&
        graph_schedule_ == schedule_version_;  // the grid and the reordering are launch shape, not arguments
    if (!reuse) {
        drop_graph();
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            launch_steps(device_commands, nullptr, steps, false, own_buffers());
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g != nullptr) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream_, &graph_), "end capture");
        hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "graph instantiate");
        graph_commands_ = device_commands;
        graph_steps_ = steps;
        graph_args_ = now;
        graph_status_ = d_status_.ptr;
        graph_cstatus_ = d_cstatus_.ptr;
        graph_summary_ = d_summary_.ptr;
        graph_schedule_ = schedule_version_;
    }
    hip_check(hipEventRecord(ev_begin_, stream_), "event");
    hip_check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
    hip_check(hipEventRecord(ev_end_, stream_), "event");
    last_steps_ = steps;
    last_stepped_ = (!mlp_instance() && queue_grid_ > 0u && queue_count_ < env_) ? queue_count_ : env_;
    last_status_ = static_cast<const MRWorldStatusC*>(d_status_.ptr);
    last_cstatus_ = static_cast<const MRContactStatusC*>(d_cstatus_.ptr);
    last_summary_ = static_cast<const AdvanceSummaryDev*>(d_summary_.ptr);
    in_flight_ = true;
}

void World::begin_submission() const {
    if (in_flight_)
        throw std::runtime_error("context busy: the world's resident state has a submission in flight; "
                                 "call mr_world_wait before the next one");
}

void World::enqueue_chained(const float* device_commands, const uint32_t* device_masks, uint32_t steps) {
    if (in_flight_)
        throw std::runtime_error("context busy: the world's resident state has a submission in flight; "
                                 "call mr_world_wait before the next one");
    enqueue(device_commands, device_masks, steps, true);
    in_flight_ = false;  // the chain orders itself on the stream; the caller waits through its own readback
}

void World::enqueue(const float* commands, const uint32_t* masks, uint32_t steps, bool device_inputs) {
    if (steps == 0u) throw std::runtime_error("control_step_count must be positive");
    if (commands == nullptr) throw std::runtime_error("commands are required");
    begin_submission();
    const size_t E = env_;
    const size_t per_step = E * model_.max_nv;
    const float* cmd = commands;
    const uint32_t* msk = masks;
    if (!device_inputs) {
        d_commands_.reserve((size_t)steps * per_step * sizeof(float));
        hip_check(hipMemcpyAsync(d_commands_.ptr, commands, (size_t)steps * per_step * sizeof(float),
                                 hipMemcpyHostToDevice, stream_),
                  "commands");
        cmd = static_cast<const float*>(d_commands_.ptr);
        if (masks != nullptr) {
            d_masks_.reserve((size_t)steps * E * sizeof(uint32_t));
            hip_check(hipMemcpyAsync(d_masks_.ptr, masks, (size_t)steps * E * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, stream_),
                      "masks");
            msk = static_cast<const uint32_t*>(d_masks_.ptr);
        }
    }
    if (stage_profiling_)
        hip_check(hipMemsetAsync(d_stage_cycles_.ptr, 0, (size_t)E * kStageCount * sizeof(unsigned long long), stream_),
                  "clear stage cycles");
    d_status_.reserve((size_t)steps * E * sizeof(MRWorldStatusC));
    d_cstatus_.reserve((size_t)steps * E * sizeof(MRContactStatusC));
    if (step_timing_) {
        while (step_events_.size() < (size_t)steps + 1u) {
            hipEvent_t ev;
            hip_check(hipEventCreate(&ev), "hipEventCreate");
            step_events_.push_back(ev);
        }
        d_step_us_.reserve((size_t)steps * E * sizeof(float));
    }
    d_summary_.reserve(sizeof(AdvanceSummaryDev));
    launch_steps(cmd, msk, steps, true, own_buffers());
    last_steps_ = steps;
    in_flight_ = true;
}

void World::enqueue_horizon(const float* device_commands, const uint32_t* device_masks, uint32_t steps) {
    if (steps == 0u) throw std::runtime_error("control_step_count must be positive");
    if (device_commands == nullptr) throw std::runtime_error("commands are required");
    if (queue_grid_ == 0u)
        throw std::runtime_error("enqueue_horizon: needs a work-queue schedule (mr_world_set_schedule with grid > 0)");
    if (mlp_instance())
        throw std::runtime_error("enqueue_horizon: the learned warm start runs one workgroup per environment");
    begin_submission();
    const size_t E = env_;
    if (stage_profiling_)
        hip_check(hipMemsetAsync(d_stage_cycles_.ptr, 0, E * kStageCount * sizeof(unsigned long long), stream_),
                  "clear stage cycles");
    d_status_.reserve((size_t)steps * E * sizeof(MRWorldStatusC));
    d_cstatus_.reserve((size_t)steps * E * sizeof(MRContactStatusC));
    if (step_timing_) d_step_us_.reserve((size_t)steps * E * sizeof(float));
    d_summary_.reserve(sizeof(AdvanceSummaryDev));
    launch_steps(device_commands, device_masks, steps, true, own_buffers(), true);
    last_steps_ = steps;
    in_flight_ = true;
}

void World::read_world_step_microseconds(float* out, uint32_t steps) const {
    if (!step_timing_ || steps > last_steps_ || d_step_us_.bytes < (size_t)steps * env_ * sizeof(float))
        throw std::runtime_error("step timing was not enabled for the last advance");
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(out, d_step_us_.ptr, (size_t)steps * env_ * sizeof(float), hipMemcpyDeviceToHost),
              "read world step times");
}

namespace {
constexpr size_t kStaticAlign = 256u;
size_t align_up(size_t n) { return (n + kStaticAlign - 1u) / kStaticAlign * kStaticAlign; }
} // namespace

size_t World::static_scratch_bytes() const {
    const size_t E = env_, pcap = cfg_.pair_capacity, ccap = cfg_.constraint_capacity;
    const size_t row_stride = std::max(1u, model_.max_nv);
    return align_up(E * pcap * sizeof(MRManifoldHeaderC)) +
        align_up(E * pcap * MR_MANIFOLD_POINT_CAPACITY * sizeof(MRManifoldPointC)) +
        2u * align_up(E * ccap * 3u * row_stride * sizeof(float)) + align_up(E * ccap * sizeof(ColdConstraint)) +
        align_up(E * ccap * sizeof(HotConstraint)) + align_up(sizeof(AdvanceSummaryDev));
}

void World::enqueue_static(const float* device_commands, const uint32_t* device_masks, uint32_t steps,
                           MRWorldStatusC* device_status, MRContactStatusC* device_cstatus) {
    if (steps == 0u) throw std::runtime_error("control_step_count must be positive");
    if (device_commands == nullptr) throw std::runtime_error("commands are required");
    if (device_status == nullptr || device_cstatus == nullptr)
        throw std::runtime_error("enqueue_device_static: the status buffers are required");
    if (step_timing_) throw std::runtime_error("enqueue_device_static: per-launch step timing must be off");
    int device = -1;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    if (device != device_) throw std::runtime_error("enqueue_device_static: the world's device must be current");
    void* base = nullptr;
    size_t capacity = 0;
    hip_check(static_scratch(&base, &capacity), "static scratch address");
    const size_t need = static_scratch_bytes();
    if (need > capacity)
        throw std::runtime_error("enqueue_device_static: the world needs " + std::to_string(need) +
                                 " bytes of scratch; the static array holds " + std::to_string(capacity));
    begin_submission();
    const size_t E = env_, pcap = cfg_.pair_capacity, ccap = cfg_.constraint_capacity;
    const size_t row_stride = std::max(1u, model_.max_nv);
    unsigned char* at = static_cast<unsigned char*>(base);
    auto carve = [&](size_t bytes) {
        unsigned char* p = at;
        at += align_up(bytes);
        return p;
    };
    StepBuffers b;
    b.stage_h = reinterpret_cast<MRManifoldHeaderC*>(carve(E * pcap * sizeof(MRManifoldHeaderC)));
    b.stage_p = reinterpret_cast<MRManifoldPointC*>(
        carve(E * pcap * MR_MANIFOLD_POINT_CAPACITY * sizeof(MRManifoldPointC)));
    b.jrow = reinterpret_cast<float*>(carve(E * ccap * 3u * row_stride * sizeof(float)));
    b.resp = reinterpret_cast<float*>(carve(E * ccap * 3u * row_stride * sizeof(float)));
    b.cold = reinterpret_cast<ColdConstraint*>(carve(E * ccap * sizeof(ColdConstraint)));
    b.hot = reinterpret_cast<HotConstraint*>(carve(E * ccap * sizeof(HotConstraint)));
    b.summary = reinterpret_cast<AdvanceSummaryDev*>(carve(sizeof(AdvanceSummaryDev)));
    b.status = device_status;
    b.cstatus = device_cstatus;
    if (stage_profiling_)
        hip_check(hipMemsetAsync(d_stage_cycles_.ptr, 0, E * kStageCount * sizeof(unsigned long long), stream_),
                  "clear stage cycles");
    launch_steps(device_commands, device_masks, steps, true, b);
    last_steps_ = steps;
    in_flight_ = true;
}

AdvanceSummary World::wait() {
    hip_check(hipStreamSynchronize(stream_), "stream synchronize");
    in_flight_ = false;
    hip_check(hipEventElapsedTime(&last_kernel_ms_, ev_begin_, ev_end_), "event elapsed");
    AdvanceSummary s;
    s.control_step_count = last_steps_;
    s.gpu_milliseconds = last_kernel_ms_;
    AdvanceSummaryDev d;
    if (last_summary_ == nullptr) throw std::runtime_error("nothing was enqueued");
    hip_check(hipMemcpy(&d, last_summary_, sizeof(d), hipMemcpyDeviceToHost), "advance summary");
    const uint64_t n = (uint64_t)last_steps_ * last_stepped_;
    s.failed_environment_steps = d.failed;
    s.successful_environment_steps = (uint32_t)(n - d.failed);
    if (d.first_failure != ~0ull) {
        s.first_failing_control_step = (uint32_t)(d.first_failure >> 40);
        s.first_failing_environment = (uint32_t)((d.first_failure >> 8) & 0xffffffffull);
        s.first_gpu_status_code = (uint32_t)(d.first_failure & 0xffull);
    }
    s.maximum_active_pairs = d.max_active_pairs;
    s.maximum_constraints = d.max_constraints;
    return s;
}

void World::read_statuses(MRWorldStatusC* out, uint32_t steps) const {
    if (steps > last_steps_) throw std::runtime_error("fewer control steps were executed");
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(out, last_status_, (size_t)steps * env_ * sizeof(MRWorldStatusC), hipMemcpyDeviceToHost),
              "read statuses");
}

void World::read_contact_statuses(MRContactStatusC* out, uint32_t steps) const {
    if (steps > last_steps_) throw std::runtime_error("fewer control steps were executed");
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(out, last_cstatus_, (size_t)steps * env_ * sizeof(MRContactStatusC), hipMemcpyDeviceToHost),
              "read contact statuses");
}

void World::read_acceleration(float* out) const {
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(out, d_accel_.ptr, (size_t)env_ * model_.max_nv * sizeof(float), hipMemcpyDeviceToHost),
              "read acceleration");
}

void World::read_evidence(MRPairRecordC* pairs, uint32_t* counts) const {
    hip_check(hipStreamSynchronize(stream_), "sync");
    hip_check(hipMemcpy(pairs, d_ev_pairs_.ptr, (size_t)env_ * cfg_.pair_capacity * sizeof(MRPairRecordC),
                        hipMemcpyDeviceToHost),
              "read evidence");
    hip_check(hipMemcpy(counts, d_ev_counts_.ptr, (size_t)env_ * sizeof(uint32_t), hipMemcpyDeviceToHost),
              "read evidence counts");
}

} // namespace mr
