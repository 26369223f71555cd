// Host executor of the batched temporalCone world on one MI355X GPU.
#pragma once

#include "mr_builder.h"
#include "mr_kernel_args.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mr {

// Device arena that only grows.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n);
    void release();
    ~DeviceBuffer() { release(); }
};

struct AdvanceSummary {
    uint32_t control_step_count = 0;
    uint32_t successful_environment_steps = 0;
    uint32_t failed_environment_steps = 0;
    uint32_t first_failing_environment = MR_INVALID_INDEX;
    uint32_t first_failing_control_step = MR_INVALID_INDEX;
    uint32_t first_gpu_status_code = 0;
    uint32_t maximum_active_pairs = 0;
    uint32_t maximum_constraints = 0;
    double gpu_milliseconds = 0.0;
};

class World {
public:
    World(const ModelData& model, const MRWorldStepConfigC& config);
    ~World();
    World(const World&) = delete;
    World& operator=(const World&) = delete;

    uint32_t environment_count() const { return env_; }
    const ModelData& model() const { return model_; }
    const MRWorldStepConfigC& config() const { return cfg_; }
    const LdsLayout& lds() const { return lds_; }
    size_t retained_bytes() const;
    // retained_bytes split by lifetime: immutable model tables, persistent
    // state (q, v, scene bodies, manifolds, reset state), transient per-step
    // scratch and boundary buffers (commands, masks, statuses).
    size_t immutable_bytes() const;
    size_t persistent_bytes() const;
    size_t transient_bytes() const;
    hipStream_t stream() const { return stream_; }
    // Device arrays of the resident state and model, for kernels that run
    // beside the control step on the same stream (the task layer).
    struct DeviceView {
        float* q;              // [E][max_nq]
        float* v;              // [E][max_nv]
        const float* reset_q;  // [E][max_nq]
        const float* reset_v;  // [E][max_nv]
        const MRDofC* dofs;    // the model's DoF table
    };
    DeviceView device_view() const;
    // Statuses [steps][E] of the last advance, in device memory.
    const MRWorldStatusC* last_status_device() const { return last_status_; }
    float control_timestep() const { return cfg_.control_timestep; }

    // Host-array state boundary (all environment-major with model strides).
    void set_state(const float* q, const float* v, const MRBodyStateC* scene);
    void set_reset_state(const float* q, const float* v, const MRBodyStateC* scene);
    void set_manifolds(const MRManifoldHeaderC* h, const MRManifoldPointC* p, const uint32_t* count);
    void get_state(float* q, float* v, MRBodyStateC* scene) const;
    void get_manifolds(MRManifoldHeaderC* h, MRManifoldPointC* p, uint32_t* count) const;

    // Enqueues `steps` control steps on the world's stream. commands is
    // [steps][E][max_nv], masks is [steps][E] (or null). With device_inputs
    // the pointers are device memory already resident in HBM; otherwise they
    // are host arrays copied before the first launch.
    void enqueue(const float* commands, const uint32_t* masks, uint32_t steps, bool device_inputs);
    // enqueue() for work that orders itself on the world's stream (the task
    // layer's steps): no in-flight bookkeeping, but refused while a caller's
    // submission waits for mr_world_wait.
    void enqueue_chained(const float* device_commands, const uint32_t* device_masks, uint32_t steps);
    // enqueue() on device inputs that allocates nothing: the per-step scratch
    // and the advance summary come from the library's static device array
    // (static_scratch), the statuses go to the caller's device buffers
    // [steps][E]. Throws when the world's scratch exceeds the array.
    void enqueue_static(const float* device_commands, const uint32_t* device_masks, uint32_t steps,
                        MRWorldStatusC* device_status, MRContactStatusC* device_cstatus);
    // Persistent horizon: every world of the work-queue schedule runs all
    // `steps` control steps inside one launch, each world's next step starting
    // when its previous one ends (device inputs [steps][E][max_nv], [steps][E]).
    void enqueue_horizon(const float* device_commands, const uint32_t* device_masks, uint32_t steps);
    // Wall-clock microseconds of each control step of each world in the last
    // advance, [steps][E] (step timing on).
    void read_world_step_microseconds(float* out, uint32_t steps) const;
    // Bytes of static scratch enqueue_static needs for this world.
    size_t static_scratch_bytes() const;
    // Waits for the stream and summarises the statuses of the last enqueue;
    // the world then accepts the next submission.
    AdvanceSummary wait();
    bool in_flight() const { return in_flight_; }

    void read_statuses(MRWorldStatusC* out, uint32_t steps) const;  // [steps][E] of the last enqueue
    void read_contact_statuses(MRContactStatusC* out, uint32_t steps) const;
    void read_acceleration(float* out) const;                       // [E][max_nv]
    void read_evidence(MRPairRecordC* pairs, uint32_t* counts) const; // [E][pair_capacity], [E]
    uint32_t last_steps() const { return last_steps_; }
    float last_kernel_milliseconds() const { return last_kernel_ms_; }
    // Per-launch HIP-event timing of the control-step kernels (off by default).
    void set_step_timing(bool enabled);
    // Optional per-stage shader-clock counters, summed over an advance: [E][kStageCount].
    void set_stage_profiling(bool enabled);
    void read_stage_cycles(unsigned long long* out) const;
    void read_step_milliseconds(double* out, uint32_t steps) const;
    // Learned warm start: device weights (the caller keeps them alive while
    // set), or null for the plain warm start.
    void set_warm_start_mlp(const MRMlpWeightsC* weights);
    // Capture of (features, converged impulse) samples at microstep
    // capture_substep of every control step of later advances.
    void set_mlp_capture(bool enabled);
    void read_mlp_samples(float* out) const;  // [E][constraint_capacity][kMlpSampleFloats] of the last step
    const LdsLayout& mlp_lds() const { return lds_mlp_; }
    // Launch schedule. grid == 0: one workgroup per environment (the default).
    // grid > 0: the persistent work-queue instance with `grid` workgroups taking
    // the environments of `order` (distinct indices, a subset allowed) in that
    // order; environments not listed are not stepped and their status records
    // are left as they were.
    void set_schedule(const uint32_t* order, uint32_t count, uint32_t grid);
    // After every control step, re-sort the work queue by that step's cone sweeps, heaviest
    // first (needs a work-queue schedule over every environment).
    void set_schedule_reorder(bool enabled);
    uint32_t queue_grid() const { return queue_grid_; }
    // Workgroups of the control-step kernel the GPU holds at once for this world.
    uint32_t resident_workgroups() const;
    // Graph-captured horizon: `steps` control steps on device commands
    // [steps][E][max_nv] at a fixed address, captured once per (pointer, steps,
    // schedule) and replayed afterwards; wait() as after enqueue.
    void enqueue_graph(const float* device_commands, uint32_t steps);

private:
    // Device buffers one advance writes besides the persistent state.
    struct StepBuffers {
        MRManifoldHeaderC* stage_h;
        MRManifoldPointC* stage_p;
        float* jrow;
        float* resp;
        ColdConstraint* cold;
        HotConstraint* hot;
        MRWorldStatusC* status;      // [steps][E]
        MRContactStatusC* cstatus;   // [steps][E]
        AdvanceSummaryDev* summary;
    };
    StepBuffers own_buffers();
    KernelArgs base_args(uint32_t step) const;
    void response_scratch_bounds(uint32_t& slots, uint32_t& stack) const;
    void compute_lds_layout();

    ModelData model_;
    MRWorldStepConfigC cfg_;
    uint32_t env_ = 0;
    LdsLayout lds_{};
    LdsLayout lds_mlp_{};  // the MLP instance's layout (plain layout plus its regions); total 0 if it does not fit
    MRMlpWeightsC mlp_{};
    bool mlp_enabled_ = false, mlp_capture_ = false;
    bool mlp_instance() const { return mlp_enabled_ || mlp_capture_; }
    uint32_t queue_grid_ = 0, queue_count_ = 0, schedule_version_ = 0;
    bool reorder_ = false;  // re-sort the work queue by each step's cone sweeps, heaviest first
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    const float* graph_commands_ = nullptr;
    uint32_t graph_steps_ = 0;
    // what the captured launches were built from: replayed only while unchanged
    KernelArgs graph_args_{};
    const void* graph_status_ = nullptr;
    const void* graph_cstatus_ = nullptr;
    const void* graph_summary_ = nullptr;
    uint32_t graph_schedule_ = 0;  // schedule_version_ the graph was captured with
    void launch_steps(const float* cmd, const uint32_t* msk, uint32_t steps, bool events, const StepBuffers& buf,
                      bool horizon = false);
    bool last_horizon_ = false;  // the last advance ran as one persistent horizon launch
    // A resident world takes one submission at a time: set by every enqueue,
    // cleared by wait() (the reference's contextBusy rule for a resident
    // continuation, MetalWorld.mm:16590-16603).
    bool in_flight_ = false;
    void begin_submission() const;  // throws while a submission is in flight
    // where the last advance wrote its statuses and summary
    const MRWorldStatusC* last_status_ = nullptr;
    const MRContactStatusC* last_cstatus_ = nullptr;
    const AdvanceSummaryDev* last_summary_ = nullptr;
    int device_ = -1;  // the device current at creation
    float us_per_tick_ = 0.0f;  // of the device's constant-rate clock, measured when step timing is enabled
    void drop_graph();
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_begin_ = nullptr, ev_end_ = nullptr;
    // model
    DeviceBuffer d_models_, d_env_model_, d_slot_plan_, d_bodies_, d_dofs_, d_actuators_, d_shapes_, d_materials_, d_pairs_, d_heights_;
    DeviceBuffer d_geometries_, d_geometry_vertices_, d_triangles_, d_bvh_nodes_;
    bool geometry_ = false;  // some shape is a convex hull or a triangle mesh
    // state
    DeviceBuffer d_q_, d_v_, d_scene_, d_mh_, d_mp_, d_mcount_, d_mcurrent_;
    DeviceBuffer d_reset_q_, d_reset_v_, d_reset_scene_;
    // per-advance inputs / outputs
    DeviceBuffer d_commands_, d_masks_, d_status_, d_cstatus_, d_accel_, d_ev_pairs_, d_ev_counts_;
    // scratch
    DeviceBuffer d_step_us_;
    DeviceBuffer d_stage_h_, d_stage_p_, d_jrow_, d_resp_, d_cold_, d_hot_, d_stage_cycles_, d_summary_, d_samples_, d_queue_, d_queue_head_;
    bool stage_profiling_ = false;
    uint32_t last_steps_ = 0;
    uint32_t last_stepped_ = 0;  // environments each control step of the last advance stepped
    float last_kernel_ms_ = 0.0f;
    bool step_timing_ = false;
    std::vector<hipEvent_t> step_events_;
};

void hip_check(hipError_t e, const char* what);

} // namespace mr
