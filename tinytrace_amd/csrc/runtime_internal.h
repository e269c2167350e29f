// Internal state behind the C ABI, shared by the render path (runtime.hip)
// and the serving loop (server.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tinytrace.h"
#include "kernels.h"

namespace tt {

struct EventPair {
    hipEvent_t begin = nullptr, end = nullptr;
};

}  // namespace tt

namespace tt {

// Where a render runs.  Consecutive renders take the lanes in turn (two or
// three), so that one frame's last, latency-bound bounces overlap the next
// frames' first ones; within a lane, when there are hardware queues for it, bounce
// b's shadow rays run on `side` beside bounce b + 1's extension on `stream`
// (fork after shade b, join before shade b + 1).  Each lane has its own path
// buffers, counters and traversal stack spill areas.
struct Lane {
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // Cross-lane order: the next render (on the next lane) accumulates after
    // `accumulated`.  (ReSTIR buffers are ordered per target instead:
    // tt_target::restir_done.)
    hipEvent_t accumulated = nullptr;
    hipEvent_t done = nullptr;  // end of the lane's last render
    EventPair integrate;
    bool pending = false;       // a render the API stream has not waited for
    bool used = false;          // accumulated has been recorded
    uint32_t* counters = nullptr;
    uint32_t* work = nullptr;  // W_SETS cursor sets (tt::WORK_SET words each)
    uint32_t* spill = nullptr;
    uint32_t* side_spill = nullptr;
    size_t path_capacity = 0;
    void* path_memory = nullptr;
    // Serving batches with ReSTIR requests: per path slot the final and
    // initial reservoirs and the (pixel, frame) key shade 0 reads (grown on
    // demand).
    size_t restir_slot_capacity = 0;
    void* restir_slot_memory = nullptr;
    // Radiance-cache queues (tt_render_cached), grown on demand: sizes the
    // memory holds, and after the queues the judge predictions and the judge
    // statistics (6 doubles).
    size_t cache_paths = 0;
    uint32_t cache_train = 0, cache_judge = 0;
    void* cache_memory = nullptr;
    float* judge_pred = nullptr;
    double* judge_stats = nullptr;
    // Timing of a cached render's query inference (created on first use).
    EventPair cache_infer;
};

// A timed launch of a render that collects statistics.
enum StatsKind : uint8_t { STATS_EXTEND = 0, STATS_SHADOW = 1, STATS_RESTIR = 2, STATS_RENDER = 3 };
struct StatsLaunch {
    EventPair events;
    StatsKind kind;
};

}  // namespace tt

struct tt_context {
    int device = 0;
    // The API stream: uploads, downloads, target clears, display, serving
    // and the cache's host operations.  Renders run on the lanes (cached
    // ones train on the cache's own stream); code outside them takes the
    // stream through tt::api_stream, which first orders it after every
    // render enqueued so far.
    hipStream_t stream = nullptr;
    std::string error;
    int cu_count = 0;
    int grid = 0;
    static constexpr uint32_t MAX_LANES = 3;
    tt::Lane lanes[MAX_LANES];
    uint32_t lane_count = 2;
    // Lanes in use and whether they have side streams, from the hardware
    // queues the HIP runtime gives the process (GPU_MAX_HW_QUEUES): every
    // stream gets a queue of its own, since streams that share a queue run
    // one after the other, which would undo the overlaps.
    bool side_streams = false;
    uint32_t next_lane = 0;
    int prev_lane = -1;      // lane of the last render that ran on a lane
    uint32_t last_lane = 0;  // lane of the last render (its counters and timing)
    // Recorded on the API stream when a render starts after host calls put
    // work there (api_dirty, set by tt::api_stream); renders wait for it.
    // The cache's training, which renders enqueue on the API stream, does not
    // set api_dirty: the next render's tracing need not wait for it.
    hipEvent_t api_mark = nullptr;
    bool api_dirty = true;
    uint32_t* work = nullptr;
    // Timing.
    // Statistics (TT_RENDER_COLLECT_STATS), summed over the renders since the
    // last tt_render_stats_get: their timed launches (event pairs drawn from
    // a pool), path counts, and ray counters added up on the device.
    std::vector<tt::EventPair> stats_pool;
    std::vector<tt::StatsLaunch> stats_launches;
    size_t last_stats_first = 0;  // the last stats render's first launch record
    uint32_t stats_renders = 0;
    uint64_t stats_paths = 0;
    unsigned long long* stats_sum = nullptr;  // C_WORDS sums, device
    double last_integrate_ms = -1.0;
    // Scenes, targets and servers alive on this context.  Destroying the
    // context while any remain only marks it; the last of them frees it, so
    // handles may be destroyed in any order.
    uint32_t children = 0;
    bool closing = false;
    // Display pass (tt_target_display): RGBA8 staging, grown on demand, and
    // its timing events (created on first use).
    size_t display_capacity = 0;
    void* display_memory = nullptr;
    tt::EventPair post_process;
    bool post_process_recorded = false;
    // Device address of this device's copy of the library's static trace
    // scratch (tt_trace_closest_static), resolved at context creation; null
    // with the reason in static_scratch_error if that failed.
    void* static_scratch = nullptr;
    std::string static_scratch_error;
};

struct tt_scene {
    tt_context* context = nullptr;
    tt::DeviceScene view{};
    std::vector<void*> allocations;
    tt_scene_info info{};
    uint64_t serial = 0;  // unique per scene created in this process
    float bounds_lo[3] = {0.0f, 0.0f, 0.0f}, bounds_hi[3] = {0.0f, 0.0f, 0.0f};  // world box
};

struct tt_target {
    tt_context* context = nullptr;
    uint32_t width = 0, height = 0;
    float4* pixels = nullptr;
    // ReSTIR DI history (allocated on first use): two G-buffers and the
    // reservoir arrays of tt::RestirBuffers in one allocation.
    void* restir_memory = nullptr;
    uint32_t restir_cur = 0;
    bool restir_history = false;
    uint64_t restir_scene = 0;  // serial of the scene the history belongs to
    // Recorded after the last ReSTIR render's shade 0, the last reader of
    // these buffers; the next ReSTIR render on this target, on whatever lane,
    // waits for it before its ReSTIR passes rewrite them.
    hipEvent_t restir_done = nullptr;
    bool restir_done_recorded = false;
};

namespace tt {

inline tt_status fail(tt_context* c, tt_status s, const std::string& msg) {
    if (c) c->error = msg;
    return s;
}

inline tt_status hip_fail(tt_context* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? TT_OUT_OF_MEMORY : TT_HIP_ERROR,
                std::string(what) + ": " + hipGetErrorString(e));
}

void context_retain(tt_context* c);
void context_release(tt_context* c);

// A cached frame's shares over `paths` paths from tt_cache_settings (null
// or zero fields: the defaults): training and judge thresholds on a path's
// hash, record capacities with a margin, Adam steps.
struct CachePlan {
    uint32_t records;  // training records a training path takes
    bool ordered;      // training takes the records in path order
    uint32_t train_steps;
    uint32_t train_capacity, judge_capacity;
    uint32_t train_threshold, judge_threshold;
};
// TT_INVALID_ARGUMENT with *why set for settings out of range.
tt_status cache_plan(const tt_cache_settings* settings, uint32_t paths, CachePlan& out, const char** why);
// Turns on the cache in f with the plan's shares.
void apply_cache_plan(const CachePlan& plan, FrameArgs& f);
// Radiance-cache queues of a lane for `paths` paths, `train` training and
// `judge` judge records (grown on demand, kept in the lane).
tt_status ensure_cache(tt_context* c, uint32_t lane, size_t paths, uint32_t train, uint32_t judge, CacheBuffers& cb);

// Device LBVH build of every BLAS (bvh_gpu.hip).
struct BvhNode;
struct BvhTri;
struct BlasSummary;
tt_status build_blas_device(tt_context* c, const tt_scene_desc& d, const int16_t* positions, const uint32_t* indices,
                            BvhNode* nodes, BvhTri* tris, std::vector<BlasSummary>& summary, uint32_t& node_count);

// 4-wide nodes from the binary ones (bvh_wide.hip): renumbers the BLAS
// roots of the device instance records and tlas_root; *out is a new device
// allocation of wide_count nodes (nullptr when count == 0).
struct BvhWide;
tt_status collapse_wide(tt_context* c, const BvhNode* nodes, uint32_t count, void* instances,
                        uint32_t instance_count, uint32_t& tlas_root, BvhWide** out, uint32_t& wide_count);

// Radiance-cache MLP (mlp.hip).  Training runs on the cache's own stream;
// its `trained` event orders later inference (and host access) after it.
tt_context* mlp_context(const tt_mlp* m);
tt_status mlp_set_scene_bounds(tt_mlp* m, const float lo[3], const float hi[3]);
// Makes `stream` wait for the last training enqueued.
tt_status mlp_wait_trained(tt_mlp* m, hipStream_t stream);
// Inference on `stream` of min(*count_dev, count) records (count records
// when count_dev is null): out_rgb[3 q ..] = L and/or radiance[path] +=
// throughput * L with (throughput, path) from scale_path.
tt_status mlp_infer_device(tt_mlp* m, hipStream_t stream, const float4* records, const float4* scale_path,
                           const uint32_t* count_dev, uint32_t count, float* out_rgb, float4* radiance);
// The cache against the plainly traced judge records on `stream`:
// predictions into pred (3 floats a record), then stats[0..5] = record count,
// sum of predictions, sum of targets, sum of squared differences, sum of
// squared targets, sum of the relative L2 loss terms (doubles).
tt_status mlp_judge_device(tt_mlp* m, hipStream_t stream, const float4* records, const float4* targets,
                           const uint32_t* count_dev, uint32_t capacity, float* pred, double* stats);
// `steps` Adam steps over consecutive slices of min(*count_dev, capacity)
// records (count_dev null: capacity records), on the cache's stream after the
// work enqueued on `after` so far.  apply false: one gradient, no update,
// into grad_out (device, MLP_PARAMS floats).  timed: HIP events around the
// steps (mlp_train_ms).  order_span > 0: the records are first put in the
// order of the key in each target's w (the shade kernel's path x records a
// path takes + record index, below order_span), so the result does not
// depend on the order their slots were taken in.
tt_status mlp_train_device(tt_mlp* m, hipStream_t after, const float4* records, const float4* targets,
                           const uint32_t* count_dev, uint32_t capacity, uint32_t steps, bool apply, float* grad_out,
                           bool timed = false, uint32_t order_span = 0);
// Milliseconds of the last timed training (synchronises).
tt_status mlp_train_ms(tt_mlp* m, double* out);
// Mean loss of the first step of the last training, whose record count was
// `count` over `steps` steps (synchronises).
tt_status mlp_first_step_loss(tt_mlp* m, uint32_t count, uint32_t steps, double* out);

// Path buffers of at least `paths` paths (grown on demand, kept in the
// context); fills the pointers of `b`.
tt_status ensure_paths(tt_context* c, uint32_t lane, size_t paths, PathBuffers& b);
// The API stream, after enqueueing waits for every render not yet waited for.
hipStream_t api_stream(tt_context* c);
// Bytes of path state per path slot.
constexpr size_t PATH_BYTES = 11 * sizeof(float4) + sizeof(uint64_t);
// Path-buffer bytes all lanes together may hold before renders stop
// overlapping (they then run on lane 0 alone).
constexpr uint64_t LANE_PATH_BUDGET = 24ull << 30;
// Per-render arguments of the kernels from the reference Constants.
FrameArgs frame_args(const tt_constants& k, uint32_t width, uint32_t height, uint32_t rng, uint32_t seed);

}  // namespace tt

#define TT_HIP(ctx, expr)                                          \
    do {                                                           \
        hipError_t e_ = (expr);                                    \
        if (e_ != hipSuccess) return tt::hip_fail(ctx, e_, #expr); \
    } while (0)
