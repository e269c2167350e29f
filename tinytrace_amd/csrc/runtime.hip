// C ABI of tinytrace_amd: context, resident scenes, render targets, the
// render call and the test entry points.  See include/tinytrace.h for the
// reference interfaces each function replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "abi_guard.h"
#include "../../include/tinytrace.h"
#include "bvh.h"
#include "kernels.h"
#include "runtime_internal.h"

using tt::fail;
using tt::hip_fail;

namespace {

tt_status upload(tt_context* c, tt_scene* s, const void* src, size_t bytes, const void** dst) {
    if (bytes == 0) {
        *dst = nullptr;
        return TT_OK;
    }
    void* p = nullptr;
    TT_HIP(c, hipMalloc(&p, bytes));
    s->allocations.push_back(p);
    s->info.device_bytes += bytes;
    TT_HIP(c, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = p;
    return TT_OK;
}

// Records the start of a timed launch on `stream`; *index names the record
// for stats_end.
tt_status stats_begin(tt_context* c, tt::StatsKind kind, hipStream_t stream, size_t* index) {
    tt::EventPair e;
    if (c->stats_pool.empty()) {
        TT_HIP(c, hipEventCreate(&e.begin));
        if (hipEventCreate(&e.end) != hipSuccess) {
            (void)hipEventDestroy(e.begin);
            return fail(c, TT_HIP_ERROR, "hipEventCreate failed");
        }
    } else {
        e = c->stats_pool.back();
        c->stats_pool.pop_back();
    }
    c->stats_launches.push_back(tt::StatsLaunch{e, kind});
    *index = c->stats_launches.size() - 1;
    TT_HIP(c, hipEventRecord(e.begin, stream));
    return TT_OK;
}

tt_status stats_end(tt_context* c, size_t index, hipStream_t stream) {
    TT_HIP(c, hipEventRecord(c->stats_launches[index].events.end, stream));
    return TT_OK;
}

// Scratch of tt_trace_closest_static, for callers that cannot allocate: the
// ray cursor (one word, 256 B reserved), staging for up to
// STATIC_TRACE_MAX_RAYS rays and hits, and the traversal stack spill area of
// a grid of at most STATIC_TRACE_GRID workgroups.  Each device holds its own
// copy for the life of the process.  Host code never passes `scratch`
// itself to HIP (on the host it names a stand-in, not device memory): it
// uses the per-device address that hipGetSymbolAddress returns.
constexpr uint32_t STATIC_TRACE_MAX_RAYS = TT_TRACE_STATIC_MAX_RAYS;
constexpr int STATIC_TRACE_GRID = 128;
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t ST_CURSOR = 0;  // one cursor set
constexpr size_t ST_RAYS = align256(size_t(tt::WORK_SET) * sizeof(uint32_t));
constexpr size_t ST_HITS = align256(ST_RAYS + size_t(STATIC_TRACE_MAX_RAYS) * sizeof(tt_ray));
constexpr size_t ST_SPILL = align256(ST_HITS + size_t(STATIC_TRACE_MAX_RAYS) * sizeof(tt_hit));
constexpr size_t ST_BYTES =
    ST_SPILL + size_t(STATIC_TRACE_GRID) * tt::TRACE_THREADS * tt::TRACE_SPILL_ENTRIES * sizeof(uint32_t);
static_assert(ST_BYTES < (12u << 20), "static trace scratch stays below 12 MiB");

static __device__ __align__(256) uint8_t scratch[ST_BYTES];

// Device address of each device's copy of `scratch`, resolved once per device
// (with that device current) by the first context created on it.
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_address[MAX_DEVICES] = {};

// Called by tt_context_create with c->device current.  A failure does not
// fail the context: only tt_trace_closest_static needs the address, and it
// reports the reason.
void resolve_static_scratch(tt_context* c) noexcept {
    if (c->device >= MAX_DEVICES) {
        tt::set_error_nothrow(&c->static_scratch_error, "device ordinal above the static scratch table");
        return;
    }
    try {
        std::lock_guard<std::mutex> lock(scratch_mutex);
        if (!scratch_address[c->device]) {
            void* p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                c->static_scratch_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
                return;
            }
            scratch_address[c->device] = p;
        }
        c->static_scratch = scratch_address[c->device];
    } catch (...) {
        tt::set_error_nothrow(&c->static_scratch_error, "could not resolve the static scratch address");
    }
}

}  // namespace

namespace tt {

// Path buffers: throughput, radiance, 2x2 queue arrays, hits, 4 shadow
// arrays (float4 each) and the generator words.
tt_status ensure_paths(tt_context* c, uint32_t lane, size_t paths, PathBuffers& b) {
    Lane& L = c->lanes[lane];
    if (paths > L.path_capacity) {
        if (L.path_memory) {
            TT_HIP(c, hipStreamSynchronize(api_stream(c)));
            TT_HIP(c, hipFree(L.path_memory));
            L.path_memory = nullptr;
            L.path_capacity = 0;
        }
        size_t bytes = paths * PATH_BYTES;
        TT_HIP(c, hipMalloc(&L.path_memory, bytes));
        L.path_capacity = paths;
    }
    size_t n = L.path_capacity;
    float4* f = static_cast<float4*>(L.path_memory);
    b.throughput = f + 0 * n;
    b.radiance = f + 1 * n;
    b.queue[0][0] = f + 2 * n;
    b.queue[0][1] = f + 3 * n;
    b.queue[1][0] = f + 4 * n;
    b.queue[1][1] = f + 5 * n;
    b.hits = f + 6 * n;
    b.shadow[0] = f + 7 * n;
    b.shadow[1] = f + 8 * n;
    b.shadow[2] = f + 9 * n;
    b.shadow[3] = f + 10 * n;
    b.rng = reinterpret_cast<uint64_t*>(f + 11 * n);
    b.counters = L.counters;
    b.work = L.work;
    b.spill = L.spill;
    b.cache = CacheBuffers{};
    b.restir_final = nullptr;
    b.restir_initial = nullptr;
    b.restir_key = nullptr;
    return TT_OK;
}

hipStream_t api_stream(tt_context* c) {
    for (Lane& L : c->lanes)
        if (L.pending) {
            // A failed wait leaves the order to the synchronisations that
            // every reader of render results performs anyway.
            (void)hipStreamWaitEvent(c->stream, L.done, 0);
            L.pending = false;
        }
    c->api_dirty = true;
    return c->stream;
}

// Radiance-cache queues of a lane (grown on demand; a lane's previous
// render and training finish before its memory is replaced).
tt_status ensure_cache(tt_context* c, uint32_t lane, size_t paths, uint32_t train, uint32_t judge, CacheBuffers& cb) {
    Lane& L = c->lanes[lane];
    if (paths > L.cache_paths || train > L.cache_train || judge > L.cache_judge) {
        paths = std::max(paths, L.cache_paths);
        train = std::max(train, L.cache_train);
        judge = std::max(judge, L.cache_judge);
        if (L.cache_memory) {
            TT_HIP(c, hipStreamSynchronize(L.stream));
            TT_HIP(c, hipStreamSynchronize(c->stream));
            TT_HIP(c, hipFree(L.cache_memory));
            L.cache_memory = nullptr;
            L.cache_paths = 0;
            L.cache_train = L.cache_judge = 0;
        }
        const size_t records = size_t(train) + judge;
        const size_t bytes = paths * (sizeof(uint4) + 5 * sizeof(float4)) + records * 6 * sizeof(float4) +
            size_t(judge) * 3 * sizeof(float) + 8 * sizeof(double);
        TT_HIP(c, hipMalloc(&L.cache_memory, bytes));
        L.cache_paths = paths;
        L.cache_train = train;
        L.cache_judge = judge;
    }
    const size_t n = L.cache_paths, records = size_t(L.cache_train) + L.cache_judge;
    float4* f = static_cast<float4*>(L.cache_memory);
    cb.query_rec = f;
    cb.query_scale = f + 4 * n;
    cb.open = reinterpret_cast<uint4*>(f + 5 * n);
    float4* r = f + 6 * n;
    cb.train_rec = r;
    cb.train_thr = r + 4 * records;
    cb.train_target = r + 5 * records;
    L.judge_pred = reinterpret_cast<float*>(r + 6 * records);
    L.judge_stats = reinterpret_cast<double*>(L.judge_pred + ((3 * size_t(L.cache_judge) + 1) & ~size_t(1)));
    cb.train_capacity = train;
    cb.judge_base = L.cache_train;
    cb.judge_capacity = judge;
    return TT_OK;
}

tt_status cache_plan(const tt_cache_settings* cs, uint32_t paths, CachePlan& out, const char** why) {
    const uint32_t train_samples = cs && cs->train_samples ? cs->train_samples : 65536u;
    out.train_steps = cs && cs->train_steps ? cs->train_steps : 4u;
    if (out.train_steps > 64) {
        *why = "train_steps > 64";
        return TT_INVALID_ARGUMENT;
    }
    out.records = cs && cs->records_per_path ? cs->records_per_path : 1u;
    out.ordered = cs && cs->ordered_training != 0u;
    if (out.records > CACHE_OPEN_RECORDS) {
        *why = "records_per_path > 3";
        return TT_INVALID_ARGUMENT;
    }
    float judge = cs && cs->judge_fraction != 0.0f ? cs->judge_fraction : 0.01f;
    if (!(judge > 0.0f)) judge = 0.0f;  // negative (or NaN): no judge paths
    judge = std::min(judge, 1.0f);
    // A training path records up to `records` samples.
    const double train_frac = std::min(1.0, double(train_samples) / (double(out.records) * double(paths)));
    // Record capacities: the expected counts with a margin (records past
    // them are dropped).
    out.train_capacity = uint32_t(std::min<uint64_t>(uint64_t(train_samples) + train_samples / 4 + 1024,
                                                      uint64_t(out.records) * paths));
    out.judge_capacity = judge > 0.0f ? uint32_t(std::min<double>(double(paths), 1.25 * judge * paths + 1024.0)) : 0u;
    auto threshold = [](double x) { return x >= 1.0 ? 0xffffffffu : uint32_t(x * 4294967296.0); };
    out.train_threshold = threshold(train_frac);
    out.judge_threshold = threshold(train_frac + judge);
    return TT_OK;
}

void apply_cache_plan(const CachePlan& plan, FrameArgs& f) {
    f.cache_mode = 1;
    f.cache_records = plan.records;
    f.train_threshold = plan.train_threshold;
    f.judge_threshold = plan.judge_threshold;
}

FrameArgs frame_args(const tt_constants& k, uint32_t width, uint32_t height, uint32_t rng, uint32_t seed) {
    FrameArgs f;
    std::memcpy(f.inverse_view, k.inverse_view, sizeof(f.inverse_view));
    std::memcpy(f.inverse_proj, k.inverse_proj, sizeof(f.inverse_proj));
    f.camera_position[0] = k.camera_position[0];
    f.camera_position[1] = k.camera_position[1];
    f.camera_position[2] = k.camera_position[2];
    f.width = width;
    f.height = height;
    f.sample_count = k.sample_count;
    f.bounce_count = k.bounce_count;
    f.frame_index = k.frame_index;
    f.accumulated_frame_count = k.accumulated_frame_count;
    f.light_sampling = k.light_sampling;
    f.rng = rng;
    f.seed = seed;
    f.cache_mode = 0;
    f.train_threshold = 0;
    f.judge_threshold = 0;
    f.cache_records = 0;
    f.primary_tiles = 0;
    return f;
}

}  // namespace tt

extern "C" {

tt_status tt_context_create(const tt_context_desc* desc, tt_context** out) try {
    if (!out) return TT_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return TT_NO_DEVICE;
    int device = desc ? desc->device : 0;
    if (device < 0 || device >= count) return TT_NO_DEVICE;
    tt_context* c = new (std::nothrow) tt_context();
    if (!c) return TT_OUT_OF_MEMORY;
    c->device = device;
    auto cleanup = [&](tt_status s) {
        tt_context_destroy(c);
        return s;
    };
    if (hipSetDevice(device) != hipSuccess) return cleanup(TT_NO_DEVICE);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return cleanup(TT_HIP_ERROR);
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return cleanup(TT_KERNEL_LOAD);
    c->cu_count = prop.multiProcessorCount;
    c->grid = tt::trace_grid(c->cu_count);
    resolve_static_scratch(c);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(TT_HIP_ERROR);
    if (hipEventCreateWithFlags(&c->api_mark, hipEventDisableTiming) != hipSuccess) return cleanup(TT_HIP_ERROR);
    if (hipMalloc(&c->work, tt::WORK_SET * sizeof(uint32_t)) != hipSuccess) return cleanup(TT_OUT_OF_MEMORY);
    if (hipMalloc(&c->stats_sum, tt::C_WORDS * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(c->stats_sum, 0, tt::C_WORDS * sizeof(unsigned long long)) != hipSuccess)
        return cleanup(TT_OUT_OF_MEMORY);
    const size_t spill_bytes = tt::spill_words(c->grid) * sizeof(uint32_t);
    // Hardware queues: the null stream and the API stream take two; each
    // lane takes one, or two with its side stream (HIP's default is four).
    // Measured on C2: three lanes with side streams (8 queues) 8.3 ms a
    // frame, three without (5) 8.5, two with (6) 8.7, two without (4) 9.1;
    // a fourth lane gains nothing.
    const char* env = std::getenv("GPU_MAX_HW_QUEUES");
    const int queues = env && std::atoi(env) > 0 ? std::atoi(env) : 4;
    c->lane_count = queues >= 5 ? 3u : 2u;
    c->side_streams = queues >= int(2 + 2 * c->lane_count);
    for (uint32_t i = 0; i < c->lane_count; i++) {
        tt::Lane& L = c->lanes[i];
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(TT_HIP_ERROR);
        if (c->side_streams && hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking) != hipSuccess)
            return cleanup(TT_HIP_ERROR);
        for (hipEvent_t* e : {&L.fork, &L.join, &L.accumulated, &L.done})
            if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return cleanup(TT_HIP_ERROR);
        if (hipEventCreate(&L.integrate.begin) != hipSuccess || hipEventCreate(&L.integrate.end) != hipSuccess)
            return cleanup(TT_HIP_ERROR);
        if (hipMalloc(&L.counters, tt::C_WORDS * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&L.work, size_t(tt::W_SETS) * tt::WORK_SET * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&L.spill, spill_bytes) != hipSuccess || hipMalloc(&L.side_spill, spill_bytes) != hipSuccess)
            return cleanup(TT_OUT_OF_MEMORY);
    }
    *out = c;
    return TT_OK;
} catch (...) {
    return tt::exception_status(nullptr);
}

}  // extern "C"

namespace {

void context_free(tt_context* c) {
    // Teardown is best effort: errors here have no caller to report to.
    (void)hipSetDevice(c->device);
    for (tt::Lane& L : c->lanes) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.side) (void)hipStreamSynchronize(L.side);
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (tt::Lane& L : c->lanes) {
        if (L.path_memory) (void)hipFree(L.path_memory);
        if (L.restir_slot_memory) (void)hipFree(L.restir_slot_memory);
        if (L.cache_memory) (void)hipFree(L.cache_memory);
        if (L.counters) (void)hipFree(L.counters);
        if (L.work) (void)hipFree(L.work);
        if (L.spill) (void)hipFree(L.spill);
        if (L.side_spill) (void)hipFree(L.side_spill);
        for (hipEvent_t e : {L.fork, L.join, L.accumulated, L.done, L.integrate.begin, L.integrate.end,
                             L.cache_infer.begin, L.cache_infer.end})
            if (e) (void)hipEventDestroy(e);
        if (L.stream) (void)hipStreamDestroy(L.stream);
        if (L.side) (void)hipStreamDestroy(L.side);
    }
    if (c->api_mark) (void)hipEventDestroy(c->api_mark);
    if (c->display_memory) (void)hipFree(c->display_memory);
    if (c->post_process.begin) (void)hipEventDestroy(c->post_process.begin);
    if (c->post_process.end) (void)hipEventDestroy(c->post_process.end);
    if (c->work) (void)hipFree(c->work);
    for (auto& l : c->stats_launches) c->stats_pool.push_back(l.events);
    for (auto& e : c->stats_pool) {
        (void)hipEventDestroy(e.begin);
        (void)hipEventDestroy(e.end);
    }
    if (c->stats_sum) (void)hipFree(c->stats_sum);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace

void tt::context_retain(tt_context* c) { c->children++; }

void tt::context_release(tt_context* c) {
    if (--c->children == 0 && c->closing) context_free(c);
}

extern "C" {

void tt_context_destroy(tt_context* c) try {
    if (!c) return;
    if (c->children > 0) {
        c->closing = true;
        return;
    }
    context_free(c);
} catch (...) {
}

const char* tt_last_error(const tt_context* c) { return c ? c->error.c_str() : "no context"; }

tt_status tt_scene_create(tt_context* c, const tt_scene_desc* d, tt_scene** out) {
    return tt_scene_create_ex(c, d, nullptr, out);
}

tt_status tt_scene_create_ex(tt_context* c, const tt_scene_desc* d, const tt_scene_options* options, tt_scene** out) try {
    if (!c || !d || !out) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    // Validate the reference arrays before building.
    for (uint32_t m = 0; m < d->mesh_count; m++) {
        const tt_mesh_record& r = d->meshes[m];
        if (r.index_count % 3 != 0 || uint64_t(r.index_offset) + r.index_count > d->index_count ||
            uint64_t(r.vertex_offset) + r.vertex_count > d->position_count || r.vertex_count > d->vertex_count)
            return fail(c, TT_INVALID_ARGUMENT, "mesh " + std::to_string(m) + " is out of range");
        for (uint32_t k = 0; k < r.index_count; k++)
            if (d->indices[r.index_offset + k] >= r.vertex_count)
                return fail(c, TT_INVALID_ARGUMENT, "mesh " + std::to_string(m) + " has an index out of range");
    }
    for (uint32_t i = 0; i < d->instance_count; i++) {
        const tt_instance& in = d->instances[i];
        if (in.mesh_index >= d->mesh_count || in.material_index >= d->material_count)
            return fail(c, TT_INVALID_ARGUMENT, "instance " + std::to_string(i) + " refers to a missing mesh/material");
    }
    if (d->instance_count >= (1u << 24)) return fail(c, TT_INVALID_ARGUMENT, "more than 2^24 instances");
    if (d->texture_count && (!d->textures || !d->texture_data))
        return fail(c, TT_INVALID_ARGUMENT, "texture arrays missing");
    for (uint32_t i = 0; i < d->texture_count; i++) {
        const tt_texture_record& r = d->textures[i];
        if (r.format > TT_TEXTURE_FORMAT_BC5_UNORM || r.width == 0 || r.height == 0 || r.levels == 0 || r.levels > 16)
            return fail(c, TT_INVALID_ARGUMENT, "texture " + std::to_string(i) + " has a bad record");
        // Bytes of a level and the alignment its texel or block loads need.
        const bool blocks = r.format == TT_TEXTURE_FORMAT_BC1_SRGB || r.format == TT_TEXTURE_FORMAT_BC5_UNORM;
        const uint64_t unit = r.format == TT_TEXTURE_FORMAT_BC1_SRGB ? 8 : r.format == TT_TEXTURE_FORMAT_BC5_UNORM ? 16
            : r.format == TT_TEXTURE_FORMAT_RGBA8_SRGB ? 4 : (r.format == TT_TEXTURE_FORMAT_R8_UNORM ? 1 : 2);
        const uint64_t align = blocks ? unit : 4;
        for (uint32_t k = 0; k < r.levels; k++) {
            uint64_t w = r.width >> k, h = r.height >> k;
            uint64_t bytes = blocks ? ((w + 3) / 4) * ((h + 3) / 4) * unit : w * h * unit;
            if (w == 0 || h == 0 || r.offsets[k] % align != 0 || r.offsets[k] + bytes > d->texture_data_bytes)
                return fail(c, TT_INVALID_ARGUMENT, "texture " + std::to_string(i) + " level out of range");
        }
    }
    for (uint32_t i = 0; i < d->material_count; i++)
        for (int t = 0; t < 27; t++) {
            uint32_t ti = d->materials[i].textures[t];
            if (ti == TT_INVALID_INDEX) continue;
            if (ti >= d->texture_count)
                return fail(c, TT_INVALID_ARGUMENT, "material " + std::to_string(i) + " refers to a missing texture");
        }
    for (uint32_t e = 0; e < d->emissive_triangle_count; e++)
        if (d->emissive_triangles[e].instance_index >= d->instance_count)
            return fail(c, TT_INVALID_ARGUMENT, "emissive triangle refers to a missing instance");

    const uint32_t builder = options ? options->bvh_builder : TT_BVH_BUILDER_HOST_SAH;
    if (builder != TT_BVH_BUILDER_HOST_SAH && builder != TT_BVH_BUILDER_DEVICE_LBVH)
        return fail(c, TT_INVALID_ARGUMENT, "unknown BVH builder");
    uint64_t total_tris = 0;
    for (uint32_t m = 0; m < d->mesh_count; m++) total_tris += d->meshes[m].index_count / 3;
    if (total_tris >= (1ull << 27)) return fail(c, TT_INVALID_ARGUMENT, "too many triangles for the 27-bit leaf encoding");
    tt::BvhBuildResult bvh;
    if (builder == TT_BVH_BUILDER_HOST_SAH && !tt::build_bvh(*d, bvh))
        return fail(c, TT_INVALID_ARGUMENT, "BVH build failed: " + bvh.error);

    tt_scene* s = new (std::nothrow) tt_scene();
    static std::atomic<uint64_t> next_serial{1};
    if (s) s->serial = next_serial.fetch_add(1);
    if (!s) return fail(c, TT_OUT_OF_MEMORY, "scene allocation");
    s->context = c;
    tt::context_retain(c);
    tt_status st = TT_OK;
    hipError_t de = hipSetDevice(c->device);
    if (de != hipSuccess) st = hip_fail(c, de, "hipSetDevice");
    auto up = [&](const void* src, size_t bytes, const void** dst) {
        if (st == TT_OK) st = upload(c, s, src, bytes, dst);
    };
    up(d->instances, size_t(d->instance_count) * sizeof(tt_instance), &s->view.instances);
    up(d->materials, size_t(d->material_count) * sizeof(tt_material), &s->view.materials);
    up(d->positions, size_t(d->position_count) * 3 * sizeof(int16_t), &s->view.positions);
    up(d->vertices, size_t(d->vertex_count) * sizeof(tt_vertex), &s->view.vertices);
    up(d->indices, size_t(d->index_count) * sizeof(uint32_t), &s->view.indices);
    up(d->emissive_triangles, size_t(d->emissive_triangle_count) * sizeof(tt_emissive_triangle), &s->view.emissive);
    if (d->texture_count) {
        // Texture records, texel bytes and the sRGB decode table
        // (c <= 0.04045 ? c / 12.92 : ((c + 0.055) / 1.055)^2.4, c = i / 255).
        float lut[256];
        for (int i = 0; i < 256; i++) {
            double cv = double(i) / 255.0;
            lut[i] = float(cv <= 0.04045 ? cv / 12.92 : std::pow((cv + 0.055) / 1.055, 2.4));
        }
        up(d->textures, size_t(d->texture_count) * sizeof(tt_texture_record), &s->view.tex_records);
        up(d->texture_data, size_t(d->texture_data_bytes), &s->view.tex_data);
        up(lut, sizeof(lut), &s->view.srgb_lut);
        s->view.texture_count = d->texture_count;
    }
    uint32_t node_count = 0;
    if (st == TT_OK && builder == TT_BVH_BUILDER_DEVICE_LBVH) {
        // Every BLAS on the device, then the TLAS on the host after them.
        uint32_t max_internal = uint32_t(total_tris);
        size_t node_bytes = (size_t(max_internal) + 2 * size_t(d->instance_count) + 2) * sizeof(tt::BvhNode);
        size_t tri_bytes = std::max<size_t>(total_tris, 1) * sizeof(tt::BvhTri);
        void* nodes = nullptr;
        void* tris_dev = nullptr;
        hipError_t e = hipMalloc(&nodes, node_bytes);
        if (e == hipSuccess) {
            s->allocations.push_back(nodes);
            e = hipMalloc(&tris_dev, tri_bytes);
        }
        if (e == hipSuccess) s->allocations.push_back(tris_dev);
        if (e != hipSuccess) st = hip_fail(c, e, "BVH allocation");
        std::vector<tt::BlasSummary> summary;
        uint32_t blas_nodes = 0;
        if (st == TT_OK)
            st = tt::build_blas_device(c, *d, static_cast<const int16_t*>(s->view.positions),
                                       static_cast<const uint32_t*>(s->view.indices), static_cast<tt::BvhNode*>(nodes),
                                       static_cast<tt::BvhTri*>(tris_dev), summary, blas_nodes);
        if (st == TT_OK && !tt::build_tlas(*d, summary, blas_nodes, bvh))
            st = fail(c, TT_INVALID_ARGUMENT, "BVH build failed: " + bvh.error);
        if (st == TT_OK && !bvh.nodes.empty()) {
            e = hipMemcpy(static_cast<tt::BvhNode*>(nodes) + blas_nodes, bvh.nodes.data(),
                          bvh.nodes.size() * sizeof(tt::BvhNode), hipMemcpyHostToDevice);
            if (e != hipSuccess) st = hip_fail(c, e, "TLAS upload");
        }
        if (st == TT_OK) {
            s->view.nodes = nodes;
            s->view.tris = tris_dev;
            s->info.device_bytes += size_t(blas_nodes + bvh.nodes.size()) * sizeof(tt::BvhNode) +
                                    size_t(total_tris) * sizeof(tt::BvhTri);
            node_count = blas_nodes + uint32_t(bvh.nodes.size());
        }
    } else {
        up(bvh.nodes.data(), bvh.nodes.size() * sizeof(tt::BvhNode), &s->view.nodes);
        up(bvh.tris.data(), bvh.tris.size() * sizeof(tt::BvhTri), &s->view.tris);
        node_count = uint32_t(bvh.nodes.size());
    }
    up(bvh.instances.data(), bvh.instances.size() * sizeof(tt::BvhInstance), &s->view.bvh_insts);
    up(bvh.instance_boxes.data(), bvh.instance_boxes.size() * sizeof(float), &s->view.inst_boxes);
    // Traversal reads 4-wide nodes: collapse the binary ones (renumbering
    // the instance records' BLAS roots and the TLAS root), then free them.
    uint32_t tlas_root = bvh.empty ? 0xffffffffu : bvh.tlas_root;
    if (st == TT_OK) {
        const void* binary = s->view.nodes;
        tt::BvhWide* wide = nullptr;
        uint32_t wide_count = 0;
        st = tt::collapse_wide(c, static_cast<const tt::BvhNode*>(binary), binary ? node_count : 0u,
                               const_cast<void*>(s->view.bvh_insts), uint32_t(bvh.instances.size()), tlas_root, &wide,
                               wide_count);
        if (st == TT_OK) {
            auto it = std::find(s->allocations.begin(), s->allocations.end(), binary);
            if (it != s->allocations.end()) {
                s->allocations.erase(it);
                (void)hipFree(const_cast<void*>(binary));
            }
            if (wide) s->allocations.push_back(wide);
            s->view.nodes = wide;
            s->info.device_bytes -= size_t(node_count) * sizeof(tt::BvhNode);
            s->info.device_bytes += size_t(wide_count) * sizeof(tt::BvhWide);
            node_count = wide_count;
        }
    }
    if (st != TT_OK) {
        tt_scene_destroy(s);
        return st;
    }
    s->view.tlas_root = tlas_root;
    for (int a = 0; a < 3; a++) {
        float lo = INFINITY, hi = -INFINITY;
        for (size_t i = 0; i < bvh.instance_boxes.size(); i += 8) {
            lo = std::min(lo, bvh.instance_boxes[i + a]);
            hi = std::max(hi, bvh.instance_boxes[i + 4 + a]);
        }
        s->bounds_lo[a] = lo <= hi ? lo : 0.0f;
        s->bounds_hi[a] = lo <= hi ? hi : 0.0f;
    }
    s->view.emissive_count = d->emissive_triangle_count;
    if (d->emissive_triangle_count > 0) {
        // World-space light records, computed once on the device.
        size_t bytes = size_t(d->emissive_triangle_count) * 5 * sizeof(float4);
        void* lights = nullptr;
        hipError_t e = hipMalloc(&lights, bytes);
        if (e == hipSuccess) {
            s->allocations.push_back(lights);
            s->info.device_bytes += bytes;
            e = tt::launch_light_records(tt::api_stream(c), s->view, lights);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(tt::api_stream(c));
        if (e != hipSuccess) {
            tt_scene_destroy(s);
            return hip_fail(c, e, "light records");
        }
        s->view.lights = lights;
    }
    uint64_t tris = 0, unique = 0;
    for (uint32_t i = 0; i < d->instance_count; i++) tris += d->meshes[d->instances[i].mesh_index].index_count / 3;
    for (uint32_t m = 0; m < d->mesh_count; m++) unique += d->meshes[m].index_count / 3;
    s->info.triangle_count = uint32_t(std::min<uint64_t>(tris, 0xffffffffu));
    s->info.unique_triangle_count = uint32_t(unique);
    s->info.node_count = node_count;
    s->info.max_depth = bvh.max_depth;
    s->info.sah_cost = bvh.sah_cost;
    s->info.expected_bytes_per_ray = bvh.expected_bytes;
    s->info.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = s;
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

void tt_scene_destroy(tt_scene* s) try {
    if (!s) return;
    if (s->context) {
        (void)hipSetDevice(s->context->device);
        (void)hipStreamSynchronize(tt::api_stream(s->context));
    }
    for (void* p : s->allocations) (void)hipFree(p);
    tt_context* c = s->context;
    delete s;
    if (c) tt::context_release(c);
} catch (...) {
}

tt_status tt_scene_info_get(const tt_scene* s, tt_scene_info* out) {
    if (!s || !out) return TT_INVALID_ARGUMENT;
    *out = s->info;
    return TT_OK;
}

tt_status tt_target_create(tt_context* c, uint32_t width, uint32_t height, tt_target** out) try {
    if (!c || !out || width == 0 || height == 0) return fail(c, TT_INVALID_ARGUMENT, "bad target size");
    *out = nullptr;
    tt_target* t = new (std::nothrow) tt_target();
    if (!t) return fail(c, TT_OUT_OF_MEMORY, "target allocation");
    t->context = c;
    tt::context_retain(c);
    t->width = width;
    t->height = height;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipMalloc(&t->pixels, size_t(width) * height * sizeof(float4));
    if (e == hipSuccess) e = hipMemsetAsync(t->pixels, 0, size_t(width) * height * sizeof(float4), tt::api_stream(c));
    if (e != hipSuccess) {
        tt_target_destroy(t);
        return hip_fail(c, e, "target allocation");
    }
    *out = t;
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

void tt_target_destroy(tt_target* t) try {
    if (!t) return;
    if (t->context) {
        (void)hipSetDevice(t->context->device);
        (void)hipStreamSynchronize(tt::api_stream(t->context));
    }
    if (t->pixels) (void)hipFree(t->pixels);
    if (t->restir_memory) (void)hipFree(t->restir_memory);
    if (t->restir_done) (void)hipEventDestroy(t->restir_done);
    tt_context* c = t->context;
    delete t;
    if (c) tt::context_release(c);
} catch (...) {
}

}  // extern "C"

namespace {

// tt_render_ex, optionally with the radiance cache (tt_render_cached).
tt_status render(tt_context* c, const tt_scene* s, const tt_constants* k, const tt_render_settings* settings,
                 tt_target* t, tt_mlp* cache, const tt_cache_settings* cache_settings, tt_cache_stats* cache_stats) {
    if (!c || !s || !k || !t) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (s->context != c || t->context != c) return fail(c, TT_STALE_HANDLE, "handle from another context");
    if (k->screen_size[0] != t->width || k->screen_size[1] != t->height)
        return fail(c, TT_INVALID_ARGUMENT, "constants.screen_size does not match the target");
    if (k->sample_count == 0 || k->bounce_count == 0 || k->bounce_count > 31)
        return fail(c, TT_INVALID_ARGUMENT, "sample_count must be >= 1 and bounce_count in [1, 31]");
    if (k->light_sampling != TT_LIGHT_SAMPLING_ON_HIT && k->light_sampling != TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION &&
        k->light_sampling != TT_LIGHT_SAMPLING_RESTIR)
        return fail(c, TT_UNSUPPORTED, "light sampling mode not supported");
    tt_render_settings def{TT_RNG_REFERENCE, 0, 0, 0, 0, 0};
    const tt_render_settings& st = settings ? *settings : def;
    const bool restir = k->light_sampling == TT_LIGHT_SAMPLING_RESTIR;
    if (restir && k->sample_count != 1) return fail(c, TT_UNSUPPORTED, "ReSTIR DI renders one sample per pixel");
    if (restir && st.restir_candidates > 1024) return fail(c, TT_INVALID_ARGUMENT, "restir_candidates > 1024");
    if (st.rng > 1) return fail(c, TT_INVALID_ARGUMENT, "unknown generator");
    uint64_t paths64 = uint64_t(t->width) * t->height * k->sample_count;
    if (paths64 >= (1ull << 31)) return fail(c, TT_INVALID_ARGUMENT, "too many paths in one render");
    uint32_t paths = uint32_t(paths64);
    TT_HIP(c, hipSetDevice(c->device));
    bool stats = (st.flags & TT_RENDER_COLLECT_STATS) != 0;
    // A render takes the next lane, unless every lane's path buffers at this
    // size would pass the budget: such renders are long enough that
    // overlapping their tails gains little, and they run on lane 0.
    const bool fits = uint64_t(paths) * tt::PATH_BYTES * c->lane_count <= tt::LANE_PATH_BUDGET;
    const uint32_t lane = !fits ? 0u : c->next_lane;
    tt::Lane& L = c->lanes[lane];
    tt::PathBuffers b;
    tt_status es = tt::ensure_paths(c, lane, paths, b);
    if (es != TT_OK) return es;
    if (fits)
        for (uint32_t i = 0; i < c->lane_count; i++) {
            // Grow every lane together, so later renders allocate nothing.
            tt::PathBuffers unused;
            if (i != lane && c->lanes[i].path_capacity < paths) {
                es = tt::ensure_paths(c, i, paths, unused);
                if (es != TT_OK) return es;
            }
        }
    b.max_bounces = k->bounce_count;

    tt::RestirBuffers rb{};
    b.restir_final = nullptr;
    if (restir) {
        size_t px = size_t(t->width) * t->height;
        if (!t->restir_done) TT_HIP(c, hipEventCreateWithFlags(&t->restir_done, hipEventDisableTiming));
        if (!t->restir_memory) {
            // Two G-buffers of 5 float4 records, four reservoir arrays.
            TT_HIP(c, hipMalloc(&t->restir_memory, px * (10 * sizeof(float4) + 4 * sizeof(uint4))));
            t->restir_cur = 0;
            t->restir_history = false;
        }
        float4* g = static_cast<float4*>(t->restir_memory);
        uint4* r = reinterpret_cast<uint4*>(g + 10 * px);
        rb.gbuffer[0] = g;
        rb.gbuffer[1] = g + 5 * px;
        rb.initial = r;
        rb.temporal = r + px;
        rb.final_[0] = r + 2 * px;
        rb.final_[1] = r + 3 * px;
        // History from another scene would index its lights and materials.
        if (t->restir_scene != s->serial) t->restir_history = false;
        rb.cur = t->restir_cur;
        rb.history_valid = t->restir_history ? 1u : 0u;
        rb.candidates = st.restir_candidates ? st.restir_candidates : 32u;
        rb.temporal_reuse = st.restir_temporal ? 1u : 0u;
        rb.spatial_reuse = st.restir_spatial ? 1u : 0u;
        rb.base = 0;
        b.restir_final = rb.final_[rb.cur];
        b.restir_initial = (rb.temporal_reuse || rb.spatial_reuse) ? rb.initial : rb.final_[rb.cur];
    }

    tt::FrameArgs f = tt::frame_args(*k, t->width, t->height, st.rng, st.seed);
    f.primary_tiles = 1;
    uint32_t train_steps = 0, train_capacity = 0, judge_capacity = 0, train_order_span = 0;
    if (cache) {
        if (tt::mlp_context(cache) != c) return fail(c, TT_STALE_HANDLE, "cache from another context");
        tt::CachePlan plan;
        const char* why = nullptr;
        if (tt::cache_plan(cache_settings, paths, plan, &why) != TT_OK) return fail(c, TT_INVALID_ARGUMENT, why);
        train_steps = plan.train_steps;
        train_capacity = plan.train_capacity;
        judge_capacity = plan.judge_capacity;
        train_order_span = plan.ordered ? paths * plan.records : 0u;
        tt::apply_cache_plan(plan, f);
        tt_status cst = tt::ensure_cache(c, lane, paths, train_capacity, judge_capacity, b.cache);
        if (cst != TT_OK) return cst;
        cst = tt::mlp_set_scene_bounds(cache, s->bounds_lo, s->bounds_hi);
        if (cst != TT_OK) return cst;
    }

    // After the host calls' work on the API stream (uploads, target clears);
    // the lane's previous render is ahead on its own stream.
    if (c->api_dirty) {
        TT_HIP(c, hipEventRecord(c->api_mark, c->stream));
        c->api_dirty = false;
    }
    hipStream_t stream = L.stream;
    TT_HIP(c, hipStreamWaitEvent(stream, c->api_mark, 0));
    // If an enqueue fails part way, still order what was enqueued before
    // what comes later: join the side stream back, mark the target's ReSTIR
    // buffers as read once this render's passes end, and mark the lane's end,
    // so later API calls and renders wait for it.
    struct PartialGuard {
        tt::Lane* lane;
        hipStream_t stream;
        hipStream_t side;
        tt_target* target;
        bool on_lane;
        bool restir_enqueued;
        bool armed;
        ~PartialGuard() {
            if (!armed) return;
            if (side && side != stream) {
                (void)hipEventRecord(lane->join, side);
                (void)hipStreamWaitEvent(stream, lane->join, 0);
            }
            if (restir_enqueued) {
                (void)hipEventRecord(target->restir_done, stream);
                target->restir_done_recorded = true;
            }
            if (on_lane) {
                (void)hipEventRecord(lane->done, stream);
                lane->pending = true;
            }
        }
    } partial{&L, stream, nullptr, t, true, false, true};
    // The previous render, if on another lane, may still be running: this
    // one accumulates after it (the accumulations of consecutive renders form
    // a chain).  ReSTIR buffers are ordered by the target's own event.
    const tt::Lane* other = c->prev_lane >= 0 && uint32_t(c->prev_lane) != lane ? &c->lanes[c->prev_lane] : nullptr;
    const bool after_other = other && other->used;
    TT_HIP(c, hipEventRecord(L.integrate.begin, stream));
    size_t whole = 0;
    if (stats) {
        c->last_stats_first = c->stats_launches.size();
        es = stats_begin(c, tt::STATS_RENDER, stream, &whole);
        if (es != TT_OK) return es;
    }
    TT_HIP(c, tt::launch_zero(stream, L.counters, tt::C_WORDS));
    TT_HIP(c, tt::launch_zero_work(stream, L.work, tt::W_SETS));
    TT_HIP(c, tt::launch_generate(stream, f, b, paths));
    size_t ev = 0;
    // Bounce b's shadow rays (side stream) overlap bounce b + 1's extension
    // (main stream): both only read what shade b wrote.  Shade b + 1 waits
    // for them, since the shadow pass adds to the paths' radiance and its
    // buffers are refilled by the next shade.
    tt::PathBuffers side_buffers = b;
    side_buffers.spill = L.side_spill;
    hipStream_t side = c->side_streams ? L.side : stream;
    partial.side = side;
    const bool shadows = f.light_sampling != TT_LIGHT_SAMPLING_ON_HIT && s->view.emissive_count > 0;
    auto extend = [&](uint32_t bounce) -> tt_status {
        if (stats) {
            tt_status ss = stats_begin(c, tt::STATS_EXTEND, stream, &ev);
            if (ss != TT_OK) return ss;
        }
        TT_HIP(c, tt::launch_extend(stream, s->view, b, bounce, c->grid));
        return stats ? stats_end(c, ev, stream) : TT_OK;
    };
    es = extend(0);
    if (es != TT_OK) return es;
    for (uint32_t bounce = 0; bounce < k->bounce_count; bounce++) {
        if (restir && bounce == 0 && s->view.emissive_count > 0) {
            // The previous ReSTIR render on this target, on whatever lane,
            // has finished reading the buffers these passes rewrite.
            if (t->restir_done_recorded) TT_HIP(c, hipStreamWaitEvent(stream, t->restir_done, 0));
            partial.restir_enqueued = true;
            if (stats) {
                es = stats_begin(c, tt::STATS_RESTIR, stream, &ev);
                if (es != TT_OK) return es;
            }
            TT_HIP(c, tt::launch_restir(stream, s->view, f, b, rb, c->grid));
            if (stats) {
                es = stats_end(c, ev, stream);
                if (es != TT_OK) return es;
            }
        }
        if (shadows && bounce > 0 && side != stream) TT_HIP(c, hipStreamWaitEvent(stream, L.join, 0));
        TT_HIP(c, tt::launch_shade(stream, s->view, f, b, bounce, paths));
        // Shade 0 is the last reader of the ReSTIR buffers.
        if (bounce == 0 && partial.restir_enqueued) {
            TT_HIP(c, hipEventRecord(t->restir_done, stream));
            t->restir_done_recorded = true;
            partial.restir_enqueued = false;
        }
        if (shadows) {
            if (side != stream) {
                TT_HIP(c, hipEventRecord(L.fork, stream));
                TT_HIP(c, hipStreamWaitEvent(side, L.fork, 0));
            }
            if (stats) {
                es = stats_begin(c, tt::STATS_SHADOW, side, &ev);
                if (es != TT_OK) return es;
            }
            TT_HIP(c, tt::launch_shadow(side, s->view, side_buffers, bounce, c->grid));
            if (stats) {
                es = stats_end(c, ev, side);
                if (es != TT_OK) return es;
            }
            if (side != stream) TT_HIP(c, hipEventRecord(L.join, side));
        }
        if (bounce + 1 < k->bounce_count) {
            es = extend(bounce + 1);
            if (es != TT_OK) return es;
        }
    }
    if (shadows && side != stream) TT_HIP(c, hipStreamWaitEvent(stream, L.join, 0));
    if (cache) {
        // Cached radiance for the paths that ended at a query, and the
        // cache's error at the judge records: both with the weights of the
        // last training enqueued (the previous frame's).
        if (cache_stats) {
            if (!L.cache_infer.begin) TT_HIP(c, hipEventCreate(&L.cache_infer.begin));
            if (!L.cache_infer.end) TT_HIP(c, hipEventCreate(&L.cache_infer.end));
            // After the last training, so the pair times the inference alone.
            tt_status ws = tt::mlp_wait_trained(cache, stream);
            if (ws != TT_OK) return ws;
            TT_HIP(c, hipEventRecord(L.cache_infer.begin, stream));
        }
        tt_status cs = tt::mlp_infer_device(cache, stream, b.cache.query_rec, b.cache.query_scale,
                                            L.counters + tt::C_CACHE_QUERY, paths, nullptr, b.radiance);
        if (cs != TT_OK) return cs;
        if (cache_stats) TT_HIP(c, hipEventRecord(L.cache_infer.end, stream));
        if (judge_capacity) {
            cs = tt::mlp_judge_device(cache, stream, b.cache.train_rec + 4 * size_t(b.cache.judge_base),
                                      b.cache.train_target + b.cache.judge_base, L.counters + tt::C_CACHE_JUDGE,
                                      judge_capacity, L.judge_pred, L.judge_stats);
            if (cs != TT_OK) return cs;
        }
    }
    if (after_other) TT_HIP(c, hipStreamWaitEvent(stream, other->accumulated, 0));
    TT_HIP(c, tt::launch_accumulate(stream, f, b, t->pixels));
    TT_HIP(c, hipEventRecord(L.accumulated, stream));
    if (stats) {
        TT_HIP(c, tt::launch_stats_add(stream, L.counters, c->stats_sum, tt::C_WORDS));
        es = stats_end(c, whole, stream);
        if (es != TT_OK) return es;
        c->stats_renders++;
        c->stats_paths += paths;
    }
    if (cache) {
        // Train on this frame's samples once its tracing is done: train_steps
        // Adam steps over consecutive slices of the records, on the API
        // stream.  The lane's queues are free again once training has read
        // them, so the lane's end waits for it.
        es = tt::mlp_train_device(cache, stream, b.cache.train_rec, b.cache.train_target,
                                  L.counters + tt::C_CACHE_TRAIN, train_capacity, train_steps, true, nullptr,
                                  cache_stats != nullptr, train_order_span);
        if (es != TT_OK) return es;
        es = tt::mlp_wait_trained(cache, stream);
        if (es != TT_OK) return es;
    }
    TT_HIP(c, hipEventRecord(L.integrate.end, stream));
    TT_HIP(c, hipEventRecord(L.done, stream));
    partial.armed = false;
    L.used = true;
    c->last_lane = lane;
    L.pending = true;
    c->prev_lane = int(lane);
    c->next_lane = (lane + 1u) % c->lane_count;
    if (cache && cache_stats) {
        // Waits for the frame and its training.
        uint32_t counts[3] = {0u, 0u, 0u};
        double js[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        TT_HIP(c, hipMemcpyAsync(counts, L.counters + tt::C_CACHE_QUERY, sizeof(counts), hipMemcpyDeviceToHost,
                                 stream));
        if (judge_capacity)
            TT_HIP(c, hipMemcpyAsync(js, L.judge_stats, sizeof(js), hipMemcpyDeviceToHost, stream));
        TT_HIP(c, hipStreamSynchronize(stream));
        const uint32_t samples = std::min(counts[1], train_capacity);
        double loss = 0.0;
        es = tt::mlp_first_step_loss(cache, samples, train_steps, &loss);
        if (es != TT_OK) return es;
        *cache_stats = tt_cache_stats{};
        cache_stats->queries = counts[0];
        cache_stats->samples = samples;
        cache_stats->loss = loss;
        const double n = js[0];
        cache_stats->judge_samples = uint32_t(n);
        if (n > 0.0) {
            cache_stats->judge_cached = js[1] / (3.0 * n);
            cache_stats->judge_traced = js[2] / (3.0 * n);
            cache_stats->judge_loss = js[5] / (3.0 * n);
        }
        float ms = 0.0f;
        TT_HIP(c, hipEventElapsedTime(&ms, L.cache_infer.begin, L.cache_infer.end));
        cache_stats->infer_ms = ms;
        es = tt::mlp_train_ms(cache, &cache_stats->train_ms);
        if (es != TT_OK) return es;
    }
    c->last_integrate_ms = -1.0;
    if (restir && s->view.emissive_count > 0) {
        t->restir_cur ^= 1u;
        t->restir_history = true;
        t->restir_scene = s->serial;
    } else {
        t->restir_history = false;
    }
    return TT_OK;
}

}  // namespace

extern "C" {

tt_status tt_render_ex(tt_context* c, const tt_scene* s, const tt_constants* k, const tt_render_settings* settings,
                       tt_target* t) try {
    return render(c, s, k, settings, t, nullptr, nullptr, nullptr);
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_render_cached(tt_context* c, const tt_scene* s, const tt_constants* k, const tt_render_settings* settings,
                           tt_mlp* cache, const tt_cache_settings* cache_settings, tt_target* t,
                           tt_cache_stats* stats) try {
    if (!cache) return fail(c, TT_INVALID_ARGUMENT, "null cache");
    return render(c, s, k, settings, t, cache, cache_settings, stats);
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_render(tt_context* c, const tt_scene* s, const tt_constants* k, tt_target* t) {
    return tt_render_ex(c, s, k, nullptr, t);
}

tt_status tt_target_download(tt_context* c, const tt_target* t, float* rgba, size_t bytes) try {
    if (!c || !t || !rgba) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (t->context != c) return fail(c, TT_STALE_HANDLE, "target from another context");
    size_t need = size_t(t->width) * t->height * sizeof(float4);
    if (bytes < need) return fail(c, TT_INVALID_ARGUMENT, "download buffer too small");
    TT_HIP(c, hipSetDevice(c->device));
    TT_HIP(c, hipMemcpyAsync(rgba, t->pixels, need, hipMemcpyDeviceToHost, tt::api_stream(c)));
    TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_timing_ms(tt_context* c, const char* name, double* out) try {
    if (!c || !name || !out) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    const tt::EventPair& integrate = c->lanes[c->last_lane].integrate;
    TT_HIP(c, hipEventSynchronize(integrate.end));
    if (std::strcmp(name, "integrate") == 0) {
        float ms = 0.0f;
        TT_HIP(c, hipEventElapsedTime(&ms, integrate.begin, integrate.end));
        *out = ms;
        return TT_OK;
    }
    if (std::strcmp(name, "post_process") == 0) {
        if (!c->post_process_recorded) return fail(c, TT_INVALID_ARGUMENT, "no tt_target_display yet");
        TT_HIP(c, hipEventSynchronize(c->post_process.end));
        float ms = 0.0f;
        TT_HIP(c, hipEventElapsedTime(&ms, c->post_process.begin, c->post_process.end));
        *out = ms;
        return TT_OK;
    }
    if (std::strcmp(name, "trace") == 0) {
        if (c->stats_renders == 0)
            return fail(c, TT_INVALID_ARGUMENT, "no render with TT_RENDER_COLLECT_STATS since the statistics were read");
        TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
        double sum = 0.0;
        for (size_t i = c->last_stats_first; i < c->stats_launches.size(); i++) {
            const tt::StatsLaunch& l = c->stats_launches[i];
            if (l.kind != tt::STATS_EXTEND && l.kind != tt::STATS_SHADOW) continue;
            float ms = 0.0f;
            TT_HIP(c, hipEventElapsedTime(&ms, l.events.begin, l.events.end));
            sum += ms;
        }
        *out = sum;
        return TT_OK;
    }
    return fail(c, TT_INVALID_ARGUMENT, std::string("unknown timing ") + name);
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_render_stats_get(tt_context* c, tt_render_stats* out) try {
    if (!c || !out) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (c->stats_renders == 0)
        return fail(c, TT_INVALID_ARGUMENT, "no render with TT_RENDER_COLLECT_STATS since the statistics were read");
    TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
    unsigned long long words[tt::C_WORDS];
    TT_HIP(c, hipMemcpy(words, c->stats_sum, sizeof(words), hipMemcpyDeviceToHost));
    *out = tt_render_stats{};
    out->paths = c->stats_paths;
    for (uint32_t b = 0; b < 32; b++) {
        out->extension_rays += words[tt::C_EXT + b];
        out->shadow_rays += words[tt::C_SHADOW + b];
    }
    out->restir_rays = words[tt::C_RESTIR_VIS];
    for (const tt::StatsLaunch& l : c->stats_launches) {
        float ms = 0.0f;
        TT_HIP(c, hipEventElapsedTime(&ms, l.events.begin, l.events.end));
        switch (l.kind) {
            case tt::STATS_EXTEND:
                out->extend_launches++;
                out->extend_ms += ms;
                break;
            case tt::STATS_SHADOW:
                out->shadow_launches++;
                out->shadow_ms += ms;
                break;
            case tt::STATS_RESTIR:
                out->restir_ms += ms;
                break;
            case tt::STATS_RENDER:
                out->integrate_ms += ms;
                break;
        }
    }
    // Start the next sum.
    TT_HIP(c, hipMemset(c->stats_sum, 0, sizeof(words)));
    for (const tt::StatsLaunch& l : c->stats_launches) c->stats_pool.push_back(l.events);
    c->stats_launches.clear();
    c->last_stats_first = 0;
    c->stats_renders = 0;
    c->stats_paths = 0;
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_trace_closest_device(tt_context* c, const tt_scene* s, const tt_ray* rays, tt_hit* hits, uint32_t count,
                                  uint32_t flags) try {
    if (!c || !s || (count && (!rays || !hits))) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (s->context != c) return fail(c, TT_STALE_HANDLE, "scene from another context");
    if (count == 0) return TT_OK;
    TT_HIP(c, hipSetDevice(c->device));
    TT_HIP(c, tt::launch_zero_work(tt::api_stream(c), c->work, 1));
    TT_HIP(c, tt::launch_trace_rays(tt::api_stream(c), s->view, rays, hits, count, (flags & TT_TRACE_CULL_BACK_FACES) != 0,
                                    c->work, c->lanes[0].spill, c->grid));
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_trace_closest(tt_context* c, const tt_scene* s, const tt_ray* rays, tt_hit* hits, uint32_t count,
                           uint32_t flags) try {
    if (!c || !s || (count && (!rays || !hits))) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (count == 0) return TT_OK;
    TT_HIP(c, hipSetDevice(c->device));
    tt_ray* dr = nullptr;
    tt_hit* dh = nullptr;
    TT_HIP(c, hipMalloc(&dr, size_t(count) * sizeof(tt_ray)));
    hipError_t e = hipMalloc(&dh, size_t(count) * sizeof(tt_hit));
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rays, size_t(count) * sizeof(tt_ray), hipMemcpyHostToDevice, tt::api_stream(c));
    tt_status st = TT_OK;
    if (e == hipSuccess) st = tt_trace_closest_device(c, s, dr, dh, count, flags);
    if (e == hipSuccess && st == TT_OK)
        e = hipMemcpyAsync(hits, dh, size_t(count) * sizeof(tt_hit), hipMemcpyDeviceToHost, tt::api_stream(c));
    if (e == hipSuccess && st == TT_OK) e = hipStreamSynchronize(tt::api_stream(c));
    (void)hipFree(dr);
    if (dh) (void)hipFree(dh);
    if (e != hipSuccess) return hip_fail(c, e, "tt_trace_closest");
    return st;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_trace_closest_static(tt_context* c, const tt_scene* s, const tt_ray* rays, tt_hit* hits, uint32_t count,
                                  uint32_t flags) try {
    if (!c || !s || (count && (!rays || !hits))) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (s->context != c) return fail(c, TT_STALE_HANDLE, "scene from another context");
    if (count > STATIC_TRACE_MAX_RAYS)
        return fail(c, TT_INVALID_ARGUMENT,
                    "tt_trace_closest_static traces at most " + std::to_string(STATIC_TRACE_MAX_RAYS) +
                        " rays per call (got " + std::to_string(count) + ")");
    if (count == 0) return TT_OK;
    if (!c->static_scratch) return fail(c, TT_KERNEL_LOAD, "static trace scratch unavailable: " + c->static_scratch_error);
    TT_HIP(c, hipSetDevice(c->device));
    uint8_t* base = static_cast<uint8_t*>(c->static_scratch);
    uint32_t* cursor = reinterpret_cast<uint32_t*>(base + ST_CURSOR);
    tt_ray* staged_rays = reinterpret_cast<tt_ray*>(base + ST_RAYS);
    tt_hit* staged_hits = reinterpret_cast<tt_hit*>(base + ST_HITS);
    uint32_t* spill = reinterpret_cast<uint32_t*>(base + ST_SPILL);
    // The spill stride is the launch's grid size, so any grid up to
    // STATIC_TRACE_GRID stays inside the spill area.
    const int grid = int(std::min<uint32_t>(STATIC_TRACE_GRID, (count + tt::TRACE_THREADS - 1) / tt::TRACE_THREADS));
    hipStream_t stream = tt::api_stream(c);
    // The cursors keep the previous call's final values: clear them with a
    // kernel (no memset node, so the sequence can also be captured).
    TT_HIP(c, tt::launch_zero_work(stream, cursor, 1));
    TT_HIP(c, hipMemcpyAsync(staged_rays, rays, size_t(count) * sizeof(tt_ray), hipMemcpyHostToDevice, stream));
    TT_HIP(c, tt::launch_trace_rays(stream, s->view, staged_rays, staged_hits, count,
                                    (flags & TT_TRACE_CULL_BACK_FACES) != 0, cursor, spill, grid));
    TT_HIP(c, hipMemcpyAsync(hits, staged_hits, size_t(count) * sizeof(tt_hit), hipMemcpyDeviceToHost, stream));
    TT_HIP(c, hipStreamSynchronize(stream));
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_bsdf_evaluate(tt_context* c, const tt_material_constants* material, const tt_bsdf_query* queries,
                           tt_bsdf_result* results, uint32_t count, uint32_t rng_state) try {
    if (!c || !material || (count && (!queries || !results))) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (count == 0) return TT_OK;
    TT_HIP(c, hipSetDevice(c->device));
    tt_bsdf_query* dq = nullptr;
    tt_bsdf_result* dr = nullptr;
    TT_HIP(c, hipMalloc(&dq, size_t(count) * sizeof(tt_bsdf_query)));
    hipError_t e = hipMalloc(&dr, size_t(count) * sizeof(tt_bsdf_result));
    if (e == hipSuccess) e = hipMemcpyAsync(dq, queries, size_t(count) * sizeof(tt_bsdf_query), hipMemcpyHostToDevice, tt::api_stream(c));
    if (e == hipSuccess) e = tt::launch_bsdf_evaluate(tt::api_stream(c), material, dq, dr, count, rng_state);
    if (e == hipSuccess)
        e = hipMemcpyAsync(results, dr, size_t(count) * sizeof(tt_bsdf_result), hipMemcpyDeviceToHost, tt::api_stream(c));
    if (e == hipSuccess) e = hipStreamSynchronize(tt::api_stream(c));
    (void)hipFree(dq);
    if (dr) (void)hipFree(dr);
    if (e != hipSuccess) return hip_fail(c, e, "tt_bsdf_evaluate");
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

tt_status tt_debug_math(tt_context* c, int op, const float* x, const float* y, float* out, uint32_t count) try {
    if (!c || !x || !y || !out) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (count == 0) return TT_OK;
    TT_HIP(c, hipSetDevice(c->device));
    float* d = nullptr;
    TT_HIP(c, hipMalloc(&d, size_t(count) * 3 * sizeof(float)));
    hipError_t e = hipMemcpyAsync(d, x, count * sizeof(float), hipMemcpyHostToDevice, tt::api_stream(c));
    if (e == hipSuccess) e = hipMemcpyAsync(d + count, y, count * sizeof(float), hipMemcpyHostToDevice, tt::api_stream(c));
    if (e == hipSuccess) e = tt::launch_math(tt::api_stream(c), op, d, d + count, d + 2 * count, count);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * count, count * sizeof(float), hipMemcpyDeviceToHost, tt::api_stream(c));
    if (e == hipSuccess) e = hipStreamSynchronize(tt::api_stream(c));
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(c, e, "tt_debug_math");
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

}  // extern "C"
