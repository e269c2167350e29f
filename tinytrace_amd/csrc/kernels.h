// Host-side launchers of the gfx950 kernels (defined in kernels.hip and
// restir.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tinytrace.h"

namespace tt {

struct DeviceScene {
    const void* nodes;
    const void* tris;
    const void* bvh_insts;
    const void* inst_boxes;
    uint32_t tlas_root;
    uint32_t emissive_count;
    const void* instances;
    const void* materials;
    const void* positions;
    const void* vertices;
    const void* indices;
    const void* emissive;
    const void* lights;  // per emissive triangle, 5 x float4 (device_surface.h)
    // Textures (none when texture_count == 0): tt_texture_record array,
    // texel bytes and the 256-entry sRGB-to-linear table.
    const void* tex_records;
    const void* tex_data;
    const void* srgb_lut;
    uint32_t texture_count;
};

// Per-render camera/integrator constants passed by value.
struct FrameArgs {
    float inverse_view[16];
    float inverse_proj[16];
    float camera_position[3];
    uint32_t width, height;
    uint32_t sample_count, bounce_count;
    uint32_t frame_index, accumulated_frame_count;
    uint32_t light_sampling;
    uint32_t rng, seed;
    // Radiance cache: 0 off; else a path's cache vertex is its first vertex
    // after the primary one with specular roughness > 0.3.  A hash h of
    // (path, frame) decides there: h < train_threshold makes a training path,
    // h < judge_threshold (>= train_threshold) a judge path, and any other
    // path ends at a cache query.  Training and judge paths run on plainly.
    uint32_t cache_mode;
    uint32_t train_threshold;
    uint32_t judge_threshold;
    uint32_t cache_records;  // training records a training path takes (1 .. CACHE_OPEN_RECORDS)
    // Primary rays in tile order (whole-frame renders): pixel (x, y)'s rays
    // take the queue slots of primary_slot(), so a wave's 64 slots are an
    // 8 x 8 pixel tile.  0: slot = path (serving chunks).
    uint32_t primary_tiles;
};

// Queue slot of pixel (x, y)'s first primary ray when primary rays are in
// tile order: the image is cut into bands of 8 rows (the last band may be
// lower); a band's pixels follow column by column, so 64 consecutive slots
// cover 8 columns of one full band, an 8 x 8 tile.  A bijection of
// [0, width * height).
__host__ __device__ inline uint32_t primary_slot(uint32_t x, uint32_t y, uint32_t width, uint32_t height) {
    const uint32_t band = y / 8u, row0 = band * 8u;
    const uint32_t rows = height - row0 < 8u ? height - row0 : 8u;
    return row0 * width + x * rows + (y - row0);
}

// The pixel whose primary_slot() is `slot`.
__host__ __device__ inline uint32_t slot_pixel(uint32_t slot, uint32_t width, uint32_t height) {
    const uint32_t row0 = (slot / (8u * width)) * 8u;  // a band starts at slot row0 * width
    const uint32_t rows = height - row0 < 8u ? height - row0 : 8u;
    const uint32_t off = slot - row0 * width;
    return (row0 + off % rows) * width + off / rows;
}

// Training records a training path can keep open: its cache vertex and the
// next rough vertices, each the outgoing radiance at that vertex.
constexpr uint32_t CACHE_OPEN_RECORDS = 3;
constexpr uint32_t CACHE_NO_RECORD = 0xffffffffu;
// open[p].w: number of open records (bits 0-1) and this flag for a path
// past its cache vertex (a training or judge path).
constexpr uint32_t CACHE_DECIDED = 4u;

// Radiance-cache queues of one render (null when the cache is off).  Judge
// records share the training arrays at slots [judge_base, judge_base +
// judge_capacity); a record whose slot would pass its capacity is dropped.
struct CacheBuffers {
    uint4* open;            // [paths] open record slots (x, y, z) and state (w)
    float4* query_rec;      // [paths x 4] raw query records (mlp.h)
    float4* query_scale;    // [paths] (throughput, path index)
    float4* train_rec;      // [records x 4] raw training / judge records
    float4* train_thr;      // [records] (throughput at the record, path index)
    float4* train_target;   // [records] radiance the path carries from there on
    uint32_t train_capacity;
    uint32_t judge_base;
    uint32_t judge_capacity;
};

// Wavefront buffers; `paths` = width * height * sample_count.
struct PathBuffers {
    float4* throughput;   // [paths]
    float4* radiance;     // [paths]
    uint64_t* rng;        // [paths] generator state
    float4* queue[2][2];  // [ping-pong][origin|direction], [paths] each
    float4* hits;         // [paths]
    float4* shadow[4];    // origin, direction, direct, emission: [paths] each
    uint32_t* counters;   // see kernels.hip
    uint32_t* work;       // W_SETS cursor sets of the persistent traversal launches
    uint32_t* spill;      // traversal stack spill
    const uint4* restir_final;  // per-path reservoirs used at the primary vertex
    const uint4* restir_initial;  // the path's own candidate after the visibility pass
    // (pixel, frame index) of each path slot for the ReSTIR lobe generator;
    // null when the slot is the pixel and the frame is the render's.
    // Serving batches place passes of several requests in one pool.
    const uint2* restir_key;
    uint32_t max_bounces;
    CacheBuffers cache;
};

// Per-target ReSTIR DI state (Bitterli et al. 2020): a G-buffer record
// (5 x float4) and reservoirs (uint4, FP16 W and M) per pixel.
struct RestirBuffers {
    float4* gbuffer[2];     // [cur] this frame, [cur ^ 1] the previous one
    uint4* initial;         // after candidate RIS + visibility
    uint4* temporal;        // after temporal reuse
    uint4* final_[2];       // after spatial reuse; [cur ^ 1] = previous frame
    uint32_t cur;
    uint32_t history_valid;
    uint32_t candidates;
    uint32_t temporal_reuse;
    uint32_t spatial_reuse;
    // Path slot of pixel 0: the passes read pixel p's primary path (queues,
    // hit, throughput) at slot base + p; G-buffers and reservoirs are indexed
    // by pixel.  0 for renders.
    uint32_t base;
};

int trace_grid(int cu_count);
size_t spill_words(int grid);
// Threads per traversal workgroup and spilled stack entries per thread (the
// traversal's own constants, checked against them in kernels.hip), for host
// code that sizes spill areas at compile time.
constexpr size_t TRACE_THREADS = 128;
constexpr size_t TRACE_SPILL_ENTRIES = 116;
// Counter words (u32, zeroed at the start of every render).
constexpr uint32_t C_EXT = 0;            // [b]: rays in the queue consumed at bounce b
constexpr uint32_t C_SHADOW = 32;        // [b]: shadow rays emitted at bounce b
constexpr uint32_t C_RESTIR_VIS = 128;   // ReSTIR visibility queries
constexpr uint32_t C_CACHE_QUERY = 130;  // radiance-cache queries
constexpr uint32_t C_CACHE_TRAIN = 131;  // radiance-cache training records
constexpr uint32_t C_CACHE_JUDGE = 132;  // radiance-cache judge records
constexpr uint32_t C_WORDS = 160;

// Work cursors of a persistent traversal launch: WORK_PARTS cursors, each
// on a 128-B line of its own (atomics on one line serialise: one cursor for
// a whole launch cost 0.43 ms of atomics for 2M rays on MI355X, 32 cursors
// 26 us).  A render's launches take W_SETS sets: extend and shadow of each
// bounce, and the ReSTIR visibility pass.
constexpr uint32_t WORK_PARTS = 32;
constexpr uint32_t WORK_STRIDE = 32;                      // words
constexpr uint32_t WORK_SET = WORK_PARTS * WORK_STRIDE;   // words per launch
constexpr uint32_t W_EXT = 0;       // [b]: extend of bounce b
constexpr uint32_t W_SHADOW = 32;   // [b]: shadow rays of bounce b
constexpr uint32_t W_RESTIR = 64;   // ReSTIR visibility
constexpr uint32_t W_SETS = 65;

hipError_t launch_zero(hipStream_t s, uint32_t* p, uint32_t words);
// Clears the cursors of `sets` consecutive cursor sets.
hipError_t launch_zero_work(hipStream_t s, uint32_t* work, uint32_t sets);
// key[i] = (first + i, frame) for i < count.
hipError_t launch_slot_keys(hipStream_t s, uint2* key, uint32_t first, uint32_t frame, uint32_t count);
hipError_t launch_stats_add(hipStream_t s, const uint32_t* counters, unsigned long long* sum, uint32_t words);
hipError_t launch_trace_rays(hipStream_t s, const DeviceScene& scene, const tt_ray* rays, tt_hit* hits, uint32_t count,
                             bool cull, uint32_t* work, uint32_t* spill, int grid);
hipError_t launch_generate(hipStream_t s, const FrameArgs& f, const PathBuffers& b, uint32_t paths);
hipError_t launch_extend(hipStream_t s, const DeviceScene& scene, const PathBuffers& b, uint32_t bounce, int grid);
hipError_t launch_shade(hipStream_t s, const DeviceScene& scene, const FrameArgs& f, const PathBuffers& b,
                        uint32_t bounce, uint32_t paths);
hipError_t launch_shadow(hipStream_t s, const DeviceScene& scene, const PathBuffers& b, uint32_t bounce, int grid);
hipError_t launch_accumulate(hipStream_t s, const FrameArgs& f, const PathBuffers& b, float4* target);
// Serving batches: one sample of pixels [pixel0, pixel0 + count) in path
// slots [base, base + count).
hipError_t launch_generate_chunk(hipStream_t s, const FrameArgs& f, const PathBuffers& b, uint32_t base,
                                 uint32_t pixel0, uint32_t count);
hipError_t launch_accumulate_chunk(hipStream_t s, const FrameArgs& f, const PathBuffers& b, uint32_t base,
                                   uint32_t pixel0, uint32_t count, float4* target);
hipError_t launch_bsdf_evaluate(hipStream_t s, const tt_material_constants* material, const tt_bsdf_query* q,
                                tt_bsdf_result* r, uint32_t count, uint32_t rng_state);
// World-space light records of every emissive triangle (scene creation).
hipError_t launch_light_records(hipStream_t s, const DeviceScene& scene, void* out);
hipError_t launch_math(hipStream_t s, int op, const float* x, const float* y, float* out, uint32_t count);

// ReSTIR DI passes at the primary vertex (restir.hip).
hipError_t launch_restir(hipStream_t s, const DeviceScene& scene, const FrameArgs& f, const PathBuffers& b,
                         const RestirBuffers& r, int grid);

}  // namespace tt
