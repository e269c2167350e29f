/*
 * tinytrace_amd C ABI: the drop-in boundary for tinytrace's path-integration
 * hot path on MI355X (gfx950).
 *
 * The reference renders through `Renderer::render_to_target`
 * (src/lib.rs:124-161), which uploads `Constants` and dispatches the
 * integrate megakernel through `Integrator::integrate`
 * (src/integrate/mod.rs:39-61) over a scene that `scene::Scene::new`
 * (src/scene.rs:28-186) uploaded and whose BLAS/TLAS the Vulkan driver built
 * (crates/tinytrace_backend/src/resource.rs:804-878, :1014-1083).  Results are
 * read back with the blocking `Context::download`
 * (crates/tinytrace_backend/src/copy.rs:183-298).  The functions below carry
 * exactly those three capabilities (scene upload + BVH build, integrate +
 * accumulate, download) plus the host-side helpers the Rust library keeps
 * above them (camera constants, asset processing).
 *
 * Conventions
 *  - Every function returns tt_status; no C++ exception crosses the ABI.
 *  - Host arrays passed in are copied before the call returns.
 *  - Handles are owned by the context that created them.  They may be
 *    destroyed in any order: a destroyed context stays alive until its last
 *    scene, target and server is destroyed.  One context per host thread.
 *  - Device work of a context runs on several HIP streams: an API stream
 *    (uploads, downloads, clears, display, serving, the radiance cache's
 *    training and host operations) and
 *    two or three render lanes, each with an optional side stream for
 *    shadow rays, so that back-to-back renders overlap.  The library orders
 *    them with events so that results equal those of calls run one at a
 *    time: every call is ordered after the renders enqueued before it that
 *    it depends on.  tt_target_download is the synchronisation point.
 *  - Structs marked "reference layout" are byte-identical to the reference's
 *    #[repr(C)] / std430 types and are checked with static asserts.
 */
#ifndef TINYTRACE_H
#define TINYTRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  Mirrors tinytrace::Error (src/error.rs:3-9) and
 * tinytrace_backend::Error (crates/tinytrace_backend/src/error.rs:7-28);
 * misuse that panics in the reference (stale handles, missing bindings)
 * returns TT_INVALID_ARGUMENT / TT_STALE_HANDLE instead. */
typedef enum tt_status {
    TT_OK = 0,
    TT_NO_DEVICE = 1,        /* Error::NoSuitableDevice */
    TT_HIP_ERROR = 2,        /* Error::VulkanResult */
    TT_KERNEL_LOAD = 3,      /* Error::Shader */
    TT_INVALID_ARGUMENT = 4, /* reference panics */
    TT_STALE_HANDLE = 5,     /* reference Handle epoch check panics */
    TT_OUT_OF_MEMORY = 6,
    TT_OVER_BUDGET = 7,      /* serving admission */
    TT_UNSUPPORTED = 8,      /* feature outside the implemented path */
    TT_INTERNAL_ERROR = 9    /* a C++ exception inside the library other than a
                                host allocation failure (a bug); the message is
                                in tt_last_error.  Host allocation and thread
                                failures return TT_OUT_OF_MEMORY. */
} tt_status;

/* LightSampling (src/config.rs:4-12).  TT_LIGHT_SAMPLING_RESTIR is new: ReSTIR
 * DI replaces uniform next-event estimation. */
enum {
    TT_LIGHT_SAMPLING_ON_HIT = 1,
    TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION = 2,
    TT_LIGHT_SAMPLING_RESTIR = 3
};

/* Sentinel for "no texture" (src/includes/scene.glsl:76) and "no hit". */
#define TT_INVALID_INDEX 0xffffffffu

/* ---------------------------------------------------------------------- */
/* Reference byte layouts                                                  */
/* ---------------------------------------------------------------------- */

/* Constants (src/lib.rs:234-250, src/includes/constants.glsl:8-15), 368 B.
 * Matrices are column-major like glam::Mat4. */
typedef struct tt_constants {
    float view[16];
    float proj[16];
    float proj_view[16];
    float inverse_view[16];
    float inverse_proj[16];
    float camera_position[4];
    uint32_t frame_index;
    uint32_t accumulated_frame_count;
    uint32_t sample_count;
    uint32_t bounce_count;
    uint32_t screen_size[2];
    uint32_t tonemap;
    uint32_t light_sampling;
} tt_constants;

/* ProcessedVertex (src/asset/mod.rs:42-47): f16x2 texture coordinates and the
 * 32-bit tangent frame of src/asset/normal.rs:7-16. 8 B. */
typedef struct tt_vertex {
    uint16_t texture_coordinates[2];
    uint32_t tangent_frame;
} tt_vertex;

/* ProcessedMesh (src/asset/mod.rs:49-64), 32 B. */
typedef struct tt_mesh_record {
    float center[3];
    float radius;
    uint32_t vertex_offset;
    uint32_t vertex_count;
    uint32_t index_offset;
    uint32_t index_count;
} tt_mesh_record;

/* MaterialConstants (src/asset/material.rs:203-231), 192 B. */
typedef struct tt_material_constants {
    float base_color[4];
    float specular_color[4];
    float transmission_color[4];
    float coat_color[4];
    float fuzz_color[4];
    float emission_color[4];
    float geometry_normal[4];
    float base_weight;
    float base_metalness;
    float base_diffuse_roughness;
    float specular_weight;
    float specular_roughness;
    float specular_roughness_anisotropy;
    float specular_ior;
    float specular_rotation;
    float transmission_weight;
    float transmission_depth;
    float coat_weight;
    float coat_roughness;
    float coat_roughness_anisotropy;
    float coat_ior;
    float coat_darkening;
    float coat_rotation;
    float fuzz_weight;
    float fuzz_roughness;
    float emission_luminance;
    float geometry_opacity;
} tt_material_constants;

/* ProcessedMaterial (src/asset/material.rs:268-274): constants, 27 texture
 * indices (TT_INVALID_INDEX = none) and padding, 304 B. */
typedef struct tt_material {
    tt_material_constants constants;
    uint32_t textures[27];
    uint32_t padding;
} tt_material;

/* Instance (src/scene.rs:236-246, src/includes/scene.glsl:155-158), 208 B. */
typedef struct tt_instance {
    float transform[16];
    float inverse_transform[16];
    float normal_transform[16];
    uint32_t vertex_offset;
    uint32_t index_offset;
    uint32_t material_index;
    uint32_t mesh_index;
} tt_instance;

/* EmissiveTriangle (src/asset/mod.rs:66-73, src/includes/scene.glsl:11-16),
 * 36 B: snorm16 object-space positions, u16 visibility hash, f16 UVs. */
typedef struct tt_emissive_triangle {
    int16_t positions[3][3];
    uint16_t hash;
    uint16_t texture_coordinates[3][2];
    uint32_t instance_index;
} tt_emissive_triangle;

/* Everything `scene::Scene::new` uploads (src/scene.rs:28-186), already
 * flattened into instances and the emissive-triangle list. */
typedef struct tt_scene_desc {
    const int16_t* positions; /* 3 x snorm16 per vertex (BLAS stride 6 B) */
    uint32_t position_count;
    const tt_vertex* vertices;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count;
    const tt_mesh_record* meshes;
    uint32_t mesh_count;
    const tt_material* materials;
    uint32_t material_count;
    const tt_instance* instances;
    uint32_t instance_count;
    const tt_emissive_triangle* emissive_triangles;
    uint32_t emissive_triangle_count;
    /* Processed textures (scene.textures, src/asset/texture.rs), indexed by
     * tt_material.textures; none when texture_count == 0. */
    const struct tt_texture_record* textures;
    uint32_t texture_count;
    const uint8_t* texture_data;
    uint64_t texture_data_bytes;
} tt_scene_desc;

/* Texel formats of processed textures (ProcessedTexture::from_*,
 * src/asset/texture.rs:57-119).  Asset processing stores colour as BC1 read
 * through the sRGB curve (texture.rs:86-101) and normal maps as BC5 of the
 * octahedral encoding (:103-117), both block-compressed by this library's
 * own encoder, and scalars as R8 unorm or R16 float.  The sampler also reads
 * the uncompressed RGBA8 sRGB and RG8 forms.
 * Block layouts (4 x 4 texels, blocks row-major, ceil(w / 4) per row):
 *  - BC1, 8 B: RGB565 endpoints c0, c1 (little-endian), then 2-bit indices,
 *    texel (x, y) at bit 2 (4y + x).  Endpoints expand by bit replication;
 *    c0 > c1: palette c0, c1, (2 c0 + c1 + 1) / 3, (c0 + 2 c1 + 1) / 3 per
 *    8-bit channel; else c0, c1, (c0 + c1 + 1) / 2, black with alpha 0.
 *  - BC5, 16 B: two BC4 blocks, red then green.  BC4: endpoints e0, e1
 *    (bytes), then 3-bit indices, texel (x, y) at bit 3 (4y + x) of the
 *    48-bit field; e0 > e1: palette e0, e1 and ((8 - k) e0 + (k - 1) e1 + 3)
 *    / 7 for k = 2..7; else e0, e1, ((6 - k) e0 + (k - 1) e1 + 2) / 5 for
 *    k = 2..5, then 0 and 255. */
enum {
    TT_TEXTURE_FORMAT_RGBA8_SRGB = 0,
    TT_TEXTURE_FORMAT_RG8_UNORM = 1,
    TT_TEXTURE_FORMAT_R8_UNORM = 2,
    TT_TEXTURE_FORMAT_R16_FLOAT = 3,
    TT_TEXTURE_FORMAT_BC1_SRGB = 4,
    TT_TEXTURE_FORMAT_BC5_UNORM = 5
};

/* One processed texture: its mip chain (generate_mip_chain,
 * texture.rs:121-143) in texture_data; level k is (width >> k) x
 * (height >> k) texels, rows from the top (v = 0), each level at a 16-byte
 * aligned offset. */
typedef struct tt_texture_record {
    uint32_t format;
    uint32_t width, height;
    uint32_t levels;
    uint32_t offsets[16]; /* byte offset of each level in texture_data */
} tt_texture_record;

/* ---------------------------------------------------------------------- */
/* Context                                                                 */
/* ---------------------------------------------------------------------- */

typedef struct tt_context tt_context;
typedef struct tt_scene tt_scene;
typedef struct tt_target tt_target;

typedef struct tt_context_desc {
    int32_t device; /* HIP device ordinal */
    uint32_t flags; /* reserved, 0 */
} tt_context_desc;

/* Replaces Context::new (crates/tinytrace_backend/src/lib.rs:171-213). */
tt_status tt_context_create(const tt_context_desc* desc, tt_context** out);
void tt_context_destroy(tt_context* context);
/* Message of the last failed call on this context ("" if none). */
const char* tt_last_error(const tt_context* context);

/* ---------------------------------------------------------------------- */
/* Scene upload and BVH build                                              */
/* ---------------------------------------------------------------------- */

/* Replaces scene::Scene::new (src/scene.rs:28-186) with the driver's
 * BLAS/TLAS build (resource.rs:804-878, :1014-1083): uploads the arrays and
 * builds one BLAS per mesh plus a TLAS over the instances. */
tt_status tt_scene_create(tt_context* context, const tt_scene_desc* desc, tt_scene** out);

/* BVH builders: binned SAH on the host threads (best traversal speed) or
 * LBVH on the device (Karras 2012: Morton codes, radix sort, parallel
 * hierarchy, bottom-up boxes; builds in milliseconds).  The TLAS over
 * instances is built on the host either way. */
enum { TT_BVH_BUILDER_HOST_SAH = 0, TT_BVH_BUILDER_DEVICE_LBVH = 1 };
typedef struct tt_scene_options {
    uint32_t bvh_builder; /* TT_BVH_BUILDER_* */
    uint32_t reserved;
} tt_scene_options;
/* tt_scene_create with options (NULL: the defaults). */
tt_status tt_scene_create_ex(tt_context* context, const tt_scene_desc* desc, const tt_scene_options* options,
                             tt_scene** out);
void tt_scene_destroy(tt_scene* scene);

typedef struct tt_scene_info {
    uint32_t triangle_count;       /* sum over instances */
    uint32_t unique_triangle_count;/* sum over meshes */
    uint32_t node_count;           /* BLAS + TLAS nodes */
    uint32_t max_depth;            /* TLAS depth + deepest BLAS depth */
    double sah_cost;               /* expected traversal cost, SAH model */
    double expected_bytes_per_ray; /* SAH-expected node + triangle bytes */
    double build_ms;               /* host build + upload */
    uint64_t device_bytes;         /* scene bytes resident in HBM */
} tt_scene_info;
tt_status tt_scene_info_get(const tt_scene* scene, tt_scene_info* out);

/* ---------------------------------------------------------------------- */
/* Render target, render, download                                         */
/* ---------------------------------------------------------------------- */

/* Replaces create_render_target (src/lib.rs:213-226): RGBA32F. */
tt_status tt_target_create(tt_context* context, uint32_t width, uint32_t height, tt_target** out);
void tt_target_destroy(tt_target* target);

enum { TT_RNG_REFERENCE = 0, /* jenkins_hash + xorshift32, random.glsl:4-42 */
       TT_RNG_PCG = 1 /* pcg32 (PCG-XSH-RR 64/32, O'Neill 2014), SplitMix64-seeded */ };

enum { TT_RENDER_COLLECT_STATS = 1u /* count rays, time stages with events */ };

/* Settings the reference does not have (new with this library). */
typedef struct tt_render_settings {
    uint32_t rng;                /* TT_RNG_* */
    uint32_t seed;               /* mixed into the pcg32 path seed */
    uint32_t restir_candidates;  /* RIS candidates (light_sampling == RESTIR) */
    uint32_t restir_temporal;    /* 0 or 1 temporal reuse pass */
    uint32_t restir_spatial;     /* 0 or 1 spatial reuse pass */
    uint32_t flags;              /* TT_RENDER_* */
} tt_render_settings;

/* Integrate + accumulate into `target`, as `Integrator::integrate`
 * (src/integrate/mod.rs:39-61) dispatching integrate.glsl::main
 * (src/integrate/integrate.glsl:323-347).  Asynchronous: renders enqueued
 * back to back overlap on the device (a frame's last bounces with the next
 * frame's first) with the results of renders run one at a time; every other
 * call on the context is ordered after the renders enqueued before it. */
tt_status tt_render(tt_context* context, const tt_scene* scene, const tt_constants* constants,
                    tt_target* target);
tt_status tt_render_ex(tt_context* context, const tt_scene* scene, const tt_constants* constants,
                       const tt_render_settings* settings, tt_target* target);

/* Blocking download of RGBA32F (width*height*16 bytes), as
 * Renderer::render_to_texture's download (src/lib.rs:201-210). */
tt_status tt_target_download(tt_context* context, const tt_target* target, float* rgba,
                             size_t bytes);

/* Replaces Renderer::prepare_to_present's post-process dispatch
 * (src/lib.rs:180-193, src/post_process/display.glsl:6-12) for headless use:
 * the target's radiance through neutral_tonemap when constants->tonemap != 0
 * (src/post_process/tonemap.glsl:1-26), sRGB-encoded and rounded to RGBA8
 * (alpha 255) as a store to an sRGB swapchain image would, copied to `rgba8`
 * (width*height*4 bytes; blocking).  Timed as "post_process". */
tt_status tt_target_display(tt_context* context, const tt_target* target, const tt_constants* constants,
                            uint8_t* rgba8, size_t bytes);

/* Named GPU timings in milliseconds, mirroring Renderer::timings
 * (src/lib.rs:167-178) and Context::insert_timestamp
 * (crates/tinytrace_backend/src/timing.rs:61-69).  Names: "integrate" (last
 * render), "post_process" (last tt_target_display); with
 * TT_RENDER_COLLECT_STATS also "trace" (sum over the last render's extension
 * and shadow closest-hit launches, until tt_render_stats_get reads them). */
tt_status tt_timing_ms(tt_context* context, const char* name, double* out);

typedef struct tt_render_stats {
    uint64_t paths;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint32_t extend_launches;  /* closest-hit launches over extension rays */
    uint32_t shadow_launches;  /* closest-hit launches over shadow rays */
    double extend_ms;          /* sum of their durations (HIP events) */
    double shadow_ms;
    double integrate_ms;       /* whole render */
    uint64_t restir_rays;      /* ReSTIR visibility-reuse queries */
    double restir_ms;          /* ReSTIR passes (candidates, visibility, reuse) */
} tt_render_stats;
/* Statistics summed over the renders with TT_RENDER_COLLECT_STATS since the
 * last call (synchronises; the next call starts a new sum).  Renders enqueued
 * back to back overlap on the device, so the summed durations of launches
 * (and of whole renders, integrate_ms) can exceed the wall time. */
tt_status tt_render_stats_get(tt_context* context, tt_render_stats* out);

/* ---------------------------------------------------------------------- */
/* Serving loop (new with this library; the reference renders one request  */
/* at a time through Renderer::render_to_texture, src/lib.rs:201-210)      */
/* ---------------------------------------------------------------------- */

/* A server keeps scenes resident and renders many requests concurrently:
 * every step runs one batch, a shared pool of path slots filled with
 * one-sample chunks (pass k, pixel range; whole passes for ReSTIR requests)
 * of the admitted requests of one scene and integrator configuration, and
 * accumulates each chunk into its request's target in pass order.  A
 * request's image therefore depends only on the request (keyed generators:
 * pixel, pass, seed) and is bit-identical whether it runs alone or in any
 * mix.  Requests wait in arrival order until their target (and ReSTIR state)
 * fits the HBM budget next to the path pool.  Scenes with textures are not
 * served (TT_UNSUPPORTED): their level of detail needs each path's own
 * camera. */
typedef struct tt_server tt_server;

typedef struct tt_server_desc {
    uint64_t hbm_budget_bytes; /* path pool + resident targets (0: 64 GiB) */
    uint32_t pool_paths;       /* path slots per batch (0: 4 Mi) */
    uint32_t batches_in_flight; /* each with its own pool (0: as many as the
                                   context has render lanes, while their pools
                                   take at most half the budget) */
} tt_server_desc;

typedef struct tt_server_request {
    uint32_t scene;          /* slot from tt_server_add_scene */
    uint32_t rng;            /* TT_RNG_* */
    uint32_t seed;
    uint32_t passes;         /* samples per pixel, one per pass */
    tt_constants constants;  /* camera, screen size, bounce_count,
                                light_sampling (ON_HIT, NEE or RESTIR);
                                frame_index and accumulated_frame_count are
                                set per pass (k and k + 1: the result is the
                                exact mean) */
    /* ReSTIR DI (light_sampling RESTIR), as tt_render_settings: the request
     * keeps its G-buffers and reservoirs (224 B per pixel, counted in the
     * HBM budget) resident across its passes; pass k reuses pass k - 1's
     * reservoirs.  Its passes run whole, so width * height must not exceed
     * the server's pool_paths.  The image equals tt_render_ex's progressive
     * render of the same passes on a fresh target, bit for bit. */
    uint32_t restir_candidates;  /* 0: 32 */
    uint32_t restir_temporal;
    uint32_t restir_spatial;
    uint32_t reserved;
} tt_server_request;

enum { TT_REQUEST_QUEUED = 0, TT_REQUEST_ADMITTED = 1, TT_REQUEST_DONE = 2 };

typedef struct tt_request_status {
    uint32_t state;          /* TT_REQUEST_* */
    uint32_t passes_done;
    uint64_t submitted_ns;   /* steady clock, ns since the server was created */
    uint64_t admitted_ns;
    uint64_t done_ns;        /* image downloaded to host memory */
} tt_request_status;

typedef struct tt_server_stats {
    uint64_t batches;
    uint64_t paths;          /* path samples rendered */
    uint64_t queued, active, done;
    uint64_t budget_used_bytes;
    double busy_ms;          /* wall time with at least one batch in flight */
    /* Batches on scenes with a radiance cache (tt_server_set_cache): paths
     * ended at a cache query, training samples trained on, judge paths and
     * their mean traced radiance and the cache's mean prediction at their
     * cache vertices (per channel), as tt_cache_stats, summed over batches. */
    uint64_t cache_queries;
    uint64_t cache_samples;
    uint64_t judge_samples;
    double judge_traced;
    double judge_cached;
} tt_server_stats;

tt_status tt_server_create(tt_context* context, const tt_server_desc* desc, tt_server** out);
void tt_server_destroy(tt_server* server);
/* Keeps `scene` resident for requests (the scene must outlive the server). */
tt_status tt_server_add_scene(tt_server* server, const tt_scene* scene, uint32_t* slot);
/* TT_OVER_BUDGET when the request's target can never fit next to the pool. */
tt_status tt_server_submit(tt_server* server, const tt_server_request* request, uint64_t* id);
/* Admits what fits and enqueues one batch; once batches_in_flight batches
 * are in flight (or nothing is left to enqueue) waits for the oldest and
 * downloads the requests it completed; *completed = their number (0 and no
 * work when idle). */
tt_status tt_server_step(tt_server* server, uint32_t* completed);
tt_status tt_server_status(tt_server* server, uint64_t id, tt_request_status* out);
/* Copies a done request's RGBA32F image (width*height*16 bytes) and forgets
 * the request. */
tt_status tt_server_take_result(tt_server* server, uint64_t id, float* rgba, size_t bytes);
tt_status tt_server_stats_get(tt_server* server, tt_server_stats* out);

/* ---------------------------------------------------------------------- */
/* Radiance cache (new with this library): a 96-5x64-3 MLP trained online   */
/* ---------------------------------------------------------------------- */

/* Network: 96 encoded features (12 sine/cosine frequencies of the position
 * per axis, 2 of the direction, normal, base colour, roughness, metalness,
 * bounce) -> 5 x (64, ReLU) -> 3 (softplus) = outgoing radiance.  22720 fp32
 * master parameters (no biases), bf16 copies on the matrix cores, Adam.  A
 * cached render ends each path at its cache vertex, its first vertex after
 * the primary one whose specular roughness exceeds 0.3, and adds
 * throughput x cache there.  Training paths run on plainly and teach the
 * cache the radiance their tails carry; judge paths run on plainly and
 * measure the cache's error. */
typedef struct tt_mlp tt_mlp;

typedef struct tt_mlp_desc {
    uint64_t seed;        /* weight initialisation (He-uniform) */
    float learning_rate;  /* Adam, 0: 1e-3 (beta1 0.9, beta2 0.99, eps 1e-8) */
    uint32_t reserved;
} tt_mlp_desc;

tt_status tt_mlp_create(tt_context* context, const tt_mlp_desc* desc, tt_mlp** out);
void tt_mlp_destroy(tt_mlp* mlp);
/* World box that positions are normalised to. */
tt_status tt_mlp_set_bounds(tt_mlp* mlp, const float lo[3], const float hi[3]);
/* Read (`get`) and/or replace (`set`) the 22720 fp32 parameters. */
tt_status tt_mlp_params(tt_mlp* mlp, float* get, const float* set, uint32_t count);
/* Host-array operations on n raw records (16 floats each: position,
 * direction, normal, base colour, roughness, metalness, bounce, unused):
 *   op 0: encode -> out[n x 96];  op 1: infer -> out[n x 3];
 *   op 2: loss gradient for targets[n x 4] -> out[22720], *loss;
 *   op 3: one Adam step on targets -> *loss (before the step). */
tt_status tt_mlp_run(tt_mlp* mlp, int op, const float* records, const float* targets, uint32_t n, float* out,
                     double* loss);

typedef struct tt_cache_settings {
    /* Training samples per frame (0: 65536).  A training path records its
     * cache vertex and up to records_per_path - 1 later rough vertices, so
     * about train_samples / records_per_path paths train.  Samples from later
     * vertices carry fewer remaining bounces than a query needs, and with
     * more than one record per path the cache learns slower and settles
     * lower (measured on the Cornell box: 2% low with one, 5% with two, 16%
     * with three after 28 frames). */
    uint32_t train_samples;
    uint32_t train_steps;  /* Adam steps per frame over them (0: 4; at most 64) */
    /* Share of the paths reaching a cache vertex that are traced plainly as
     * the error judge (0: 0.01; negative: none). */
    float judge_fraction;
    uint32_t records_per_path;  /* training samples per training path (0: 1; 1 to 3) */
    /* Nonzero: training takes the samples in path order instead of the order
     * shading finished in, so the same frames with the same seeds give the
     * same weights and images bit for bit; costs a scatter, a scan over
     * paths x records_per_path entries and a gather per training (0.1 ms a
     * frame on the 512 x 512 Cornell box). */
    uint32_t ordered_training;
} tt_cache_settings;

typedef struct tt_cache_stats {
    uint32_t queries;        /* paths ended at a cache query */
    uint32_t samples;        /* training samples recorded and trained on */
    double loss;             /* mean training loss of the frame's first Adam step */
    uint32_t judge_samples;  /* judge paths, each traced plainly from its cache vertex */
    uint32_t reserved;
    double judge_traced;     /* their mean traced radiance from that vertex (per channel) */
    double judge_cached;     /* the cache's mean prediction there, with the weights the frame's queries used */
    double judge_loss;       /* mean (y - t)^2 / (y^2 + 0.01) of prediction y against traced t */
    double infer_ms;         /* the frame's query inference (HIP events on its stream) */
    double train_ms;         /* the frame's Adam steps (HIP events on the API stream) */
} tt_cache_stats;

/* tt_render_ex with the radiance cache: renders (any light sampling; with
 * ReSTIR the primary vertex uses the reservoirs and the cache ends paths
 * later), queries the cache for the paths it ends and for the judge
 * records, then trains it on the frame's training samples.  Cached renders
 * overlap on the render lanes like others; a frame's queries wait for the
 * previous frame's training.  With `stats` the call waits for the frame and
 * its training and fills them; without, it returns once enqueued. */
tt_status tt_render_cached(tt_context* context, const tt_scene* scene, const tt_constants* constants,
                           const tt_render_settings* settings, tt_mlp* cache, const tt_cache_settings* cache_settings,
                           tt_target* target, tt_cache_stats* stats);

/* Serves the requests on `scene_slot` with the radiance cache, as
 * tt_render_cached renders: every batch of them ends its paths at cache
 * queries, infers them with the weights of the last training enqueued,
 * then trains the cache on its training paths (`settings` per batch as per
 * frame; null: the defaults); judge statistics go to tt_server_stats.  A
 * null cache serves the slot plainly again.  The cache must belong to the
 * server's context and stay alive while requests on the slot are served.
 * A cached request's image depends on what the cache learnt from the batches
 * before, so it is no longer the same bits alone and in a mix. */
tt_status tt_server_set_cache(tt_server* server, uint32_t scene_slot, tt_mlp* cache,
                              const tt_cache_settings* settings);

/* ---------------------------------------------------------------------- */
/* Test entry points                                                       */
/* ---------------------------------------------------------------------- */

/* One ray: world-space origin/direction and the accepted t interval. */
typedef struct tt_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} tt_ray;

/* Closest hit: instance = TT_INVALID_INDEX on a miss. */
typedef struct tt_hit {
    float t;
    uint32_t instance;  /* rayQueryGetIntersectionInstanceCustomIndexEXT */
    uint32_t primitive; /* rayQueryGetIntersectionPrimitiveIndexEXT */
    float barycentrics[2];
} tt_hit;

enum { TT_TRACE_CULL_BACK_FACES = 1u };

/* Closest-hit traversal of host rays, standing in for the rayQueryEXT loops
 * of integrate.glsl:112-127 (extension rays, back faces culled) and
 * :168-185 (shadow rays, no culling).  Synchronous. */
tt_status tt_trace_closest(tt_context* context, const tt_scene* scene, const tt_ray* rays,
                           tt_hit* hits, uint32_t count, uint32_t flags);
/* Same on device-resident arrays (no copies, asynchronous). */
tt_status tt_trace_closest_device(tt_context* context, const tt_scene* scene,
                                  const tt_ray* device_rays, tt_hit* device_hits, uint32_t count,
                                  uint32_t flags);

/* No counterpart in the reference.  tt_trace_closest for callers that cannot
 * allocate device memory (a serving loop or harness that forbids hipMalloc
 * after start-up): same rays, same hits, bit for bit, but the call allocates
 * and frees nothing.  Rays and hits are staged through a fixed scratch array
 * that the library holds on each device for the life of the process (about
 * 10.5 MiB: the ray cursor, staging for TT_TRACE_STATIC_MAX_RAYS rays and
 * hits, and the traversal stack spill of a 128-workgroup grid).  count above
 * TT_TRACE_STATIC_MAX_RAYS returns TT_INVALID_ARGUMENT.  Every call on a
 * device shares that device's scratch: calls that overlap on one device
 * (from several contexts or threads) race and must be serialised by the
 * caller.  Synchronous. */
#define TT_TRACE_STATIC_MAX_RAYS 65536u
tt_status tt_trace_closest_static(tt_context* context, const tt_scene* scene, const tt_ray* rays,
                                  tt_hit* hits, uint32_t count, uint32_t flags);

typedef struct tt_bsdf_query {
    float wi[3];
    float wo[3];
} tt_bsdf_query;

typedef struct tt_bsdf_result {
    float bsdf[3];
    float density;
} tt_bsdf_result;

/* bsdf_prepare + bsdf_evaluate for one material at many (wi, wo) pairs, as
 * the BdsfSampler kernel (src/material/test.rs:38-83): every query starts
 * from generator state `rng_state`. */
tt_status tt_bsdf_evaluate(tt_context* context, const tt_material_constants* material,
                           const tt_bsdf_query* queries, tt_bsdf_result* results, uint32_t count,
                           uint32_t rng_state);

/* The device's scalar math (op: 0 sin, 1 cos, 2 exp, 3 log, 4 acos, 5 asin,
 * 6 pow(x, y), 7 sqrt, 8 x / y, 9 1 / x) on host arrays, so tests can check
 * that the kernels' arithmetic is bit-identical to the host's.  No reference
 * counterpart (the reference uses GLSL built-ins). */
tt_status tt_debug_math(tt_context* context, int op, const float* x, const float* y, float* out,
                        uint32_t count);

/* ---------------------------------------------------------------------- */
/* Host side of the library (camera, asset processing)                     */
/* ---------------------------------------------------------------------- */

/* Camera (src/camera.rs:6-26). */
typedef struct tt_camera {
    float position[3];
    float yaw;
    float pitch;
    float z_near;
    float z_far;
    float fov;
} tt_camera;

/* Config (src/config.rs:23-43). */
typedef struct tt_config {
    uint32_t bounce_count;
    uint32_t sample_count;
    uint32_t tonemap;
    uint32_t light_sampling;
} tt_config;

void tt_camera_default(tt_camera* out);
void tt_config_default(tt_config* out);

/* Constants as Renderer::render_to_target builds them (src/lib.rs:125-142):
 * view = look_at_rh, proj = perspective_infinite_reverse_rh with
 * z_axis[3] = 1 (src/camera.rs:42-50), inverses by cofactor expansion. */
tt_status tt_constants_build(const tt_camera* camera, const tt_config* config, uint32_t width,
                             uint32_t height, uint32_t frame_index,
                             uint32_t accumulated_frame_count, tt_constants* out);

/* Asset processing: asset::Scene (src/asset/mod.rs:76-205). */
typedef struct tt_asset_scene tt_asset_scene;

/* asset::Mesh (src/asset/mod.rs:20-27). */
typedef struct tt_mesh_desc {
    const float* positions;           /* 3 per vertex */
    const float* texture_coordinates; /* 2 per vertex */
    const uint32_t* indices;          /* optional */
    const float* normals;             /* optional, 3 per vertex */
    const float* tangents;            /* optional, 4 per vertex (generated when absent) */
    uint32_t vertex_count;
    uint32_t index_count;
} tt_mesh_desc;

/* asset::Material with constant parameters (src/asset/material.rs:8-175). */
typedef struct tt_material_desc {
    float base_weight;
    float base_color[3];
    float base_metalness;
    float base_diffuse_roughness;
    float specular_weight;
    float specular_color[3];
    float specular_roughness;
    float specular_roughness_anisotropy;
    float specular_ior;
    float specular_rotation;
    float transmission_weight;
    float transmission_color[3];
    float transmission_depth;
    float coat_weight;
    float coat_color[3];
    float coat_roughness;
    float coat_roughness_anisotropy;
    float coat_ior;
    float coat_darkening;
    float coat_rotation;
    float fuzz_weight;
    float fuzz_color[3];
    float fuzz_roughness;
    float emission_luminance;
    float emission_color[3];
    float geometry_opacity;
    float geometry_normal[3];
} tt_material_desc;

/* Defaults of asset::Material (src/asset/material.rs:35-175). */
void tt_material_desc_default(tt_material_desc* out);

/* Texture channels of a material: the fields of MaterialTextures in order
 * (src/asset/material.rs:203-231), i.e. the order of tt_material.textures. */
enum {
    TT_TEXTURE_BASE_COLOR = 0,
    TT_TEXTURE_SPECULAR_COLOR = 1,
    TT_TEXTURE_TRANSMISSION_COLOR = 2,
    TT_TEXTURE_COAT_COLOR = 3,
    TT_TEXTURE_FUZZ_COLOR = 4,
    TT_TEXTURE_EMISSION_COLOR = 5,
    TT_TEXTURE_GEOMETRY_NORMAL = 6,
    TT_TEXTURE_BASE_WEIGHT = 7,
    TT_TEXTURE_BASE_METALNESS = 8,
    TT_TEXTURE_BASE_DIFFUSE_ROUGHNESS = 9,
    TT_TEXTURE_SPECULAR_WEIGHT = 10,
    TT_TEXTURE_SPECULAR_ROUGHNESS = 11,
    TT_TEXTURE_SPECULAR_ROUGHNESS_ANISOTROPY = 12,
    TT_TEXTURE_SPECULAR_IOR = 13,
    TT_TEXTURE_SPECULAR_ROTATION = 14,
    TT_TEXTURE_TRANSMISSION_WEIGHT = 15,
    TT_TEXTURE_TRANSMISSION_DEPTH = 16,
    TT_TEXTURE_COAT_WEIGHT = 17,
    TT_TEXTURE_COAT_ROUGHNESS = 18,
    TT_TEXTURE_COAT_ROUGHNESS_ANISOTROPY = 19,
    TT_TEXTURE_COAT_IOR = 20,
    TT_TEXTURE_COAT_DARKENING = 21,
    TT_TEXTURE_COAT_ROTATION = 22,
    TT_TEXTURE_FUZZ_WEIGHT = 23,
    TT_TEXTURE_FUZZ_ROUGHNESS = 24,
    TT_TEXTURE_EMISSION_LUMINANCE = 25,
    TT_TEXTURE_GEOMETRY_OPACITY = 26,
    TT_TEXTURE_CHANNELS = 27
};

/* A material parameter given as a texture (Param::Texture): width x height
 * texels, rows from the top (v = 0); 3 floats per texel for the colour and
 * normal channels (0..6), 1 float otherwise. */
typedef struct tt_texture_desc {
    uint32_t channel; /* TT_TEXTURE_* */
    uint32_t width, height;
    const float* texels;
} tt_texture_desc;

tt_status tt_asset_scene_create(tt_asset_scene** out);
void tt_asset_scene_destroy(tt_asset_scene* scene);
/* Textures inserted from now on (materials and glTF imports): colour and
 * normal maps block-compressed as BC1 / BC5 (enable 1, the default, as the
 * reference's asset processing stores them) or uncompressed as RGBA8 / RG8
 * (enable 0: the texels' exact quantized bytes). */
tt_status tt_asset_set_block_compression(tt_asset_scene* scene, uint32_t enable);
/* Scene::insert_mesh (src/asset/mod.rs:89-161). */
tt_status tt_asset_insert_mesh(tt_asset_scene* scene, const tt_mesh_desc* mesh, uint32_t* index);
/* Scene::insert_material (src/asset/material.rs:305-357). */
tt_status tt_asset_insert_material(tt_asset_scene* scene, const tt_material_desc* material,
                                   uint32_t* index);
/* The same with textured parameters: each texture is processed into its mip
 * chain and format and replaces its channel's constant, which becomes the
 * channel type's default (zero) as Param::constant_or_default gives it. */
tt_status tt_asset_insert_material_textured(tt_asset_scene* scene, const tt_material_desc* material,
                                            const tt_texture_desc* textures, uint32_t texture_count,
                                            uint32_t* index);
/* Node tree (src/asset/mod.rs:35-40).  Node 0 is the root; `transform` is
 * column-major. */
tt_status tt_asset_add_node(tt_asset_scene* scene, uint32_t parent, const float transform[16],
                            uint32_t* index);
tt_status tt_asset_add_model(tt_asset_scene* scene, uint32_t node, uint32_t mesh,
                             uint32_t material);
/* Replaces GltfImporter::new(path) + insert_in_scene (src/asset/gltf.rs:50-72,
 * 419-452): appends the default scene (else the first) of a .gltf file with
 * external or base64 data-URI buffers under the root node, as the reference
 * does (meshes per primitive, materials per index with their PNG textures,
 * node tree depth first).  TT_UNSUPPORTED for binary .glb, images other than
 * PNG, non-triangle primitives and accessor formats the reference rejects;
 * TT_INVALID_ARGUMENT for unreadable files, corrupt images and malformed
 * documents.  On failure the scene may hold part of the import. */
tt_status tt_asset_import_gltf(tt_asset_scene* scene, const char* path);

/* Depth-first flatten into instances and emissive triangles, as
 * scene::Scene::new does (src/scene.rs:32-64).  The arrays stay valid until
 * the asset scene is modified or destroyed. */
tt_status tt_asset_flatten(tt_asset_scene* scene, tt_scene_desc* out);

#ifdef __cplusplus
}
#endif

#endif /* TINYTRACE_H */
