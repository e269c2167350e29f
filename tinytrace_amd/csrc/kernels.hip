// gfx950 kernels of the wavefront path integrator.
//
// The reference runs one megakernel thread per pixel
// (src/integrate/integrate.glsl:323-347).  Here one render is a sequence of
// launches on one stream, per bounce b:
//   extend  : closest hit of every ray in queue[b % 2] (persistent waves
//             fetch 64 rays at a time from partitioned atomic cursors);
//             primary rays (b = 0) of a whole frame sit in 8 x 8 pixel-tile
//             order (tt::primary_slot)
//   shade   : one lane per hit: next_path_segment (integrate.glsl:270-321)
//             up to the two ray queries, appending the continuation ray to
//             queue[(b+1) % 2] and the shadow ray to the shadow queue, each
//             workgroup's rays binned by direction octant (one atomic per
//             workgroup and queue)
//   shadow  : closest hit of every shadow ray, the reference's visibility
//             rule (integrate.glsl:175-185) and the radiance update
// framed by generate (camera rays, integrate.glsl:323-333) and accumulate
// (sample average + progressive accumulation, integrate.glsl:338-346).
// With ReSTIR DI, the passes of restir.hip run between extend(0) and
// shade(0), and shade(0) takes its light sample from the pixel's reservoir.
//
// Path state is SoA in HBM: throughput/radiance float4, generator u32,
// queues float4 (origin.xyz + path index bits, direction.xyz); every access
// is one 16-B load or store per lane, consecutive across the wave.
#include "bvh.h"
#include "device_surface.h"

using namespace ttd;

static_assert(TRACE_STACK_LIMIT == int(tt::BVH_WIDE_STACK_LIMIT), "traversal stack covers the deepest wide path");
static_assert(size_t(TRACE_BLOCK) == tt::TRACE_THREADS && size_t(TRACE_SPILL_STACK) == tt::TRACE_SPILL_ENTRIES,
              "host-side traversal constants match the device's");

namespace {

using tt::C_EXT;
using tt::C_SHADOW;

// ------------------------------------------------------------ test rays --
template <bool CULL>
__global__ __launch_bounds__(TRACE_BLOCK) void trace_rays_kernel(tt::DeviceScene d, const tt_ray* rays, tt_hit* hits,
                                                                  uint32_t count, uint32_t* work, uint32_t* spill) {
    SceneView s = view(d);
    Stack st = make_stack(spill);
    WorkFetch wf = work_fetch(work, count);
    uint32_t first, batch;
    while (next_batch(wf, first, batch)) {
        uint32_t i = first + lane_id();
        if (lane_id() < batch) {
            tt_ray r = rays[i];
            HitRecord h;
            bool found = trace_closest<CULL>(s, f3(r.origin[0], r.origin[1], r.origin[2]),
                                             f3(r.direction[0], r.direction[1], r.direction[2]), r.tmin, r.tmax, st, h);
            tt_hit o;
            o.t = found ? h.t : r.tmax;
            o.instance = found ? h.instance : 0xffffffffu;
            o.primitive = found ? h.primitive : 0xffffffffu;
            o.barycentrics[0] = h.b1;
            o.barycentrics[1] = h.b2;
            hits[i] = o;
        }
    }
}

// -------------------------------------------------------------- generate --
// integrate.glsl:129-132 (pixel_ndc), :17-20 (camera_ray_direction),
// :323-333 (seed, per-sample generator copy).  Path p = pixel * spp + sample.
// Path p of pixel `pixel`, sample `sample` of this pass, its ray in queue
// slot `slot`.
TT_DEV void generate_path(const tt::FrameArgs& f, const tt::PathBuffers& b, uint32_t p, uint32_t slot, uint32_t pixel,
                          uint32_t sample) {
    uint32_t x = pixel % f.width, y = pixel / f.width;
    Generator g = init_generator(pixel, f.frame_index, f.rng, f.seed);
    for (uint32_t k = 0; k < 2u * sample; k++) next_u32(g);
    float ox = random_float(g);
    float oy = random_float(g);
    ox = ox - 0.5f;
    oy = oy - 0.5f;
    float ndc_x = ((float(x) + ox) / float(f.width)) * 2.0f - 1.0f;
    float ndc_y = ((float(y) + oy) / float(f.height)) * 2.0f - 1.0f;
    f3 dir = camera_ray_direction(f.inverse_view, f.inverse_proj, ndc_x, ndc_y);
    // The sample's device coordinates ride in the spare w components for the
    // first hit's texture level of detail: x with the direction, y with the
    // throughput (whose w then holds the level of detail).
    b.queue[0][0][slot] = make_float4(f.camera_position[0], f.camera_position[1], f.camera_position[2], __uint_as_float(p));
    b.queue[0][1][slot] = make_float4(dir.x, dir.y, dir.z, ndc_x);
    b.throughput[p] = make_float4(1.0f, 1.0f, 1.0f, ndc_y);
    b.radiance[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.rng[p] = g.state;
    if (b.cache.open) b.cache.open[p] = make_uint4(tt::CACHE_NO_RECORD, tt::CACHE_NO_RECORD, tt::CACHE_NO_RECORD, 0u);
}

__global__ __launch_bounds__(256) void generate_kernel(tt::FrameArgs f, tt::PathBuffers b, uint32_t paths) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= paths) return;
    const uint32_t pixel = p / f.sample_count, sample = p % f.sample_count;
    uint32_t slot = p;
    if (f.primary_tiles)
        slot = tt::primary_slot(pixel % f.width, pixel / f.width, f.width, f.height) * f.sample_count + sample;
    generate_path(f, b, p, slot, pixel, sample);
    if (p == 0) b.counters[C_EXT + 0] = paths;
}

// Serving: one sample of pixels [pixel0, pixel0 + count) into path slots
// [base, base + count); the batch's path count accumulates in C_EXT[0].
__global__ __launch_bounds__(256) void generate_chunk_kernel(tt::FrameArgs f, tt::PathBuffers b, uint32_t base,
                                                             uint32_t pixel0, uint32_t count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    generate_path(f, b, base + i, base + i, pixel0 + i, 0u);
    if (i == 0) atomicAdd(b.counters + C_EXT + 0, count);
}

// ---------------------------------------------------------------- extend --
// Five waves per SIMD (at most 102 VGPRs, no spills): with the LDS stack that
// is the most a CU holds, and the persistent grid is sized to fill it.
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void extend_kernel(tt::DeviceScene d, tt::PathBuffers b, uint32_t bounce) {
    SceneView s = view(d);
    Stack st = make_stack(b.spill);
    const uint32_t count = b.counters[C_EXT + bounce];
    const float4* qo = b.queue[bounce & 1][0];
    const float4* qd = b.queue[bounce & 1][1];
    WorkFetch wf = work_fetch(b.work + (tt::W_EXT + bounce) * tt::WORK_SET, count);
    uint32_t first, batch;
    while (next_batch(wf, first, batch)) {
        uint32_t i = first + lane_id();
        if (lane_id() < batch) {
            float4 o = qo[i], dd = qd[i];
            HitRecord h;
            bool found = trace_closest<true>(s, f3(o.x, o.y, o.z), f3(dd.x, dd.y, dd.z), 1.0e-3f, 1000.0f, st, h);
            b.hits[i] = make_float4(h.b1, h.b2, __uint_as_float(found ? h.instance : 0xffffffffu),
                                    __uint_as_float(h.primitive));
        }
    }
}

// ----------------------------------------------------------------- shade --
// Radiance increment a seen from a vertex with throughput t (0 where t is 0).
TT_DEV f3 cache_ratio(f3 a, f3 t) {
    return f3(t.x > 0.0f ? a.x / t.x : 0.0f, t.y > 0.0f ? a.y / t.y : 0.0f, t.z > 0.0f ? a.z / t.z : 0.0f);
}

// open[p].w flag of a judge path (one record, at its cache vertex).
constexpr uint32_t CACHE_JUDGE_PATH = 8u;

// Adds radiance `a` (as the path's throughput scaled it) to the targets of
// the path's open records, each divided by the throughput at its vertex.
TT_DEV void add_to_records(const tt::CacheBuffers& cb, uint4 open, f3 a) {
    const uint32_t n = open.w & 3u;
    for (uint32_t j = 0; j < n; j++) {
        uint32_t slot = j == 0 ? open.x : (j == 1 ? open.y : open.z);
        float4 T = cb.train_thr[slot];
        float4 t = cb.train_target[slot];
        f3 d = cache_ratio(a, f3(T.x, T.y, T.z));
        cb.train_target[slot] = make_float4(t.x + d.x, t.y + d.y, t.z + d.z, t.w);  // w: the record's order key
    }
}

// TEXTURED: the instantiation for scenes with textures (texture fetches at
// hits and on textured light samples); untextured scenes run the other,
// which carries no texture code.  The untextured one is held to four waves
// per SIMD (128 VGPRs, a 32-B spill), which beats three waves without spills;
// the textured one would spill 160 B and keeps the compiler's choice.
template <bool TEXTURED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TEXTURED ? 1 : 4))) void shade_kernel(tt::DeviceScene d, tt::FrameArgs f, tt::PathBuffers b,
                                                    uint32_t bounce) {
    SceneView s = view(d);
    const uint32_t count = b.counters[C_EXT + bounce];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= count) return;  // the whole workgroup
    // The continuation ray and the shadow ray, appended after the workgroup
    // has binned them.
    bool cont_out = false;
    f3 cont_org(0.0f), cont_dir(0.0f);
    uint32_t cont_path = 0;
    bool shw_out = false;
    float4 shw0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), shw1 = shw0, shw2 = shw0, shw3 = shw0;
    if (i < count) {
        const bool last = bounce + 1 >= f.bounce_count;
        const bool restir = f.light_sampling == TT_LIGHT_SAMPLING_RESTIR && bounce == 0;
        const bool nee_mode = f.light_sampling == TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION ||
            (f.light_sampling == TT_LIGHT_SAMPLING_RESTIR && bounce > 0);

        float4 qd = b.queue[bounce & 1][1][i];
        float4 qo = b.queue[bounce & 1][0][i];
        float4 hit = b.hits[i];
        uint32_t p = __float_as_uint(qo.w);
        f3 ray_dir(qd.x, qd.y, qd.z);
        float4 thr4 = b.throughput[p];
        float4 acc4 = b.radiance[p];
        // Texture level of detail: the path's (create_path_state starts it at
        // 1), set at the first hit from ray differentials.
        float lod = bounce == 0 ? 1.0f : thr4.w;
        f3 throughput(thr4.x, thr4.y, thr4.z);
        f3 accumulated(acc4.x, acc4.y, acc4.z);
        Generator g{b.rng[p], f.rng};
        uint32_t instance = __float_as_uint(hit.z);

        // Radiance cache: the path's open training / judge records, the radiance
        // this vertex adds (for their targets) and a query or record here.
        const bool cache = f.cache_mode != 0u;
        const f3 thr_in = throughput;
        uint4 open = cache ? b.cache.open[p] : make_uint4(0u, 0u, 0u, 0u);
        bool emit_query = false, emit_train = false, emit_judge = false;
        f3 added;
        float rec[16];

        bool emit_shadow = false, emit_ray = false;
        Mat m;
        Surface sf;
        f3 shadow_org, shadow_dir, direct, emission_after;
        uint32_t light_instance = 0, light_hash = 0;
        bool has_emission_after = false;
        f3 next_dir, next_org;

        if (instance == 0xffffffffu) {
            // Miss: constant white sky (integrate.glsl:272-276).
            added = f3(1.0f) * throughput;
            accumulated = accumulated + added;
        } else {
            if constexpr (TEXTURED) {
                TextureHit tex{true, bounce == 0,
                               PrimaryRay{f.inverse_view, f.inverse_proj, f3(qo.x, qo.y, qo.z), float(f.width),
                                          float(f.height), qd.w, thr4.w},
                               lod};
                sf = surface_at_hit(s, instance, __float_as_uint(hit.w), hit.x, hit.y, ray_dir, m, tex);
                lod = tex.lod;
            } else {
                sf = surface_at_hit(s, instance, __float_as_uint(hit.w), hit.x, hit.y, ray_dir, m);
            }
            if (cache && bounce >= 1 && m.specular_roughness > 0.3f) {
                if (!(open.w & tt::CACHE_DECIDED)) {
                    // The cache vertex: a query ends the path here; a training or
                    // judge path records the vertex and runs on.
                    uint32_t h = pcg32_output(splitmix64(pcg32_key(p, f.frame_index) ^ 0x6a09e667f3bcc909ull));
                    emit_train = h < f.train_threshold;
                    emit_judge = !emit_train && h < f.judge_threshold;
                    emit_query = !emit_train && !emit_judge;
                } else if (!(open.w & CACHE_JUDGE_PATH) && (open.w & 3u) < f.cache_records) {
                    // A later rough vertex of a training path: one more sample.
                    emit_train = true;
                }
                if (emit_train || emit_judge || emit_query) {
                    f3 v[4] = {sf.position, -ray_dir, sf.basis.normal, m.base_color};
                    for (int k = 0; k < 4; k++) {
                        rec[3 * k + 0] = v[k].x;
                        rec[3 * k + 1] = v[k].y;
                        rec[3 * k + 2] = v[k].z;
                    }
                    rec[12] = m.specular_roughness;
                    rec[13] = m.base_metalness;
                    rec[14] = float(bounce);
                    rec[15] = 0.0f;
                }
            }
        }
        if (instance != 0xffffffffu && !emit_query) {
            f3 wi = sf.wi;
            Lobes lobes;
            if (restir) {
                uint2 key = b.restir_key ? b.restir_key[p] : make_uint2(p, f.frame_index);
                Generator lg = restir_lobe_generator(key.x, key.y);
                lobes = bsdf_prepare(m, wi, lg);
            } else {
                lobes = bsdf_prepare(m, wi, g);
            }

            bool gate = s.emissive_count > 0 && m.specular_roughness > 0.15f;
            bool nee = (nee_mode || restir) && gate;
            f3 emissive = m.emission_color * m.emission_luminance;
            bool sampled = false, known_visible = false;
            LightPoint lp;
            f3 contribution;
            if (nee && nee_mode) {
                // sample_random_light (scene.glsl:114-136) and the unshadowed
                // part of direct_light_contribution (integrate.glsl:156-193).
                uint32_t li = random_uint(g, s.emissive_count);
                f2 r2 = random_vec2(g);
                if constexpr (TEXTURED) lp = light_point_textured(s, li, sample_triangle(r2));
                else lp = light_point(s, li, sample_triangle(r2));
                float light_probability = 1.0f / float(s.emissive_count);
                f3 to_light = lp.position - sf.position;
                f3 wo = transform_from_basis(sf.basis, to_light);
                float distance_squared = length_squared(to_light);
                float distance = sqrt_(distance_squared);
                f3 dir = to_light / distance;
                float density = (1.0f / lp.area) * distance_squared / saturate(dot(-dir, lp.normal)) * light_probability;
                float bsdf_density;
                f3 bsdf = bsdf_evaluate(m, lobes, wi, wo, bsdf_density);
                contribution = bsdf * lp.emission * abs_(cos_theta(wo)) / fmax_(density, PDF_EPSILON);
                shadow_dir = dir;
                sampled = true;
            } else if (nee && restir) {
                // ReSTIR DI: the reservoir's sample y and contribution weight W
                // estimate the direct light as f(y) Le(y) G(y) W.
                Reservoir r = unpack_reservoir(b.restir_final[p]);
                if (r.light != 0xffffffffu && r.W > 0.0f) {
                    if constexpr (TEXTURED) lp = light_point_textured(s, r.light, reservoir_bary(r.bary));
                    else lp = light_point(s, r.light, reservoir_bary(r.bary));
                    f3 to_light = lp.position - sf.position;
                    f3 wo = transform_from_basis(sf.basis, to_light);
                    float distance_squared = length_squared(to_light);
                    float distance = sqrt_(distance_squared);
                    f3 dir = to_light / distance;
                    float cos_light = saturate(dot(-dir, lp.normal));
                    float bsdf_density;
                    f3 bsdf = bsdf_evaluate(m, lobes, wi, wo, bsdf_density);
                    float scale = ((abs_(cos_theta(wo)) * cos_light) / distance_squared) * r.W;
                    contribution = (bsdf * lp.emission) * scale;
                    shadow_dir = dir;
                    sampled = true;
                    // The pixel's own candidate that passed the visibility pass
                    // is the same ray from the same point: its answer stands.
                    uint4 own = b.restir_initial[p];
                    known_visible = own.x == r.light && own.y == r.bary && (own.z & 0xffffu) != 0u;
                }
            }
            if (sampled && known_visible) {
                // As the shadow kernel would: direct light, then the deferred
                // emission of the primary vertex.
                direct = throughput * contribution;
                accumulated = accumulated + direct;
                accumulated = accumulated + emissive * throughput;
            } else if (sampled) {
                direct = throughput * contribution;
                if (direct.x == 0.0f && direct.y == 0.0f && direct.z == 0.0f) {
                    // Visibility only scales the contribution, so a zero one (the
                    // light's back faces the point, or the direction lies below
                    // the surface) needs no shadow ray: the result is the
                    // occluded one, and at the primary vertex that adds the
                    // deferred emission, as the shadow kernel would.
                    if (bounce == 0) accumulated = accumulated + emissive * throughput;
                } else {
                    shadow_org = sf.position;
                    light_instance = lp.instance;
                    light_hash = lp.hash;
                    emit_shadow = true;
                    if (bounce == 0) {
                        emission_after = emissive * throughput;
                        has_emission_after = true;
                    }
                }
            } else if (bounce == 0 || !nee) {
                added = emissive * throughput;
                accumulated = accumulated + added;
            }
            if (!last) {
                f3 wo, value;
                float density;
                if (bsdf_sample(m, lobes, wi, g, wo, density, value)) {
                    next_dir = transform_to_basis(sf.basis, wo);
                    next_org = sf.position;
                    throughput = throughput * (value / fmax_(density, PDF_EPSILON) * abs_(cos_theta(wo)));
                    emit_ray = true;
                }
            }
        }

        b.radiance[p] = make_float4(accumulated.x, accumulated.y, accumulated.z, 0.0f);
        b.throughput[p] = make_float4(throughput.x, throughput.y, throughput.z, lod);
        b.rng[p] = g.state;

        cont_out = emit_ray;
        cont_org = next_org;
        cont_dir = next_dir;
        cont_path = p;
        if (emit_shadow) {
            shw_out = true;
            shw0 = make_float4(shadow_org.x, shadow_org.y, shadow_org.z, __uint_as_float(p));
            shw1 = make_float4(shadow_dir.x, shadow_dir.y, shadow_dir.z, __uint_as_float(light_instance));
            shw2 = make_float4(direct.x, direct.y, direct.z,
                               __uint_as_float(light_hash | (has_emission_after ? 0x10000u : 0u)));
            shw3 = make_float4(emission_after.x, emission_after.y, emission_after.z, 0.0f);
        }
        if (cache) {
            const tt::CacheBuffers& cb = b.cache;
            uint32_t qs = wave_append(b.counters + tt::C_CACHE_QUERY, emit_query);
            if (emit_query) {
                for (int k = 0; k < 4; k++)
                    cb.query_rec[4 * size_t(qs) + k] = make_float4(rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
                cb.query_scale[qs] = make_float4(thr_in.x, thr_in.y, thr_in.z, __uint_as_float(p));
            }
            // What this vertex adds reaches the records opened before it.
            add_to_records(cb, open, added);
            uint32_t ts = wave_append(b.counters + tt::C_CACHE_TRAIN, emit_train);
            uint32_t js = wave_append(b.counters + tt::C_CACHE_JUDGE, emit_judge);
            uint32_t slot = tt::CACHE_NO_RECORD;
            if (emit_train && ts < cb.train_capacity) slot = ts;
            if (emit_judge && js < cb.judge_capacity) slot = cb.judge_base + js;
            if (slot != tt::CACHE_NO_RECORD) {
                for (int k = 0; k < 4; k++)
                    cb.train_rec[4 * size_t(slot) + k] = make_float4(rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
                cb.train_thr[slot] = make_float4(thr_in.x, thr_in.y, thr_in.z, __uint_as_float(p));
                f3 t = cache_ratio(added, thr_in);
                uint32_t n = open.w & 3u;
                // w: path x records a path takes + the path's record index,
                // the key training orders records by (slots follow the order
                // waves finish in).
                cb.train_target[slot] = make_float4(t.x, t.y, t.z, __uint_as_float(p * f.cache_records + n));
                if (n == 0) open.x = slot;
                else if (n == 1) open.y = slot;
                else open.z = slot;
                open.w += 1u;
            }
            if (emit_train || emit_judge) {
                open.w |= tt::CACHE_DECIDED | (emit_judge ? CACHE_JUDGE_PATH : 0u);
                cb.open[p] = open;
            }
        }
    }

    // Continuation and shadow rays binned by direction octant across the
    // workgroup: one atomic per queue reserves the workgroup's rays, which
    // take consecutive slots octant by octant, so the traversal kernels'
    // 64-ray batches mix fewer directions.  Slot order only decides which
    // lanes trace which rays.
    __shared__ uint32_t octant_count[16];  // continuation rays 0-7, shadow rays 8-15
    __shared__ uint32_t octant_base[16];
    if (threadIdx.x < 16) octant_count[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t c_oct = (cont_dir.x < 0.0f ? 1u : 0u) | (cont_dir.y < 0.0f ? 2u : 0u) | (cont_dir.z < 0.0f ? 4u : 0u);
    const uint32_t s_oct = 8u | (shw1.x < 0.0f ? 1u : 0u) | (shw1.y < 0.0f ? 2u : 0u) | (shw1.z < 0.0f ? 4u : 0u);
    uint32_t c_rank = 0, s_rank = 0;
    if (cont_out) c_rank = atomicAdd(&octant_count[c_oct], 1u);
    if (shw_out) s_rank = atomicAdd(&octant_count[s_oct], 1u);
    __syncthreads();
    if (threadIdx.x < 2) {
        // Thread 0 the continuation queue, thread 1 the shadow queue.
        const uint32_t q0 = 8u * threadIdx.x;
        uint32_t total = 0;
        for (uint32_t k = 0; k < 8; k++) {
            octant_base[q0 + k] = total;
            total += octant_count[q0 + k];
        }
        uint32_t* counter = threadIdx.x == 0 ? b.counters + C_EXT + bounce + 1 : b.counters + C_SHADOW + bounce;
        const uint32_t first = total ? atomicAdd(counter, total) : 0u;
        for (uint32_t k = 0; k < 8; k++) octant_base[q0 + k] += first;
    }
    __syncthreads();
    if (cont_out) {
        const uint32_t slot = octant_base[c_oct] + c_rank;
        b.queue[(bounce + 1) & 1][0][slot] = make_float4(cont_org.x, cont_org.y, cont_org.z, __uint_as_float(cont_path));
        b.queue[(bounce + 1) & 1][1][slot] = make_float4(cont_dir.x, cont_dir.y, cont_dir.z, 0.0f);
    }
    if (shw_out) {
        const uint32_t slot = octant_base[s_oct] + s_rank;
        b.shadow[0][slot] = shw0;
        b.shadow[1][slot] = shw1;
        b.shadow[2][slot] = shw2;
        b.shadow[3][slot] = shw3;
    }
}

// ---------------------------------------------------------------- shadow --
// integrate.glsl:168-192: closest hit without culling in [1e-3, 1e4];
// visible iff the hit is the sampled light's instance and
// (primitive + index_offset / 3) % 0xffff equals the light's hash.  The query
// is confined to the light instance's box interval and ends at the first
// occluder in front of it (ShadowQuery), with the same outcome.
__global__ __launch_bounds__(TRACE_BLOCK) void shadow_kernel(tt::DeviceScene d, tt::PathBuffers b, uint32_t bounce) {
    SceneView s = view(d);
    Stack st = make_stack(b.spill);
    const uint32_t count = b.counters[C_SHADOW + bounce];
    WorkFetch wf = work_fetch(b.work + (tt::W_SHADOW + bounce) * tt::WORK_SET, count);
    uint32_t first, batch;
    while (next_batch(wf, first, batch)) {
        uint32_t i = first + lane_id();
        if (lane_id() < batch) {
            float4 o = b.shadow[0][i], dd = b.shadow[1][i];
            float4 direct = b.shadow[2][i];
            float4 e = b.shadow[3][i];
            uint32_t meta = __float_as_uint(direct.w);
            bool visible = light_visible(s, f3(o.x, o.y, o.z), f3(dd.x, dd.y, dd.z), __float_as_uint(dd.w),
                                         meta & 0xffffu, st);
            bool emission = (meta & 0x10000u) != 0;
            if (visible && b.cache.open) {
                // Direct light reaching a training or judge path: it counts
                // for the records at this vertex and before it.
                uint32_t p = __float_as_uint(o.w);
                add_to_records(b.cache, b.cache.open[p], f3(direct.x, direct.y, direct.z));
            }
            if (visible || emission) {
                uint32_t p = __float_as_uint(o.w);
                float4 a = b.radiance[p];
                f3 acc(a.x, a.y, a.z);
                if (visible) acc = acc + f3(direct.x, direct.y, direct.z);
                if (emission) acc = acc + f3(e.x, e.y, e.z);
                b.radiance[p] = make_float4(acc.x, acc.y, acc.z, 0.0f);
            }
        }
    }
}

// ------------------------------------------------------------ accumulate --
// integrate.glsl:338-346: with n = accumulated_frame_count != 0 the target
// becomes (previous * (n - 1) + new) / n.
TT_DEV void accumulate_pixel(const tt::FrameArgs& f, float4* target, uint32_t pixel, f3 acc) {
    if (f.accumulated_frame_count != 0) {
        float4 prev = target[pixel];
        acc = (f3(prev.x, prev.y, prev.z) * float(f.accumulated_frame_count - 1) + acc) / float(f.accumulated_frame_count);
    }
    target[pixel] = make_float4(acc.x, acc.y, acc.z, 1.0f);
}

__global__ __launch_bounds__(256) void accumulate_kernel(tt::FrameArgs f, tt::PathBuffers b, float4* target) {
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= f.width * f.height) return;
    f3 acc;
    for (uint32_t s = 0; s < f.sample_count; s++) {
        float4 r = b.radiance[size_t(pixel) * f.sample_count + s];
        acc = acc + f3(r.x, r.y, r.z);
    }
    acc = acc / float(f.sample_count);
    accumulate_pixel(f, target, pixel, acc);
}

// Serving: the chunk generated by generate_chunk_kernel (one sample).
__global__ __launch_bounds__(256) void accumulate_chunk_kernel(tt::FrameArgs f, tt::PathBuffers b, uint32_t base,
                                                               uint32_t pixel0, uint32_t count, float4* target) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 r = b.radiance[base + i];
    accumulate_pixel(f, target, pixel0 + i, f3(r.x, r.y, r.z) / 1.0f);
}

// ---------------------------------------------------------------- lights --
__global__ void light_records_kernel(tt::DeviceScene d, float4* out) {
    SceneView s = view(d);
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= s.emissive_count) return;
    store_light_record(out + 5 * size_t(li), light_triangle(s, li));
}

// ------------------------------------------------------------------ misc --
__global__ void zero_kernel(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Cursor c of consecutive cursor sets sits at word c * WORK_STRIDE.
__global__ void zero_work_kernel(uint32_t* work, uint32_t cursors) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cursors) work[size_t(i) * tt::WORK_STRIDE] = 0;
}

__global__ void slot_keys_kernel(uint2* key, uint32_t first, uint32_t frame, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = make_uint2(first + i, frame);
}

// sum[i] += counters[i]: a render's ray counters into the statistics sums
// (renders on several lanes may add at once).
__global__ void stats_add_kernel(const uint32_t* counters, unsigned long long* sum, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && counters[i]) atomicAdd(sum + i, (unsigned long long)counters[i]);
}

// BdsfSampler (src/material/test.rs:38-83): prepare from a fixed generator
// state, then evaluate.
__global__ void bsdf_evaluate_kernel(tt_material_constants mc, const tt_bsdf_query* q, tt_bsdf_result* out,
                                     uint32_t count, uint32_t rng_state) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 words[12];
    __builtin_memcpy(words, &mc, sizeof(words));
    Mat m = load_material(words);
    Generator g{rng_state, 0u};
    f3 wi(q[i].wi[0], q[i].wi[1], q[i].wi[2]), wo(q[i].wo[0], q[i].wo[1], q[i].wo[2]);
    Lobes l = bsdf_prepare(m, wi, g);
    float density;
    f3 v = bsdf_evaluate(m, l, wi, wo, density);
    tt_bsdf_result r;
    r.bsdf[0] = v.x;
    r.bsdf[1] = v.y;
    r.bsdf[2] = v.z;
    r.density = density;
    out[i] = r;
}

__global__ void math_kernel(int op, const float* x, const float* y, float* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = x[i];
    float r;
    switch (op) {
        case 0: r = ttm::sin(v); break;
        case 1: r = ttm::cos(v); break;
        case 2: r = ttm::exp(v); break;
        case 3: r = ttm::log(v); break;
        case 4: r = ttm::acos(v); break;
        case 5: r = ttm::asin(v); break;
        case 6: r = ttm::pow(v, y[i]); break;
        case 7: r = ttm::sqrt(v); break;
        case 8: r = v / y[i]; break;
        case 10: r = ttm::half_to_float(ttm::float_to_half(v)); break;
        case 11: {
            float c;
            ttm::sincospi(v, r, c);
            break;
        }
        case 12: {
            float sn;
            ttm::sincospi(v, sn, r);
            break;
        }
        default: r = 1.0f / v; break;
    }
    out[i] = r;
}

uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

namespace tt {

// Ten 128-thread workgroups per CU: five waves per SIMD, the occupancy that
// extend's register budget and the LDS stack allow.
#ifndef TT_TRACE_GRID_PER_CU
#define TT_TRACE_GRID_PER_CU 10
#endif
int trace_grid(int cu_count) { return cu_count * TT_TRACE_GRID_PER_CU; }
size_t spill_words(int grid) { return size_t(grid) * TRACE_BLOCK * TRACE_SPILL_STACK; }

hipError_t launch_zero(hipStream_t s, uint32_t* p, uint32_t words) {
    zero_kernel<<<blocks(words, 256), 256, 0, s>>>(p, words);
    return hipGetLastError();
}

hipError_t launch_zero_work(hipStream_t s, uint32_t* work, uint32_t sets) {
    const uint32_t cursors = sets * WORK_PARTS;
    zero_work_kernel<<<blocks(cursors, 256), 256, 0, s>>>(work, cursors);
    return hipGetLastError();
}

hipError_t launch_slot_keys(hipStream_t s, uint2* key, uint32_t first, uint32_t frame, uint32_t count) {
    if (count == 0) return hipSuccess;
    slot_keys_kernel<<<blocks(count, 256), 256, 0, s>>>(key, first, frame, count);
    return hipGetLastError();
}

hipError_t launch_stats_add(hipStream_t s, const uint32_t* counters, unsigned long long* sum, uint32_t words) {
    stats_add_kernel<<<blocks(words, 256), 256, 0, s>>>(counters, sum, words);
    return hipGetLastError();
}

hipError_t launch_trace_rays(hipStream_t s, const DeviceScene& scene, const tt_ray* rays, tt_hit* hits, uint32_t count,
                             bool cull, uint32_t* work, uint32_t* spill, int grid) {
    if (cull)
        trace_rays_kernel<true><<<grid, TRACE_BLOCK, 0, s>>>(scene, rays, hits, count, work, spill);
    else
        trace_rays_kernel<false><<<grid, TRACE_BLOCK, 0, s>>>(scene, rays, hits, count, work, spill);
    return hipGetLastError();
}

hipError_t launch_generate(hipStream_t s, const FrameArgs& f, const PathBuffers& b, uint32_t paths) {
    generate_kernel<<<blocks(paths, 256), 256, 0, s>>>(f, b, paths);
    return hipGetLastError();
}

hipError_t launch_generate_chunk(hipStream_t s, const FrameArgs& f, const PathBuffers& b, uint32_t base,
                                 uint32_t pixel0, uint32_t count) {
    generate_chunk_kernel<<<blocks(count, 256), 256, 0, s>>>(f, b, base, pixel0, count);
    return hipGetLastError();
}

hipError_t launch_accumulate_chunk(hipStream_t s, const FrameArgs& f, const PathBuffers& b, uint32_t base,
                                   uint32_t pixel0, uint32_t count, float4* target) {
    accumulate_chunk_kernel<<<blocks(count, 256), 256, 0, s>>>(f, b, base, pixel0, count, target);
    return hipGetLastError();
}

hipError_t launch_light_records(hipStream_t s, const DeviceScene& scene, void* out) {
    if (scene.emissive_count == 0) return hipSuccess;
    light_records_kernel<<<blocks(scene.emissive_count, 64), 64, 0, s>>>(scene, static_cast<float4*>(out));
    return hipGetLastError();
}

hipError_t launch_extend(hipStream_t s, const DeviceScene& scene, const PathBuffers& b, uint32_t bounce, int grid) {
    extend_kernel<<<grid, TRACE_BLOCK, 0, s>>>(scene, b, bounce);
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const DeviceScene& scene, const FrameArgs& f, const PathBuffers& b,
                        uint32_t bounce, uint32_t paths) {
    if (scene.texture_count > 0)
        shade_kernel<true><<<blocks(paths, 256), 256, 0, s>>>(scene, f, b, bounce);
    else
        shade_kernel<false><<<blocks(paths, 256), 256, 0, s>>>(scene, f, b, bounce);
    return hipGetLastError();
}

hipError_t launch_shadow(hipStream_t s, const DeviceScene& scene, const PathBuffers& b, uint32_t bounce, int grid) {
    shadow_kernel<<<grid, TRACE_BLOCK, 0, s>>>(scene, b, bounce);
    return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t s, const FrameArgs& f, const PathBuffers& b, float4* target) {
    accumulate_kernel<<<blocks(f.width * f.height, 256), 256, 0, s>>>(f, b, target);
    return hipGetLastError();
}

hipError_t launch_bsdf_evaluate(hipStream_t s, const tt_material_constants* material, const tt_bsdf_query* q,
                                tt_bsdf_result* r, uint32_t count, uint32_t rng_state) {
    bsdf_evaluate_kernel<<<blocks(count, 64), 64, 0, s>>>(*material, q, r, count, rng_state);
    return hipGetLastError();
}

hipError_t launch_math(hipStream_t s, int op, const float* x, const float* y, float* out, uint32_t count) {
    math_kernel<<<blocks(count, 256), 256, 0, s>>>(op, x, y, out, count);
    return hipGetLastError();
}

}  // namespace tt
