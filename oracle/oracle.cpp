// ORACLE (test infrastructure only).  CPU restatement of tinytrace's
// path-integration hot path; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it, as the checker.  Never linked into or
// called by the product.
//
// C entry points (ctypes):
//   oracle_scene_create / destroy   load a tt_scene_desc (float or double)
//   oracle_render                    whole image + progressive accumulation
//                                    (integrate.glsl:323-347)
//   oracle_render_pixels             sample-averaged radiance of listed pixels
//   oracle_trace                     closest hits (integrate.glsl:112-127, :168-185)
//   oracle_bsdf_evaluate / _sample   bsdf_prepare + evaluate/sample (bsdf.glsl)
//   oracle_rng                       generator draws (random.glsl)
//   oracle_math_f32                  the shared transcendental functions
//   oracle_camera_direction          camera_ray_direction (integrate.glsl:17-20)
#include <atomic>
#include <chrono>
#include <cstdio>
#include <memory>
#include <thread>

#include "oracle_restir.hpp"

using namespace oracle;

namespace {

struct OracleScene {
    tt_scene_desc desc;
    int precision;
    std::unique_ptr<Scene<float>> f;
    std::unique_ptr<Scene<double>> d;
};

template <typename R>
Camera<R> make_camera(const tt_constants& k) {
    Camera<R> c;
    c.inverse_view = load_m4<R>(k.inverse_view);
    c.inverse_proj = load_m4<R>(k.inverse_proj);
    c.position = V3<R>(R(k.camera_position[0]), R(k.camera_position[1]), R(k.camera_position[2]));
    c.width = k.screen_size[0];
    c.height = k.screen_size[1];
    return c;
}

template <typename F>
void parallel_for(uint32_t count, uint32_t chunk, int threads, F&& body) {
    if (threads <= 0) threads = int(std::thread::hardware_concurrency());
    if (threads < 1) threads = 1;
    std::atomic<uint32_t> next{0};
    auto worker = [&](int tid) {
        for (;;) {
            uint32_t begin = next.fetch_add(chunk);
            if (begin >= count) break;
            uint32_t end = std::min(count, begin + chunk);
            for (uint32_t i = begin; i < end; i++) body(tid, i);
        }
    };
    if (threads == 1) {
        worker(0);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(worker, t);
    for (auto& t : pool) t.join();
}

template <typename R>
void render_pixels_impl(const Scene<R>& s, const tt_constants& k, const Settings& st, const uint32_t* pixels,
                        uint32_t n, float* out_rgb, int threads, uint64_t* rays) {
    Camera<R> cam = make_camera<R>(k);
    int nthreads = threads <= 0 ? int(std::thread::hardware_concurrency()) : threads;
    std::vector<TraceCounts> counts(std::max(1, nthreads));
    parallel_for(n, 64, nthreads, [&](int tid, uint32_t i) {
        uint32_t p = pixels[i];
        uint32_t x = p % k.screen_size[0], y = p / k.screen_size[0];
        V3<R> c = render_pixel(s, cam, k, st, x, y, counts[tid]);
        out_rgb[3 * i + 0] = float(c.x);
        out_rgb[3 * i + 1] = float(c.y);
        out_rgb[3 * i + 2] = float(c.z);
    });
    if (rays) {
        rays[0] = rays[1] = 0;
        for (auto& c : counts) {
            rays[0] += c.extension;
            rays[1] += c.shadow;
        }
    }
}

template <typename R>
void render_image_impl(const Scene<R>& s, const tt_constants& k, const Settings& st, float* rgba, int threads,
                       uint64_t* rays) {
    Camera<R> cam = make_camera<R>(k);
    uint32_t w = k.screen_size[0], h = k.screen_size[1];
    int nthreads = threads <= 0 ? int(std::thread::hardware_concurrency()) : threads;
    std::vector<TraceCounts> counts(std::max(1, nthreads));
    // 16x16 tiles from an atomic counter.
    uint32_t tiles_x = (w + 15) / 16, tiles_y = (h + 15) / 16;
    parallel_for(tiles_x * tiles_y, 1, nthreads, [&](int tid, uint32_t tile) {
        uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
        for (uint32_t y = ty * 16; y < std::min(h, ty * 16 + 16); y++)
            for (uint32_t x = tx * 16; x < std::min(w, tx * 16 + 16); x++) {
                V3<R> c = render_pixel(s, cam, k, st, x, y, counts[tid]);
                float* px = rgba + 4 * (size_t(y) * w + x);
                // integrate.glsl:341-346
                if (k.accumulated_frame_count != 0) {
                    V3<R> prev{R(px[0]), R(px[1]), R(px[2])};
                    c = (prev * R(float(k.accumulated_frame_count - 1)) + c) / R(float(k.accumulated_frame_count));
                }
                px[0] = float(c.x);
                px[1] = float(c.y);
                px[2] = float(c.z);
                px[3] = 1.0f;
            }
    });
    if (rays) {
        rays[0] = rays[1] = 0;
        for (auto& c : counts) {
            rays[0] += c.extension;
            rays[1] += c.shadow;
        }
    }
}

}  // namespace

extern "C" {

void* oracle_scene_create(const tt_scene_desc* desc, int precision, int use_bvh) {
    auto* s = new OracleScene();
    s->desc = *desc;
    s->precision = precision;
    if (precision == 64) {
        s->d.reset(new Scene<double>());
        s->d->load(&s->desc, use_bvh != 0);
    } else {
        s->f.reset(new Scene<float>());
        s->f->load(&s->desc, use_bvh != 0);
    }
    return s;
}

void oracle_scene_destroy(void* handle) { delete static_cast<OracleScene*>(handle); }

// Whole image; `rgba` holds the previous target contents (read when
// accumulated_frame_count != 0) and receives the new ones.
int oracle_render(void* handle, const tt_constants* k, uint32_t rng, uint32_t seed, float* rgba, int threads,
                  uint64_t* rays) {
    auto* s = static_cast<OracleScene*>(handle);
    if (k->light_sampling != TT_LIGHT_SAMPLING_ON_HIT && k->light_sampling != TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION)
        return 1;
    Settings st{rng, seed};
    if (s->precision == 64)
        render_image_impl(*s->d, *k, st, rgba, threads, rays);
    else
        render_image_impl(*s->f, *k, st, rgba, threads, rays);
    return 0;
}

// Sample-averaged radiance (before progressive accumulation) of the listed
// linear pixel indices.
int oracle_render_pixels(void* handle, const tt_constants* k, uint32_t rng, uint32_t seed, const uint32_t* pixels,
                         uint32_t n, float* out_rgb, int threads, uint64_t* rays) {
    auto* s = static_cast<OracleScene*>(handle);
    if (k->light_sampling != TT_LIGHT_SAMPLING_ON_HIT && k->light_sampling != TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION)
        return 1;
    Settings st{rng, seed};
    if (s->precision == 64)
        render_pixels_impl(*s->d, *k, st, pixels, n, out_rgb, threads, rays);
    else
        render_pixels_impl(*s->f, *k, st, pixels, n, out_rgb, threads, rays);
    return 0;
}

// One ReSTIR DI frame (float oracle).  gbuffer: 2 x W*H x RECORD floats,
// reservoirs: 2 x W*H x 4 words (final reservoirs); [cur] is written, [cur^1]
// is last frame's and is read when `history` != 0.  rgba as oracle_render.
int oracle_render_restir(void* handle, const tt_constants* k, uint32_t rng, uint32_t seed, uint32_t candidates,
                         int temporal, int spatial, float* gbuffer, uint32_t* reservoirs, uint32_t cur,
                         uint32_t history, float* rgba, int threads, uint64_t* rays) {
    auto* os = static_cast<OracleScene*>(handle);
    if (os->precision != 32 || k->sample_count != 1) return 1;
    const Scene<float>& s = *os->f;
    Camera<float> cam = make_camera<float>(*k);
    uint32_t w = k->screen_size[0], h = k->screen_size[1];
    uint32_t px = w * h;
    uint32_t frame = k->frame_index;
    int nthreads = threads <= 0 ? int(std::thread::hardware_concurrency()) : threads;
    std::vector<TraceCounts> counts(std::max(1, nthreads));
    std::vector<Generator> gens(px);
    std::vector<V3<float>> dirs(px);
    std::vector<Hit> hits(px);
    std::vector<uint8_t> found(px);
    std::vector<V2<float>> ndcs(px);
    float* gcur = gbuffer + size_t(cur) * px * RECORD;
    const float* gprev = gbuffer + size_t(cur ^ 1u) * px * RECORD;
    uint32_t* fcur = reservoirs + size_t(cur) * px * 4;
    const uint32_t* fprev = reservoirs + size_t(cur ^ 1u) * px * 4;
    std::vector<uint32_t> initial(size_t(px) * 4), temporal_out(size_t(px) * 4);
    bool has_lights = s.desc->emissive_triangle_count > 0;
    Settings st{rng, seed};

    // Camera rays and primary hits; G-buffer records.
    parallel_for(px, 64, nthreads, [&](int tid, uint32_t p) {
        uint32_t x = p % w, y = p / w;
        Generator g = init_generator_from_pixel(x, y, w, frame, st.rng, st.seed);
        float ox = random_float<float>(g);
        float oy = random_float<float>(g);
        ox = ox - 0.5f;
        oy = oy - 0.5f;
        float ndc_x = ((float(x) + ox) / float(w)) * 2.0f - 1.0f;
        float ndc_y = ((float(y) + oy) / float(h)) * 2.0f - 1.0f;
        V3<float> dir = camera_ray_direction(cam, ndc_x, ndc_y);
        gens[p] = g;
        dirs[p] = dir;
        counts[tid].extension++;
        found[p] = s.closest_hit(cam.position, dir, true, 1.0e-3f, 1000.0f, hits[p]) ? 1 : 0;
        float* g16 = gcur + size_t(p) * RECORD;
        ndcs[p] = V2<float>{ndc_x, ndc_y};
        if (!found[p]) {
            g16[4] = g16[5] = g16[6] = g16[7] = 0.0f;
            return;
        }
        RayHit<float> hit = get_ray_hit(s, hits[p]);
        ShadingF q;
        if (s.textured()) q.lod = first_hit_lod(cam, hit, ndc_x, ndc_y);
        q.u = hit.tu;
        q.v = hit.tv;
        q.m = s.material_at(hit.material_index, q.u, q.v, q.lod);
        q.material = hit.material_index;
        q.position = hit.world_position;
        q.basis = create_surface_basis(hit.tangent_space, q.m.geometry_normal);
        q.wi = transform_from_basis(q.basis, -dir);
        q.valid = has_lights && q.m.specular_roughness > 0.15f;
        q.depth = length(q.position - cam.position);
        write_record(g16, q, -dir, frame);
    });
    if (has_lights) {
        // Candidates and visibility.
        uint32_t* after_candidates = (temporal || spatial) ? initial.data() : fcur;
        parallel_for(px, 64, nthreads, [&](int tid, uint32_t p) {
            Res res;
            const float* g16 = gcur + size_t(p) * RECORD;
            if (found[p] && g16[7] != 0.0f) {
                ShadingF q = read_record(s, g16, p);
                Generator gen = pass_generator(p, frame, seed, 1u);
                float sum = 0.0f, p_sel = 0.0f;
                uint32_t sel_light = 0xffffffffu, sel_bary = 0u;
                for (uint32_t j = 0; j < candidates; j++) {
                    uint32_t li = random_uint(gen, s.desc->emissive_triangle_count);
                    V2<float> r2 = random_vec2<float>(gen);
                    uint32_t bary = quantize_bary(sample_triangle(r2));
                    float area;
                    float ph = restir_target(s, q, li, bary, &area);
                    float wgt = ph * (float(s.desc->emissive_triangle_count) * area);
                    sum = sum + wgt;
                    float u = random_float<float>(gen);
                    if (u * sum < wgt) {
                        sel_light = li;
                        sel_bary = bary;
                        p_sel = ph;
                    }
                }
                res.M = float(candidates);
                if (sel_light != 0xffffffffu && p_sel > 0.0f) {
                    res.light = sel_light;
                    res.bary = sel_bary;
                    res.W = (sum / float(candidates)) / p_sel;
                    LightPointF lp = light_point(s, sel_light, reservoir_bary(sel_bary));
                    V3<float> to_light = lp.position - q.position;
                    float dist = Fn<float>::sqrt(length_squared(to_light));
                    V3<float> dir = to_light / dist;
                    if (!light_visible(s, q.position, dir, lp, counts[tid])) res.W = 0.0f;
                }
            }
            pack(res, after_candidates + size_t(p) * 4);
        });
        const uint32_t* current = after_candidates;
        if (temporal) {
            uint32_t* dst = spatial ? temporal_out.data() : fcur;
            parallel_for(px, 64, nthreads, [&](int, uint32_t p) {
                Res rq = unpack(current + size_t(p) * 4);
                ShadingF q = read_record(s, gcur + size_t(p) * RECORD, p);
                if (q.valid && history) {
                    ShadingF o = read_record(s, gprev + size_t(p) * RECORD, p);
                    if (similar(q, o)) {
                        Res ro = unpack(fprev + size_t(p) * 4);
                        ro.M = rmin(ro.M, 20.0f * float(candidates));
                        Generator gen = pass_generator(p, frame, seed, 2u);
                        rq = combine(s, q, rq, o, ro, gen);
                    }
                }
                pack(rq, dst + size_t(p) * 4);
            });
            current = dst;
        }
        if (spatial) {
            parallel_for(px, 64, nthreads, [&](int, uint32_t p) {
                Res rq = unpack(current + size_t(p) * 4);
                ShadingF q = read_record(s, gcur + size_t(p) * RECORD, p);
                if (q.valid) {
                    Generator gen = pass_generator(p, frame, seed, 3u);
                    float angle = 2.0f * K<float>::PI * random_float<float>(gen);
                    float radius = 30.0f * Fn<float>::sqrt(random_float<float>(gen));
                    int x = int(p % w), y = int(p / w);
                    int nx = int(std::floor(float(x) + radius * ttm::cos(angle) + 0.5f));
                    int ny = int(std::floor(float(y) + radius * ttm::sin(angle) + 0.5f));
                    nx = nx < 0 ? 0 : (nx >= int(w) ? int(w) - 1 : nx);
                    ny = ny < 0 ? 0 : (ny >= int(h) ? int(h) - 1 : ny);
                    uint32_t n = uint32_t(ny) * w + uint32_t(nx);
                    if (n != p) {
                        ShadingF o = read_record(s, gcur + size_t(n) * RECORD, n);
                        if (similar(q, o)) rq = combine(s, q, rq, o, unpack(current + size_t(n) * 4), gen);
                    }
                }
                pack(rq, fcur + size_t(p) * 4);
            });
        }
    }
    // Paths, with the final reservoir at bounce 0; progressive accumulation.
    parallel_for(px, 64, nthreads, [&](int tid, uint32_t p) {
        V3<float> c = trace_path_restir(s, cam, k->bounce_count, dirs[p], gens[p], found[p] != 0, hits[p],
                                        fcur + size_t(p) * 4, p, frame, counts[tid], ndcs[p].x, ndcs[p].y);
        c = c / float(k->sample_count);
        float* out = rgba + 4 * size_t(p);
        if (k->accumulated_frame_count != 0) {
            V3<float> prev{out[0], out[1], out[2]};
            c = (prev * float(k->accumulated_frame_count - 1) + c) / float(k->accumulated_frame_count);
        }
        out[0] = c.x;
        out[1] = c.y;
        out[2] = c.z;
        out[3] = 1.0f;
    });
    if (rays) {
        rays[0] = rays[1] = 0;
        for (auto& ct : counts) {
            rays[0] += ct.extension;
            rays[1] += ct.shadow;
        }
    }
    return 0;
}

int oracle_trace(void* handle, const tt_ray* rays, tt_hit* hits, uint32_t n, uint32_t flags, int threads) {
    auto* s = static_cast<OracleScene*>(handle);
    bool cull = (flags & TT_TRACE_CULL_BACK_FACES) != 0;
    parallel_for(n, 256, threads, [&](int, uint32_t i) {
        const tt_ray& r = rays[i];
        Hit h;
        bool found;
        if (s->precision == 64) {
            found = s->d->closest_hit(V3<double>(r.origin[0], r.origin[1], r.origin[2]),
                                      V3<double>(r.direction[0], r.direction[1], r.direction[2]), cull, r.tmin, r.tmax, h);
        } else {
            found = s->f->closest_hit(V3<float>(r.origin[0], r.origin[1], r.origin[2]),
                                      V3<float>(r.direction[0], r.direction[1], r.direction[2]), cull, r.tmin, r.tmax, h);
        }
        if (found) {
            hits[i].t = h.t;
            hits[i].instance = h.instance;
            hits[i].primitive = h.primitive;
            hits[i].barycentrics[0] = h.b1;
            hits[i].barycentrics[1] = h.b2;
        } else {
            hits[i].t = r.tmax;
            hits[i].instance = TT_INVALID_INDEX;
            hits[i].primitive = TT_INVALID_INDEX;
            hits[i].barycentrics[0] = 0.0f;
            hits[i].barycentrics[1] = 0.0f;
        }
    });
    return 0;
}

// bsdf_prepare + bsdf_evaluate per query, each from generator `rng_state`
// (xorshift), as src/material/test.rs:38-83.
int oracle_bsdf_evaluate(const tt_material_constants* mc, const tt_bsdf_query* q, tt_bsdf_result* out, uint32_t n,
                         uint32_t rng_state, int precision) {
    for (uint32_t i = 0; i < n; i++) {
        Generator g{rng_state, 0};
        if (precision == 64) {
            Material<double> m = load_material<double>(&mc->base_color[0]);
            V3<double> wi(q[i].wi[0], q[i].wi[1], q[i].wi[2]), wo(q[i].wo[0], q[i].wo[1], q[i].wo[2]);
            Lobes<double> l = bsdf_prepare(m, wi, g);
            double density;
            V3<double> b = bsdf_evaluate(m, l, wi, wo, density);
            out[i] = {{float(b.x), float(b.y), float(b.z)}, float(density)};
        } else {
            Material<float> m = load_material<float>(&mc->base_color[0]);
            V3<float> wi(q[i].wi[0], q[i].wi[1], q[i].wi[2]), wo(q[i].wo[0], q[i].wo[1], q[i].wo[2]);
            Lobes<float> l = bsdf_prepare(m, wi, g);
            float density;
            V3<float> b = bsdf_evaluate(m, l, wi, wo, density);
            out[i] = {{b.x, b.y, b.z}, density};
        }
    }
    return 0;
}

// bsdf_prepare + bsdf_sample from generator state rng_states[i]; then
// bsdf_evaluate at the sampled direction with the same lobes.  out: per
// query 12 floats: wo[3], value[3], density, ok, lobe, eval[3] ... eval_density.
int oracle_bsdf_sample(const tt_material_constants* mc, const float* wi_in, const uint32_t* rng_states, uint32_t n,
                       float* out, int precision) {
    for (uint32_t i = 0; i < n; i++) {
        float* o = out + 13 * i;
        Generator g{rng_states[i], 0};
        if (precision == 64) {
            Material<double> m = load_material<double>(&mc->base_color[0]);
            V3<double> wi(wi_in[3 * i], wi_in[3 * i + 1], wi_in[3 * i + 2]);
            Lobes<double> l = bsdf_prepare(m, wi, g);
            V3<double> wo, value;
            double density;
            int lobe;
            bool ok = bsdf_sample(m, l, wi, g, wo, density, lobe, value);
            double ed = 0;
            V3<double> ev = ok ? bsdf_evaluate(m, l, wi, wo, ed) : V3<double>();
            float vals[13] = {float(wo.x), float(wo.y), float(wo.z), float(value.x), float(value.y), float(value.z),
                              float(density), ok ? 1.0f : 0.0f, float(lobe), float(ev.x), float(ev.y), float(ev.z),
                              float(ed)};
            std::memcpy(o, vals, sizeof(vals));
        } else {
            Material<float> m = load_material<float>(&mc->base_color[0]);
            V3<float> wi(wi_in[3 * i], wi_in[3 * i + 1], wi_in[3 * i + 2]);
            Lobes<float> l = bsdf_prepare(m, wi, g);
            V3<float> wo, value;
            float density;
            int lobe;
            bool ok = bsdf_sample(m, l, wi, g, wo, density, lobe, value);
            float ed = 0;
            V3<float> ev = ok ? bsdf_evaluate(m, l, wi, wo, ed) : V3<float>();
            float vals[13] = {wo.x, wo.y, wo.z, value.x, value.y, value.z, density, ok ? 1.0f : 0.0f, float(lobe),
                              ev.x, ev.y, ev.z, ed};
            std::memcpy(o, vals, sizeof(vals));
        }
    }
    return 0;
}

// Generator draws: seeds with init_generator_from_pixel(x, y, width, frame)
// and writes `n` u32 draws.
void oracle_rng(uint32_t kind, uint32_t x, uint32_t y, uint32_t width, uint32_t frame, uint32_t seed, uint32_t n,
                uint32_t* out, uint64_t* initial_state) {
    Generator g = init_generator_from_pixel(x, y, width, frame, kind, seed);
    if (initial_state) *initial_state = g.state;
    for (uint32_t i = 0; i < n; i++) out[i] = next_u32(g);
}

uint32_t oracle_jenkins_hash(uint32_t x) { return jenkins_hash(x); }
// pcg32_srandom_r(initstate, initseq) followed by n pcg32_random_r draws.
void oracle_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    uint64_t inc = (initseq << 1u) | 1u;
    uint64_t s = pcg32_srandom_state(initstate, inc);
    for (uint32_t i = 0; i < n; i++) {
        out[i] = pcg32_output(s);
        s = s * PCG32_MULT + inc;
    }
}
uint64_t oracle_splitmix64(uint64_t x) { return splitmix64(x); }
float oracle_uint_to_unit_float(uint32_t v) { return uint_to_unit_float(v); }

// op: 0 sin, 1 cos, 2 exp, 3 log, 4 acos, 5 asin, 6 pow(x, y), 7 sqrt, 8 half round trip,
// 9 sin(pi x), 10 cos(pi x)
void oracle_math_f32(int op, const float* x, const float* y, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        float v = x[i];
        switch (op) {
            case 0: out[i] = ttm::sin(v); break;
            case 1: out[i] = ttm::cos(v); break;
            case 2: out[i] = ttm::exp(v); break;
            case 3: out[i] = ttm::log(v); break;
            case 4: out[i] = ttm::acos(v); break;
            case 5: out[i] = ttm::asin(v); break;
            case 6: out[i] = ttm::pow(v, y[i]); break;
            case 8: out[i] = ttm::half_to_float(ttm::float_to_half(v)); break;
            case 9: {
                float c;
                ttm::sincospi(v, out[i], c);
                break;
            }
            case 10: {
                float sn;
                ttm::sincospi(v, sn, out[i]);
                break;
            }
            default: out[i] = ttm::sqrt(v); break;
        }
    }
}

void oracle_camera_direction(const tt_constants* k, float ndc_x, float ndc_y, float* out) {
    Camera<float> c = make_camera<float>(*k);
    V3<float> d = camera_ray_direction(c, ndc_x, ndc_y);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

// Samples of the BSDF pieces used by the closed-form tests.
float oracle_fresnel_dielectric(float c, float ior) { return fresnel_dielectric(c, ior); }
float oracle_microfacet_distribution(float ax, float ay, const float* m) {
    return microfacet_distribution(V2<float>{ax, ay}, V3<float>(m[0], m[1], m[2]));
}
void oracle_sample_triangle(float u, float v, float* out) {
    V3<float> b = sample_triangle(V2<float>{u, v});
    out[0] = b.x;
    out[1] = b.y;
    out[2] = b.z;
}
void oracle_cosine_hemisphere_sample(float u, float v, float* out) {
    V3<float> d = cosine_hemisphere_sample(V2<float>{u, v});
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}
void oracle_decode_tangent_frame(uint32_t encoded, float* out) {
    TangentFrame<float> f = decode_tangent_frame<float>(encoded);
    float v[7] = {f.normal.x, f.normal.y, f.normal.z, f.tangent.x, f.tangent.y, f.tangent.z, f.bitangent_sign};
    std::memcpy(out, v, sizeof(v));
}

// Post-process display (src/post_process/display.glsl:6-12): neutral_tonemap
// (src/post_process/tonemap.glsl:1-26) when `tonemap` != 0, then the sRGB
// transfer function and 8-bit rounding of an sRGB store; RGBA8, alpha 255.
void oracle_display(const float* rgba, uint32_t count, uint32_t tonemap, uint8_t* out) {
    auto maxf = [](float a, float b) { return a < b ? b : a; };
    auto curve = [](float v) { return 1.0f - ttm::exp(-v); };
    auto mix = [](float x, float y, float a) { return x * (1.0f - a) + y * a; };
    for (uint32_t i = 0; i < count; i++) {
        float c[3] = {rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2]};
        if (tonemap) {
            // rgb_to_ycbcr: color * mat3(columns) = dot with each column.
            float y = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
            float cb = c[0] * -0.1146f + c[1] * -0.3854f + c[2] * 0.5f;
            float cr = c[0] * 0.5f + c[1] * -0.4542f + c[2] * -0.0458f;
            float bt = curve(ttm::sqrt(cb * cb + cr * cr) * 2.4f);
            float desat = maxf((bt - 0.7f) * 0.8f, 0.0f);
            desat = desat * desat;
            float tm_lum = curve(y);
            float lum = 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2];
            float scale = maxf(0.0f, tm_lum / maxf(1e-5f, lum));
            float w = bt * bt;
            for (int k = 0; k < 3; k++) {
                float tm0 = c[k] * scale;
                float tm1 = curve(mix(c[k], y, desat));
                c[k] = mix(tm0, tm1, w) * 0.97f;
            }
        }
        for (int k = 0; k < 3; k++) {
            float v = c[k] > 0.0f ? (c[k] < 1.0f ? c[k] : 1.0f) : 0.0f;
            float srgb = v <= 0.0031308f ? 12.92f * v : 1.055f * ttm::pow(v, 1.0f / 2.4f) - 0.055f;
            out[4 * i + k] = uint8_t(uint32_t(srgb * 255.0f + 0.5f));
        }
        out[4 * i + 3] = 255;
    }
}

// Test entry points of the texture path: one sample of texture `index`
// (float precision), and the level of detail of the first hit of the
// primary ray at (ndc_x, ndc_y); 0 on success, 1 without a hit.
int oracle_sample_texture(void* handle, uint32_t index, float u, float v, float lod, float* out) {
    auto* os = static_cast<OracleScene*>(handle);
    if (index >= os->desc.texture_count) return 1;
    os->f->sample(index, u, v, lod, out);
    return 0;
}

// The 8-bit channels of texel (x, y) of a BC1 or BC5 level w texels wide.
void oracle_block_texel(uint32_t format, const uint8_t* data, uint32_t w, uint32_t x, uint32_t y, uint8_t* out) {
    Scene<float>::block_texel(format, data, w, x, y, out);
}

int oracle_primary_lod(void* handle, const tt_constants* k, float ndc_x, float ndc_y, float* lod) {
    auto* os = static_cast<OracleScene*>(handle);
    const Scene<float>& s = *os->f;
    Camera<float> cam = make_camera<float>(*k);
    V3<float> dir = camera_ray_direction(cam, ndc_x, ndc_y);
    Hit h;
    if (!s.closest_hit(cam.position, dir, true, 1.0e-3f, 1000.0f, h)) return 1;
    *lod = first_hit_lod(cam, get_ray_hit(s, h), ndc_x, ndc_y);
    return 0;
}

}  // extern "C"
