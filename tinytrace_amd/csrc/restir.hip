// ReSTIR DI at the primary vertex (Bitterli et al. 2020, "Spatiotemporal
// reservoir resampling for real-time ray tracing with dynamic direct
// lighting"; reuse weights from Lin et al. 2022, "Generalized resampled
// importance sampling").  The reference has no ReSTIR: in
// LightSampling::ReSTIR mode it replaces the uniform next-event estimate of
// direct_light_contribution (integrate.glsl:156-193) at bounce 0 only.
//
// Samples y are points on emissive triangles in the area measure, stored as
// (triangle, unorm16 barycentrics).  Source pdf: triangle uniformly, point
// uniformly, p(y) = 1 / (count * area).  Target at a shading point q:
//   p_hat_q(y) = luminance(f_q(wi, wo) Le(y)) |cos wo| cos_light / d^2
// (unshadowed), where f_q uses lobes prepared from a generator seeded by the
// pixel index, so p_hat_q is a deterministic function of the pixel's G-buffer
// record and can be evaluated at neighbours and across frames.
//
// Passes, one lane per pixel (spp 1):
//   candidates : M RIS candidates -> reservoir (y, W = sum w / (M p_hat(y)), M)
//   visibility : shadow query for y; W = 0 if occluded (visibility reuse)
//   temporal   : combine with last frame's final reservoir at the same pixel
//                (M clamped to 20x the candidate count) if the surfaces match
//   spatial    : combine with one neighbour within 30 px
// Two reservoirs combine with the generalized balance heuristic
//   m_i(y) = M_i p_hat_i(y) / sum_j M_j p_hat_j(y),
//   w_i = m_i(y_i) p_hat_q(y_i) W_i,  pick with probability w_i / sum w,
//   W = sum w / p_hat_q(y),  M = sum M_i.
// W and M are stored as IEEE half (FP16 reservoirs).
#include "device_surface.h"

using namespace ttd;

namespace {

using tt::C_RESTIR_VIS;
constexpr float RESTIR_RADIUS = 30.0f;
constexpr float RESTIR_MAX_HISTORY = 20.0f;

TT_DEV Generator pass_generator(const tt::FrameArgs& f, uint32_t pixel, uint32_t salt) {
    return init_generator(pixel, f.frame_index, 1u, f.seed ^ (salt * 0x9e3779b9u));
}

// A shading point rebuilt from its G-buffer record.
struct Shading {
    Surface sf;
    Mat m;
    Lobes lobes;
    float depth;
    bool valid;
};

// Record (RECORD float4s): g0 = (position, material), g1 = (normal, valid),
// g2 = (tangent, depth), g3 = (-ray direction, frame index), g4 = (texture
// coordinate, level of detail, 0), read on textured scenes only.
constexpr size_t RECORD = 5;

TT_DEV void write_record(float4* g, const Surface& sf, bool valid, float depth, f3 wi_world, uint32_t frame, float u,
                         float v, float lod) {
    g[0] = make_float4(sf.position.x, sf.position.y, sf.position.z, __uint_as_float(sf.material));
    g[1] = make_float4(sf.basis.normal.x, sf.basis.normal.y, sf.basis.normal.z, valid ? 1.0f : 0.0f);
    g[2] = make_float4(sf.basis.tangent.x, sf.basis.tangent.y, sf.basis.tangent.z, depth);
    g[3] = make_float4(wi_world.x, wi_world.y, wi_world.z, __uint_as_float(frame));
    g[4] = make_float4(u, v, lod, 0.0f);
}

template <bool TEXTURED>
TT_DEV Shading read_record(const SceneView& s, const float4* g, uint32_t pixel) {
    Shading q;
    float4 g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
    q.valid = g1.w != 0.0f;
    if (!q.valid) return q;
    q.sf.position = f3(g0.x, g0.y, g0.z);
    q.sf.material = __float_as_uint(g0.w);
    q.sf.basis.normal = f3(g1.x, g1.y, g1.z);
    q.sf.basis.tangent = f3(g2.x, g2.y, g2.z);
    q.sf.basis.bitangent = normalize(cross(q.sf.basis.normal, q.sf.basis.tangent));
    q.sf.wi = transform_from_basis(q.sf.basis, f3(g3.x, g3.y, g3.z));
    q.depth = g2.w;
    const float4* material = s.materials + 19 * size_t(q.sf.material);
    q.m = load_material(material);
    if constexpr (TEXTURED) {
        float4 g4 = g[4];
        Mat t = q.m;
        texture_hit(texture_view(s), material, &t, g4.x, g4.y, g4.z);
        q.m = t;
    }
    Generator lg = restir_lobe_generator(pixel, __float_as_uint(g3.w));
    q.lobes = bsdf_prepare(q.m, q.sf.wi, lg);
    return q;
}

template <bool TEXTURED>
TT_DEV float target(const SceneView& s, const Shading& q, uint32_t light, uint32_t bary, float* area = nullptr) {
    LightPoint lp = TEXTURED ? light_point_textured(s, light, reservoir_bary(bary))
                             : light_point(s, light, reservoir_bary(bary));
    if (area) *area = lp.area;
    f3 to_light = lp.position - q.sf.position;
    f3 wo = transform_from_basis(q.sf.basis, to_light);
    float d2 = length_squared(to_light);
    f3 dir = to_light / sqrt_(d2);
    float cos_light = saturate(dot(-dir, lp.normal));
    float density;
    f3 f = bsdf_evaluate(q.m, q.lobes, q.sf.wi, wo, density);
    f3 c = f * lp.emission;
    float lum = dot(c, f3(0.2126f, 0.7152f, 0.0722f));
    float p = ((lum * abs_(cos_theta(wo))) * cos_light) / d2;
    return p > 0.0f && p < INFINITY ? p : 0.0f;
}

TT_DEV bool similar(const Shading& a, const Shading& b) {
    return a.valid && b.valid && dot(a.sf.basis.normal, b.sf.basis.normal) >= 0.9f &&
        abs_(a.depth - b.depth) <= 0.1f * a.depth;
}

TT_DEV Reservoir empty_reservoir() { return Reservoir{0xffffffffu, 0u, 0.0f, 0.0f}; }

// Generalized balance heuristic combination of the pixel's reservoir rq
// (at q) with another pixel's ro (at o); one draw from g.
template <bool TEXTURED>
TT_DEV Reservoir combine(const SceneView& s, const Shading& q, Reservoir rq, const Shading& o, Reservoir ro,
                         Generator& g) {
    float pq_yq = 0.0f, po_yq = 0.0f, pq_yo = 0.0f, po_yo = 0.0f;
    if (rq.light != 0xffffffffu) {
        pq_yq = target<TEXTURED>(s, q, rq.light, rq.bary);
        po_yq = target<TEXTURED>(s, o, rq.light, rq.bary);
    }
    if (ro.light != 0xffffffffu) {
        pq_yo = target<TEXTURED>(s, q, ro.light, ro.bary);
        po_yo = target<TEXTURED>(s, o, ro.light, ro.bary);
    }
    float dq = rq.M * pq_yq + ro.M * po_yq;
    float d_o = rq.M * pq_yo + ro.M * po_yo;
    float mq = dq > 0.0f ? (rq.M * pq_yq) / dq : 0.0f;
    float mo = d_o > 0.0f ? (ro.M * po_yo) / d_o : 0.0f;
    float wq = mq * pq_yq * rq.W;
    float wo = mo * pq_yo * ro.W;
    float sum = wq + wo;
    float u = random_float(g);
    Reservoir r;
    bool take_other = u * sum < wo;
    r.light = take_other ? ro.light : rq.light;
    r.bary = take_other ? ro.bary : rq.bary;
    float p = take_other ? pq_yo : pq_yq;
    r.W = p > 0.0f ? sum / p : 0.0f;
    r.M = rq.M + ro.M;
    if (r.W == 0.0f && r.light == 0xffffffffu) r = empty_reservoir();
    return r;
}

// ------------------------------------------------------------ candidates --
// TEXTURED: the instantiation for scenes with textures (textured materials
// at the primary vertex and in the records, textured light emission).  As
// for shading, the untextured instantiation is held to four waves per SIMD.
template <bool TEXTURED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TEXTURED ? 1 : 4))) void restir_candidates_kernel(tt::DeviceScene d, tt::FrameArgs f, tt::PathBuffers b,
                                                                tt::RestirBuffers r, uint4* out) {
    SceneView s = view(d);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= f.width * f.height) return;
    // Whole-frame renders take the pixels in the primary rays' tile order
    // (tt::primary_slot), so a wave's lanes are an 8 x 8 pixel tile.
    const uint32_t p = f.primary_tiles ? tt::slot_pixel(t, f.width, f.height) : t;
    float4* g = r.gbuffer[r.cur] + RECORD * size_t(p);
    const uint32_t path = r.base + p;  // the pixel's primary path
    const uint32_t qslot = f.primary_tiles ? t : path;
    float4 qd = b.queue[0][1][qslot];
    float4 hit = b.hits[qslot];
    uint32_t instance = __float_as_uint(hit.z);
    Reservoir res = empty_reservoir();
    bool emit = false;
    f3 org, dir;
    float dist = 0.0f;
    LightPoint sel_lp;
    if (instance == 0xffffffffu) {
        g[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
        f3 ray_dir(qd.x, qd.y, qd.z);
        Shading q;
        f3 cam(f.camera_position[0], f.camera_position[1], f.camera_position[2]);
        // The primary vertex as shade 0 builds it; the sample's device
        // coordinates ride in the first ray's and the throughput's w.
        TextureHit tex{TEXTURED, true,
                       PrimaryRay{f.inverse_view, f.inverse_proj, cam, float(f.width), float(f.height), qd.w,
                                  b.throughput[path].w},
                       1.0f};
        q.sf = surface_at_hit(s, instance, __float_as_uint(hit.w), hit.x, hit.y, ray_dir, q.m, tex);
        q.valid = s.emissive_count > 0 && q.m.specular_roughness > 0.15f;
        q.depth = length(q.sf.position - cam);
        write_record(g, q.sf, q.valid, q.depth, -ray_dir, f.frame_index, tex.u, tex.v, tex.lod);
        if (q.valid) {
            Generator lg = restir_lobe_generator(p, f.frame_index);
            q.lobes = bsdf_prepare(q.m, q.sf.wi, lg);
            Generator gen = pass_generator(f, p, 1u);
            float sum = 0.0f, p_sel = 0.0f;
            uint32_t sel_light = 0xffffffffu, sel_bary = 0u;
            for (uint32_t j = 0; j < r.candidates; j++) {
                uint32_t li = random_uint(gen, s.emissive_count);
                f2 r2 = random_vec2(gen);
                uint32_t bary = quantize_bary(sample_triangle(r2));
                float area;
                float ph = target<TEXTURED>(s, q, li, bary, &area);
                float w = ph * (float(s.emissive_count) * area);
                sum = sum + w;
                float u = random_float(gen);
                if (u * sum < w) {
                    sel_light = li;
                    sel_bary = bary;
                    p_sel = ph;
                }
            }
            if (sel_light != 0xffffffffu && p_sel > 0.0f) {
                res.light = sel_light;
                res.bary = sel_bary;
                res.W = (sum / float(r.candidates)) / p_sel;
                res.M = float(r.candidates);
                sel_lp = light_point(s, sel_light, reservoir_bary(sel_bary));
                f3 to_light = sel_lp.position - q.sf.position;
                dist = sqrt_(length_squared(to_light));
                dir = to_light / dist;
                org = q.sf.position;
                emit = true;
            } else {
                res.M = float(r.candidates);
            }
        }
    }
    out[p] = pack_reservoir(res);
    uint32_t slot = wave_append(b.counters + C_RESTIR_VIS, emit);
    if (emit) {
        b.shadow[0][slot] = make_float4(org.x, org.y, org.z, __uint_as_float(p));
        b.shadow[1][slot] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(sel_lp.instance));
        b.shadow[2][slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(sel_lp.hash));
    }
}

// ------------------------------------------------------------ visibility --
// The reference's shadow-ray rule (integrate.glsl:168-185) for each
// candidate winner; occluded samples keep M and get W = 0.
__global__ __launch_bounds__(TRACE_BLOCK) void restir_visibility_kernel(tt::DeviceScene d, tt::PathBuffers b,
                                                                        uint4* res) {
    SceneView s = view(d);
    Stack st = make_stack(b.spill);
    const uint32_t count = b.counters[C_RESTIR_VIS];
    WorkFetch wf = work_fetch(b.work + tt::W_RESTIR * tt::WORK_SET, count);
    uint32_t first, batch;
    while (next_batch(wf, first, batch)) {
        uint32_t i = first + lane_id();
        if (lane_id() < batch) {
            float4 o = b.shadow[0][i], dd = b.shadow[1][i], meta4 = b.shadow[2][i];
            bool visible = light_visible(s, f3(o.x, o.y, o.z), f3(dd.x, dd.y, dd.z), __float_as_uint(dd.w),
                                         __float_as_uint(meta4.w) & 0xffffu, st);
            if (!visible) {
                uint32_t p = __float_as_uint(o.w);
                uint4 v = res[p];
                v.z &= 0xffff0000u;  // W = +0 (half), M kept
                res[p] = v;
            }
        }
    }
}

// -------------------------------------------------------------- temporal --
template <bool TEXTURED>
__global__ __launch_bounds__(256) void restir_temporal_kernel(tt::DeviceScene d, tt::FrameArgs f, tt::RestirBuffers r,
                                                              const uint4* in, uint4* out) {
    SceneView s = view(d);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= f.width * f.height) return;
    const uint32_t p = f.primary_tiles ? tt::slot_pixel(t, f.width, f.height) : t;  // tile order, as candidates
    Reservoir rq = unpack_reservoir(in[p]);
    Shading q = read_record<TEXTURED>(s, r.gbuffer[r.cur] + RECORD * size_t(p), p);
    if (q.valid && r.history_valid) {
        Shading o = read_record<TEXTURED>(s, r.gbuffer[r.cur ^ 1u] + RECORD * size_t(p), p);
        if (similar(q, o)) {
            Reservoir ro = unpack_reservoir(r.final_[r.cur ^ 1u][p]);
            ro.M = fmin_(ro.M, RESTIR_MAX_HISTORY * float(r.candidates));
            Generator gen = pass_generator(f, p, 2u);
            rq = combine<TEXTURED>(s, q, rq, o, ro, gen);
        }
    }
    out[p] = pack_reservoir(rq);
}

// --------------------------------------------------------------- spatial --
template <bool TEXTURED>
__global__ __launch_bounds__(256) void restir_spatial_kernel(tt::DeviceScene d, tt::FrameArgs f, tt::RestirBuffers r,
                                                             const uint4* in, uint4* out) {
    SceneView s = view(d);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= f.width * f.height) return;
    const uint32_t p = f.primary_tiles ? tt::slot_pixel(t, f.width, f.height) : t;  // tile order, as candidates
    Reservoir rq = unpack_reservoir(in[p]);
    Shading q = read_record<TEXTURED>(s, r.gbuffer[r.cur] + RECORD * size_t(p), p);
    if (q.valid) {
        Generator gen = pass_generator(f, p, 3u);
        float angle = 2.0f * PI * random_float(gen);
        float radius = RESTIR_RADIUS * sqrt_(random_float(gen));
        int x = int(p % f.width), y = int(p / f.width);
        int nx = int(floorf(float(x) + radius * ttm::cos(angle) + 0.5f));
        int ny = int(floorf(float(y) + radius * ttm::sin(angle) + 0.5f));
        nx = nx < 0 ? 0 : (nx >= int(f.width) ? int(f.width) - 1 : nx);
        ny = ny < 0 ? 0 : (ny >= int(f.height) ? int(f.height) - 1 : ny);
        uint32_t n = uint32_t(ny) * f.width + uint32_t(nx);
        if (n != p) {
            Shading o = read_record<TEXTURED>(s, r.gbuffer[r.cur] + RECORD * size_t(n), n);
            if (similar(q, o)) rq = combine<TEXTURED>(s, q, rq, o, unpack_reservoir(in[n]), gen);
        }
    }
    out[p] = pack_reservoir(rq);
}

uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

template <bool TEXTURED>
hipError_t restir_passes(hipStream_t s, const tt::DeviceScene& scene, const tt::FrameArgs& f, const tt::PathBuffers& b,
                         const tt::RestirBuffers& r, int grid) {
    uint32_t pixels = f.width * f.height;
    uint32_t nb = blocks(pixels, 256);
    // initial -> (temporal) -> (spatial) -> final_[cur]
    uint4* after_candidates = (r.temporal_reuse || r.spatial_reuse) ? r.initial : r.final_[r.cur];
    restir_candidates_kernel<TEXTURED><<<nb, 256, 0, s>>>(scene, f, b, r, after_candidates);
    restir_visibility_kernel<<<grid, TRACE_BLOCK, 0, s>>>(scene, b, after_candidates);
    const uint4* current = after_candidates;
    if (r.temporal_reuse) {
        uint4* dst = r.spatial_reuse ? r.temporal : r.final_[r.cur];
        restir_temporal_kernel<TEXTURED><<<nb, 256, 0, s>>>(scene, f, r, current, dst);
        current = dst;
    }
    if (r.spatial_reuse) restir_spatial_kernel<TEXTURED><<<nb, 256, 0, s>>>(scene, f, r, current, r.final_[r.cur]);
    return hipGetLastError();
}

}  // namespace

namespace tt {

hipError_t launch_restir(hipStream_t s, const DeviceScene& scene, const FrameArgs& f, const PathBuffers& b,
                         const RestirBuffers& r, int grid) {
    return scene.texture_count > 0 ? restir_passes<true>(s, scene, f, b, r, grid)
                                   : restir_passes<false>(s, scene, f, b, r, grid);
}

}  // namespace tt
