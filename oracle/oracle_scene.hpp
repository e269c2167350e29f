// ORACLE (test infrastructure only; never linked into the product).
//
// Scene access, watertight ray/triangle intersection, a two-level BVH used
// only to make the oracle tractable on large scenes, and the integrator of
// src/integrate/integrate.glsl restated per pixel.
//
// Reference files followed:
//   src/integrate/integrate.glsl:17-20   camera_ray_direction
//   src/integrate/integrate.glsl:22-110  get_ray_hit
//   src/integrate/integrate.glsl:112-132 trace_ray, pixel_ndc
//   src/integrate/integrate.glsl:156-201 direct_light_contribution, create_surface_basis
//   src/integrate/integrate.glsl:270-347 next_path_segment, main
//   src/includes/scene.glsl:29-64,107-136 tangent frames, snorm16, light sampling
//   src/includes/octahedron.glsl:16-23  octahedron_decode
//   src/includes/math.glsl:23-33        transform_to_basis / from_basis, gram_schmidt
//
// Ray queries.  The reference traces with hardware rayQueryEXT whose exact
// arithmetic is not observable.  This restatement fixes it as: the world ray
// is taken to each instance's object space with the stored f32
// inverse_transform (origin: M*(o,1); direction: upper 3x3 of M times d), and
// tested against the snorm16-dequantized object-space triangle with the
// watertight test of Woop, Benthin & Wald (JCGT 2013, including its
// double-precision fallback).  Back-face culling keeps hits where U, V, W >= 0
// (the CCW geometric normal faces the ray origin).  Accepted hits satisfy
// tmin <= t <= tmax; the closest hit is the minimum of (t, instance,
// primitive) in lexicographic order, so the result does not depend on the
// traversal order and equals brute force over all triangles.
#pragma once

#include <algorithm>
#include <vector>

#include "../include/tinytrace.h"
#include "oracle_bsdf.hpp"

namespace oracle {

template <typename R>
inline R dequantize_snorm16(int16_t v) { return R(float(v) / 32767.0f); }

// octahedron.glsl:16-23
template <typename R>
inline V3<R> octahedron_decode(R u, R v) {
    R sx = u * R(2) - R(1), sy = v * R(2) - R(1);
    V3<R> n(sx, sy, R(1) - rabs(sx) - rabs(sy));
    R t = rclamp(-n.z, R(0), R(1));
    n.x = n.x + (n.x >= R(0) ? -t : t);
    n.y = n.y + (n.y >= R(0) ? -t : t);
    return normalize(n);
}

template <typename R>
struct TangentFrame {
    V3<R> normal, tangent;
    R bitangent_sign;
};

// scene.glsl:29-64 (dequantize_unorm: value / (2^bits - 1), in float)
template <typename R>
inline TangentFrame<R> decode_tangent_frame(uint32_t e) {
    R u = R(float(e & 0x3ffu) / float(1023));
    R v = R(float((e >> 10) & 0x3ffu) / float(1023));
    TangentFrame<R> f;
    f.normal = octahedron_decode(u, v);
    V3<R> ortho = rabs(f.normal.x) > rabs(f.normal.z) ? V3<R>(-f.normal.y, f.normal.x, R(0))
                                                       : V3<R>(R(0), -f.normal.z, f.normal.y);
    R angle = R(float((e >> 20) & 0x7ffu) / float(2047)) * K<R>::TAU;
    f.tangent = ortho * Fn<R>::cos(angle) + cross(f.normal, ortho) * Fn<R>::sin(angle);
    f.bitangent_sign = (e >> 31) == 1u ? R(1) : R(-1);
    return f;
}

template <typename R>
struct Basis {
    V3<R> normal, tangent, bitangent;
};

// math.glsl:23-33
template <typename R>
inline V3<R> transform_to_basis(const Basis<R>& b, V3<R> d) {
    return normalize(d.x * b.tangent + d.y * b.bitangent + d.z * b.normal);
}
template <typename R>
inline V3<R> transform_from_basis(const Basis<R>& b, V3<R> d) {
    return normalize(V3<R>(dot(d, b.tangent), dot(d, b.bitangent), dot(d, b.normal)));
}
template <typename R>
inline V3<R> gram_schmidt(V3<R> n, V3<R> t) { return normalize(t - n * dot(n, t)); }

template <typename R>
inline V3<R> interpolate(V3<R> b, V3<R> x, V3<R> y, V3<R> z) { return b.x * x + b.y * y + b.z * z; }

struct Hit {
    float t;
    uint32_t instance;
    uint32_t primitive;
    float b1, b2;
};

// Lexicographic (t, instance, primitive) ordering of candidate hits.
template <typename T>
inline bool closer(T t, uint32_t inst, uint32_t prim, T best_t, uint32_t best_inst, uint32_t best_prim) {
    if (t != best_t) return t < best_t;
    if (inst != best_inst) return inst < best_inst;
    return prim < best_prim;
}

template <typename R>
struct WoopRay {
    V3<R> org, dir;
    int kx, ky, kz;
    R sx, sy, sz;
};

template <typename R>
inline WoopRay<R> woop_setup(V3<R> org, V3<R> dir) {
    WoopRay<R> w;
    w.org = org;
    w.dir = dir;
    R ax = rabs(dir.x), ay = rabs(dir.y), az = rabs(dir.z);
    // Largest component; ties resolve to the lower axis.
    int kz = (ax >= ay) ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (dir[kz] < R(0)) std::swap(kx, ky);
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    w.sx = dir[kx] / dir[kz];
    w.sy = dir[ky] / dir[kz];
    w.sz = R(1) / dir[kz];
    return w;
}

template <typename R>
struct Wide;
template <>
struct Wide<float> {
    using T = double;
};
template <>
struct Wide<double> {
    using T = long double;
};

// Woop, Benthin, Wald 2013, Fig. 6/7.  Returns true and (t, b1, b2) on a hit
// inside [tmin, tmax].
template <typename R>
inline bool intersect_triangle(const WoopRay<R>& r, V3<R> v0, V3<R> v1, V3<R> v2, bool cull, R tmin,
                               R tmax, R& t_out, R& b1, R& b2) {
    V3<R> A = v0 - r.org, B = v1 - r.org, C = v2 - r.org;
    R Ax = A[r.kx] - r.sx * A[r.kz];
    R Ay = A[r.ky] - r.sy * A[r.kz];
    R Bx = B[r.kx] - r.sx * B[r.kz];
    R By = B[r.ky] - r.sy * B[r.kz];
    R Cx = C[r.kx] - r.sx * C[r.kz];
    R Cy = C[r.ky] - r.sy * C[r.kz];
    R U = Cx * By - Cy * Bx;
    R V = Ax * Cy - Ay * Cx;
    R W = Bx * Ay - By * Ax;
    if (U == R(0) || V == R(0) || W == R(0)) {
        using D = typename Wide<R>::T;
        U = R(D(Cx) * D(By) - D(Cy) * D(Bx));
        V = R(D(Ax) * D(Cy) - D(Ay) * D(Cx));
        W = R(D(Bx) * D(Ay) - D(By) * D(Ax));
    }
    if (cull) {
        if (U < R(0) || V < R(0) || W < R(0)) return false;
    } else {
        if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    }
    R det = U + V + W;
    if (det == R(0)) return false;
    R Az = r.sz * A[r.kz];
    R Bz = r.sz * B[r.kz];
    R Cz = r.sz * C[r.kz];
    R T = U * Az + V * Bz + W * Cz;
    R t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    b1 = V / det;
    b2 = W / det;
    return true;
}

// ------------------------------------------------------------- the scene --
template <typename R>
struct Instance {
    M4<R> transform, inverse_transform, normal_transform;
    uint32_t vertex_offset, index_offset, material_index, mesh_index;
};

template <typename R>
struct Aabb {
    R lo[3], hi[3];
    void reset() {
        for (int i = 0; i < 3; i++) {
            lo[i] = R(1e30);
            hi[i] = R(-1e30);
        }
    }
    void grow(V3<R> p) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(lo[i], p[i]);
            hi[i] = std::max(hi[i], p[i]);
        }
    }
    void grow(const Aabb& b) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(lo[i], b.lo[i]);
            hi[i] = std::max(hi[i], b.hi[i]);
        }
    }
    R area() const {
        R dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0) return 0;
        return R(2) * (dx * dy + dy * dz + dz * dx);
    }
};

// A plain binned-SAH BVH2 (oracle-internal).  Box tests are made
// conservative (boxes padded by 2^-20 relative + per-ray slack proportional to
// the ray origin's magnitude, slab distances widened by Ize 2013's 1+2*gamma3
// factor) so that they never reject a triangle that the exact test above
// accepts; the result therefore equals brute force.
template <typename R>
struct Bvh {
    struct Node {
        Aabb<R> box;
        uint32_t left_or_first;  // inner: left child; leaf: first primitive
        uint32_t count;          // 0 = inner (right child = left + 1)
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> prims;

    void build(const std::vector<Aabb<R>>& boxes) {
        size_t n = boxes.size();
        prims.resize(n);
        for (size_t i = 0; i < n; i++) prims[i] = uint32_t(i);
        nodes.clear();
        nodes.reserve(2 * n + 1);
        std::vector<V3<R>> centroids(n);
        for (size_t i = 0; i < n; i++)
            centroids[i] = V3<R>((boxes[i].lo[0] + boxes[i].hi[0]) * R(0.5), (boxes[i].lo[1] + boxes[i].hi[1]) * R(0.5),
                                 (boxes[i].lo[2] + boxes[i].hi[2]) * R(0.5));
        nodes.push_back(Node{});
        if (n == 0) {
            nodes[0].box.reset();
            nodes[0].left_or_first = 0;
            nodes[0].count = 0xffffffffu;  // empty
            return;
        }
        struct Task {
            uint32_t node, first, count;
        };
        std::vector<Task> stack{{0, 0, uint32_t(n)}};
        while (!stack.empty()) {
            Task task = stack.back();
            stack.pop_back();
            Aabb<R> box, cbox;
            box.reset();
            cbox.reset();
            for (uint32_t i = task.first; i < task.first + task.count; i++) {
                box.grow(boxes[prims[i]]);
                cbox.grow(centroids[prims[i]]);
            }
            // Pad: relative 2^-20 of the coordinate magnitude.
            for (int a = 0; a < 3; a++) {
                R pad_lo = (rabs(box.lo[a]) + R(1e-30)) * R(9.5367431640625e-07);
                R pad_hi = (rabs(box.hi[a]) + R(1e-30)) * R(9.5367431640625e-07);
                box.lo[a] = box.lo[a] - pad_lo;
                box.hi[a] = box.hi[a] + pad_hi;
            }
            nodes[task.node].box = box;
            if (task.count <= 2) {
                nodes[task.node].left_or_first = task.first;
                nodes[task.node].count = task.count;
                continue;
            }
            const int BINS = 12;
            R best_cost = R(1e30);
            int best_axis = -1, best_split = 0;
            for (int a = 0; a < 3; a++) {
                R extent = cbox.hi[a] - cbox.lo[a];
                if (!(extent > R(0))) continue;
                Aabb<R> bin_box[BINS];
                uint32_t bin_count[BINS] = {};
                for (int b = 0; b < BINS; b++) bin_box[b].reset();
                for (uint32_t i = task.first; i < task.first + task.count; i++) {
                    int b = int((centroids[prims[i]][a] - cbox.lo[a]) / extent * R(BINS));
                    b = std::min(std::max(b, 0), BINS - 1);
                    bin_box[b].grow(boxes[prims[i]]);
                    bin_count[b]++;
                }
                R right_area[BINS];
                uint32_t right_count[BINS];
                Aabb<R> acc;
                acc.reset();
                uint32_t cnt = 0;
                for (int b = BINS - 1; b > 0; b--) {
                    acc.grow(bin_box[b]);
                    cnt += bin_count[b];
                    right_area[b] = acc.area();
                    right_count[b] = cnt;
                }
                acc.reset();
                cnt = 0;
                for (int b = 0; b < BINS - 1; b++) {
                    acc.grow(bin_box[b]);
                    cnt += bin_count[b];
                    if (cnt == 0 || right_count[b + 1] == 0) continue;
                    R cost = acc.area() * R(cnt) + right_area[b + 1] * R(right_count[b + 1]);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_split = b;
                    }
                }
            }
            uint32_t mid;
            if (best_axis < 0) {
                mid = task.first + task.count / 2;
            } else {
                R extent = cbox.hi[best_axis] - cbox.lo[best_axis];
                auto it = std::partition(prims.begin() + task.first, prims.begin() + task.first + task.count,
                                         [&](uint32_t p) {
                                             int b = int((centroids[p][best_axis] - cbox.lo[best_axis]) / extent * R(BINS));
                                             b = std::min(std::max(b, 0), BINS - 1);
                                             return b <= best_split;
                                         });
                mid = uint32_t(it - prims.begin());
                if (mid == task.first || mid == task.first + task.count) mid = task.first + task.count / 2;
            }
            uint32_t left = uint32_t(nodes.size());
            nodes.push_back(Node{});
            nodes.push_back(Node{});
            nodes[task.node].left_or_first = left;
            nodes[task.node].count = 0;
            stack.push_back({left, task.first, mid - task.first});
            stack.push_back({left + 1, mid, task.first + task.count - mid});
        }
    }
};

// Per-ray data for conservative slab tests.
template <typename R>
struct SlabRay {
    V3<R> org, inv, slack;
};

template <typename R>
inline SlabRay<R> slab_setup(V3<R> org, V3<R> dir) {
    SlabRay<R> s;
    s.org = org;
    s.inv = V3<R>(R(1) / dir.x, R(1) / dir.y, R(1) / dir.z);
    R mo = std::max(rabs(org.x), std::max(rabs(org.y), rabs(org.z)));
    R e = mo * R(64) * Fn<R>::epsilon();
    s.slack = V3<R>(e * rabs(s.inv.x), e * rabs(s.inv.y), e * rabs(s.inv.z));
    return s;
}

template <typename R>
inline bool slab_test(const SlabRay<R>& s, const Aabb<R>& b, R tmin, R tmax) {
    const R g3 = R(3) * Fn<R>::epsilon() / (R(1) - R(3) * Fn<R>::epsilon());
    const R up = R(1) + R(2) * g3;
    R t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; a++) {
        R tn = (b.lo[a] - s.org[a]) * s.inv[a];
        R tf = (b.hi[a] - s.org[a]) * s.inv[a];
        if (tn > tf) std::swap(tn, tf);
        if (tn != tn) tn = R(-1e30);  // 0 * inf
        if (tf != tf) tf = R(1e30);
        tf = tf * up + s.slack[a];
        tn = tn - rabs(tn) * R(2) * g3 - s.slack[a];
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    return true;
}

template <typename R>
struct Scene {
    const tt_scene_desc* desc = nullptr;
    std::vector<Instance<R>> instances;
    std::vector<Material<R>> materials;
    // Object-space dequantized positions per vertex.
    std::vector<V3<R>> positions;
    // BLAS per mesh, TLAS over instances.
    std::vector<Bvh<R>> blas;
    Bvh<R> tlas;
    bool use_bvh = true;
    // sRGB code -> linear (c <= 0.04045 ? c / 12.92 : ((c + 0.055) / 1.055)^2.4).
    float srgb[256];

    void load(const tt_scene_desc* d, bool build_bvh) {
        desc = d;
        instances.resize(d->instance_count);
        for (uint32_t i = 0; i < d->instance_count; i++) {
            const tt_instance& s = d->instances[i];
            Instance<R>& o = instances[i];
            o.transform = load_m4<R>(s.transform);
            o.inverse_transform = load_m4<R>(s.inverse_transform);
            o.normal_transform = load_m4<R>(s.normal_transform);
            o.vertex_offset = s.vertex_offset;
            o.index_offset = s.index_offset;
            o.material_index = s.material_index;
            o.mesh_index = s.mesh_index;
        }
        materials.resize(d->material_count);
        for (uint32_t i = 0; i < d->material_count; i++)
            materials[i] = load_material<R>(&d->materials[i].constants.base_color[0]);
        positions.resize(d->position_count);
        for (uint32_t i = 0; i < d->position_count; i++)
            positions[i] = V3<R>(dequantize_snorm16<R>(d->positions[3 * i]), dequantize_snorm16<R>(d->positions[3 * i + 1]),
                                 dequantize_snorm16<R>(d->positions[3 * i + 2]));
        use_bvh = build_bvh;
        if (build_bvh) build_acceleration();
        for (int i = 0; i < 256; i++) {
            double c = double(i) / 255.0;
            srgb[i] = float(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
        }
    }

    bool textured() const { return desc->texture_count > 0; }

    // One texel (x, y) of texture level `data`, w texels wide, as the
    // sampler returns it: BC1 and RGBA8 through the sRGB table, BC5, RG8 and
    // R8 as unorm, R16 as half.
    void texel(uint32_t format, const uint8_t* data, uint32_t w, uint32_t x, uint32_t y, R out[4]) const {
        const size_t i = size_t(y) * w + x;
        if (format == TT_TEXTURE_FORMAT_BC1_SRGB || format == TT_TEXTURE_FORMAT_BC5_UNORM) {
            uint8_t c[4];
            block_texel(format, data, w, x, y, c);
            for (int k = 0; k < 3; k++)
                out[k] = format == TT_TEXTURE_FORMAT_BC1_SRGB ? R(srgb[c[k]]) : R(float(c[k]) / 255.0f);
            out[3] = R(float(c[3]) / 255.0f);
        } else if (format == TT_TEXTURE_FORMAT_RGBA8_SRGB) {
            const uint8_t* t = data + 4 * size_t(i);
            out[0] = R(srgb[t[0]]);
            out[1] = R(srgb[t[1]]);
            out[2] = R(srgb[t[2]]);
            out[3] = R(float(t[3]) / 255.0f);
        } else if (format == TT_TEXTURE_FORMAT_RG8_UNORM) {
            out[0] = R(float(data[2 * size_t(i)]) / 255.0f);
            out[1] = R(float(data[2 * size_t(i) + 1]) / 255.0f);
            out[2] = R(0);
            out[3] = R(1);
        } else if (format == TT_TEXTURE_FORMAT_R8_UNORM) {
            out[0] = R(float(data[i]) / 255.0f);
            out[1] = out[2] = R(0);
            out[3] = R(1);
        } else {
            uint16_t h = uint16_t(data[2 * size_t(i)] | (data[2 * size_t(i) + 1] << 8));
            out[0] = R(ttm::half_to_float(h));
            out[1] = out[2] = R(0);
            out[3] = R(1);
        }
    }

    // The 8-bit channels of texel (x, y) of a BC1 (RGB, alpha) or BC5 (R, G,
    // 0, 255) level, decoded as the block formats define them.
    static void block_texel(uint32_t format, const uint8_t* data, uint32_t w, uint32_t x, uint32_t y, uint8_t c[4]) {
        const size_t bytes = format == TT_TEXTURE_FORMAT_BC1_SRGB ? 8 : 16;
        const uint8_t* b = data + bytes * ((size_t(y) / 4) * ((w + 3) / 4) + x / 4);
        const uint32_t t = (y % 4) * 4 + (x % 4);  // texel number in the block
        if (format == TT_TEXTURE_FORMAT_BC1_SRGB) {
            const uint32_t e[2] = {uint32_t(b[0]) | uint32_t(b[1]) << 8, uint32_t(b[2]) | uint32_t(b[3]) << 8};
            const uint32_t index = (b[4 + t / 4] >> (2 * (t % 4))) & 3;
            uint32_t rgb[2][3];
            for (int j = 0; j < 2; j++) {
                uint32_t r5 = e[j] >> 11, g6 = (e[j] >> 5) & 63, b5 = e[j] & 31;
                rgb[j][0] = r5 * 8 + r5 / 4;
                rgb[j][1] = g6 * 4 + g6 / 16;
                rgb[j][2] = b5 * 8 + b5 / 4;
            }
            c[3] = 255;
            for (int k = 0; k < 3; k++) {
                uint32_t a = rgb[0][k], z = rgb[1][k];
                uint32_t v;
                if (index == 0) v = a;
                else if (index == 1) v = z;
                else if (e[0] > e[1]) v = index == 2 ? (2 * a + z + 1) / 3 : (a + 2 * z + 1) / 3;
                else v = index == 2 ? (a + z + 1) / 2 : 0;
                c[k] = uint8_t(v);
            }
            if (index == 3 && !(e[0] > e[1])) c[3] = 0;
            return;
        }
        for (int ch = 0; ch < 2; ch++) {
            const uint8_t* h = b + 8 * ch;
            const uint32_t e0 = h[0], e1 = h[1];
            uint64_t bits = 0;
            for (int k = 0; k < 6; k++) bits |= uint64_t(h[2 + k]) << (8 * k);
            const uint32_t index = uint32_t(bits >> (3 * t)) & 7;
            uint32_t v;
            if (index == 0) v = e0;
            else if (index == 1) v = e1;
            else if (e0 > e1) v = ((8 - index) * e0 + (index - 1) * e1 + 3) / 7;
            else if (index < 6) v = ((6 - index) * e0 + (index - 1) * e1 + 2) / 5;
            else v = index == 6 ? 0 : 255;
            c[ch] = uint8_t(v);
        }
        c[2] = 0;
        c[3] = 255;
    }

    // Repeat addressing of one axis: lower and upper texel, weight of upper.
    static void wrap(R c, uint32_t n, uint32_t& i0, uint32_t& i1, R& a) {
        R f = c - std::floor(c);
        if (!(f >= R(0) && f < R(1))) f = R(0);
        R x = f * R(float(n)) - R(0.5);
        R fx = std::floor(x);
        a = x - fx;
        int i = int(fx);
        i0 = i < 0 ? n - 1u : uint32_t(i);
        i1 = i0 + 1u == n ? 0u : i0 + 1u;
    }

    void bilinear(uint32_t format, const uint8_t* data, uint32_t w, uint32_t h, R u, R v, R out[4]) const {
        uint32_t x0, x1, y0, y1;
        R ax, ay;
        wrap(u, w, x0, x1, ax);
        wrap(v, h, y0, y1, ay);
        R c00[4], c10[4], c01[4], c11[4];
        texel(format, data, w, x0, y0, c00);
        texel(format, data, w, x1, y0, c10);
        texel(format, data, w, x0, y1, c01);
        texel(format, data, w, x1, y1, c11);
        for (int k = 0; k < 4; k++) {
            R top = c00[k] * (R(1) - ax) + c10[k] * ax;
            R bottom = c01[k] * (R(1) - ax) + c11[k] * ax;
            out[k] = top * (R(1) - ay) + bottom * ay;
        }
    }

    // textureLod with linear filtering between the two levels around the
    // clamped level of detail (the sampler of scene.rs:121-128 without its
    // anisotropic filtering).
    void sample(uint32_t index, R u, R v, R lod, R out[4]) const {
        const tt_texture_record& r = desc->textures[index];
        R top = R(float(r.levels - 1u));
        R l = lod > R(0) ? (lod < top ? lod : top) : R(0);
        uint32_t l0 = uint32_t(std::floor(l));
        uint32_t l1 = l0 + 1u < r.levels ? l0 + 1u : r.levels - 1u;
        R t = l - R(float(l0));
        R a[4], b[4];
        bilinear(r.format, desc->texture_data + r.offsets[l0], r.width >> l0, r.height >> l0, u, v, a);
        bilinear(r.format, desc->texture_data + r.offsets[l1], r.width >> l1, r.height >> l1, u, v, b);
        for (int k = 0; k < 4; k++) out[k] = a[k] * (R(1) - t) + b[k] * t;
    }

    // integrate.glsl:219-267: the material constants with its textured
    // channels replaced (the channels load_textures reads, in its order).
    Material<R> material_at(uint32_t index, R u, R v, R lod) const {
        Material<R> m = materials[index];
        if (!textured()) return m;
        const uint32_t* t = desc->materials[index].textures;
        R c[4];
        auto fetch = [&](int ch) {
            if (t[ch] == TT_INVALID_INDEX) return false;
            sample(t[ch], u, v, lod, c);
            return true;
        };
        if (fetch(TT_TEXTURE_BASE_COLOR)) m.base_color = V3<R>(c[0], c[1], c[2]);
        if (fetch(TT_TEXTURE_SPECULAR_COLOR)) m.specular_color = V3<R>(c[0], c[1], c[2]);
        if (fetch(TT_TEXTURE_TRANSMISSION_COLOR)) m.transmission_color = V3<R>(c[0], c[1], c[2]);
        if (fetch(TT_TEXTURE_COAT_COLOR)) m.coat_color = V3<R>(c[0], c[1], c[2]);
        if (fetch(TT_TEXTURE_FUZZ_COLOR)) m.fuzz_color = V3<R>(c[0], c[1], c[2]);
        if (fetch(TT_TEXTURE_GEOMETRY_NORMAL)) m.geometry_normal = octahedron_decode(c[0], c[1]);
        if (fetch(TT_TEXTURE_BASE_METALNESS)) m.base_metalness = c[0];
        if (fetch(TT_TEXTURE_BASE_DIFFUSE_ROUGHNESS)) m.base_diffuse_roughness = c[0];
        if (fetch(TT_TEXTURE_SPECULAR_WEIGHT)) m.specular_weight = c[0];
        if (fetch(TT_TEXTURE_SPECULAR_ROUGHNESS)) m.specular_roughness = c[0];
        if (fetch(TT_TEXTURE_SPECULAR_ROUGHNESS_ANISOTROPY)) m.specular_roughness_anisotropy = c[0];
        if (fetch(TT_TEXTURE_SPECULAR_IOR)) m.specular_ior = c[0];
        if (fetch(TT_TEXTURE_SPECULAR_ROTATION)) m.specular_rotation = c[0];
        if (fetch(TT_TEXTURE_TRANSMISSION_WEIGHT)) m.transmission_weight = c[0];
        if (fetch(TT_TEXTURE_COAT_WEIGHT)) m.coat_weight = c[0];
        if (fetch(TT_TEXTURE_COAT_DARKENING)) m.coat_darkening = c[0];
        if (fetch(TT_TEXTURE_COAT_ROTATION)) m.coat_rotation = c[0];
        if (fetch(TT_TEXTURE_FUZZ_WEIGHT)) m.fuzz_weight = c[0];
        if (fetch(TT_TEXTURE_FUZZ_ROUGHNESS)) m.fuzz_roughness = c[0];
        if (fetch(TT_TEXTURE_GEOMETRY_OPACITY)) m.geometry_opacity = c[0];
        return m;
    }

    // material_emissive (integrate.glsl:148-154): emission colour x
    // luminance, each from its texture where the material has one.
    V3<R> emissive_at(uint32_t index, R u, R v, R lod) const {
        const Material<R>& m = materials[index];
        V3<R> color = m.emission_color;
        R luminance = m.emission_luminance;
        if (textured()) {
            const uint32_t* t = desc->materials[index].textures;
            R c[4];
            if (t[TT_TEXTURE_EMISSION_COLOR] != TT_INVALID_INDEX) {
                sample(t[TT_TEXTURE_EMISSION_COLOR], u, v, lod, c);
                color = V3<R>(c[0], c[1], c[2]);
            }
            if (t[TT_TEXTURE_EMISSION_LUMINANCE] != TT_INVALID_INDEX) {
                sample(t[TT_TEXTURE_EMISSION_LUMINANCE], u, v, lod, c);
                luminance = c[0];
            }
        }
        return color * luminance;
    }

    V3<R> vertex_position(uint32_t mesh, uint32_t prim, int corner) const {
        const tt_mesh_record& m = desc->meshes[mesh];
        uint32_t index = desc->indices[m.index_offset + 3 * prim + corner];
        return positions[m.vertex_offset + index];
    }

    void build_acceleration() {
        blas.resize(desc->mesh_count);
        for (uint32_t m = 0; m < desc->mesh_count; m++) {
            uint32_t tris = desc->meshes[m].index_count / 3;
            std::vector<Aabb<R>> boxes(tris);
            for (uint32_t p = 0; p < tris; p++) {
                boxes[p].reset();
                for (int c = 0; c < 3; c++) boxes[p].grow(vertex_position(m, p, c));
            }
            blas[m].build(boxes);
        }
        std::vector<Aabb<R>> boxes(instances.size());
        for (size_t i = 0; i < instances.size(); i++) {
            const Aabb<R>& ob = blas[instances[i].mesh_index].nodes[0].box;
            boxes[i].reset();
            if (ob.lo[0] > ob.hi[0]) continue;  // empty mesh
            // Transform the 8 corners in double, then pad generously: the
            // object-space ray is a rounded image of the world ray.
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            const tt_instance& s = desc->instances[i];
            for (int corner = 0; corner < 8; corner++) {
                double p[3] = {double((corner & 1) ? ob.hi[0] : ob.lo[0]), double((corner & 2) ? ob.hi[1] : ob.lo[1]),
                               double((corner & 4) ? ob.hi[2] : ob.lo[2])};
                for (int r = 0; r < 3; r++) {
                    double w = double(s.transform[12 + r]);
                    for (int c = 0; c < 3; c++) w += double(s.transform[c * 4 + r]) * p[c];
                    lo[r] = std::min(lo[r], w);
                    hi[r] = std::max(hi[r], w);
                }
            }
            double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
            for (int r = 0; r < 3; r++) {
                double pad = (std::fabs(lo[r]) + std::fabs(hi[r]) + ext) * 1e-5;
                boxes[i].lo[r] = R(lo[r] - pad);
                boxes[i].hi[r] = R(hi[r] + pad);
            }
        }
        tlas.build(boxes);
    }

    // Closest hit within one instance: brute force or BLAS traversal.
    void intersect_instance(uint32_t inst, V3<R> wo, V3<R> wd, bool cull, R tmin, R tmax, Hit& best, R& best_t,
                            bool& found) const {
        const Instance<R>& in = instances[inst];
        V3<R> o = mul_point(in.inverse_transform, wo, R(1));
        V3<R> d = mul_dir(in.inverse_transform, wd);
        WoopRay<R> w = woop_setup(o, d);
        auto test = [&](uint32_t p) {
            R t, b1, b2;
            V3<R> v0 = vertex_position(in.mesh_index, p, 0), v1 = vertex_position(in.mesh_index, p, 1),
                  v2 = vertex_position(in.mesh_index, p, 2);
            if (!intersect_triangle(w, v0, v1, v2, cull, tmin, tmax, t, b1, b2)) return;
            if (!found || closer(t, inst, p, best_t, best.instance, best.primitive)) {
                found = true;
                best_t = t;
                best.t = float(t);
                best.instance = inst;
                best.primitive = p;
                best.b1 = float(b1);
                best.b2 = float(b2);
            }
        };
        uint32_t tris = desc->meshes[in.mesh_index].index_count / 3;
        if (!use_bvh) {
            for (uint32_t p = 0; p < tris; p++) test(p);
            return;
        }
        const Bvh<R>& bv = blas[in.mesh_index];
        if (bv.nodes.empty() || bv.nodes[0].count == 0xffffffffu) return;
        SlabRay<R> s = slab_setup(o, d);
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const typename Bvh<R>::Node& node = bv.nodes[stack[--sp]];
            if (!slab_test(s, node.box, tmin, found ? best_t : tmax)) continue;
            if (node.count > 0) {
                for (uint32_t k = 0; k < node.count; k++) test(bv.prims[node.left_or_first + k]);
            } else {
                stack[sp++] = node.left_or_first + 1;
                stack[sp++] = node.left_or_first;
            }
        }
    }

    bool closest_hit(V3<R> org, V3<R> dir, bool cull, R tmin, R tmax, Hit& hit) const {
        bool found = false;
        R best_t = tmax;
        if (!use_bvh) {
            for (uint32_t i = 0; i < instances.size(); i++)
                intersect_instance(i, org, dir, cull, tmin, tmax, hit, best_t, found);
            return found;
        }
        if (tlas.nodes.empty() || tlas.nodes[0].count == 0xffffffffu) return false;
        SlabRay<R> s = slab_setup(org, dir);
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const typename Bvh<R>::Node& node = tlas.nodes[stack[--sp]];
            if (!slab_test(s, node.box, tmin, found ? best_t : tmax)) continue;
            if (node.count > 0) {
                for (uint32_t k = 0; k < node.count; k++)
                    intersect_instance(tlas.prims[node.left_or_first + k], org, dir, cull, tmin, tmax, hit, best_t, found);
            } else {
                stack[sp++] = node.left_or_first + 1;
                stack[sp++] = node.left_or_first;
            }
        }
        return found;
    }
};

// --------------------------------------------------------- the integrator --
struct Settings {
    uint32_t rng;
    uint32_t seed;
};

template <typename R>
struct Camera {
    M4<R> inverse_view, inverse_proj;
    V3<R> position;
    uint32_t width, height;
};

// integrate.glsl:17-20
template <typename R>
inline V3<R> camera_ray_direction(const Camera<R>& c, R ndc_x, R ndc_y) {
    V3<R> v;
    for (int i = 0; i < 3; i++)
        v[i] = c.inverse_proj.c[0][i] * (-ndc_x) + c.inverse_proj.c[1][i] * (-ndc_y) + c.inverse_proj.c[2][i] * R(1) +
            c.inverse_proj.c[3][i] * R(1);
    return normalize(mul_point(c.inverse_view, v, R(0)));
}

template <typename R>
struct RayHit {
    V3<R> world_position;
    Basis<R> tangent_space;
    uint32_t instance_index, material_index;
    // Texture coordinates of the vertices and the hit, world vertices and
    // barycentrics (for the first hit's ray differentials).
    R u[3], v[3], tu, tv;
    V3<R> p[3], bary;
};

// integrate.glsl:22-110.
template <typename R>
inline RayHit<R> get_ray_hit(const Scene<R>& s, const Hit& h) {
    RayHit<R> hit;
    hit.instance_index = h.instance;
    const Instance<R>& in = s.instances[h.instance];
    hit.material_index = in.material_index;
    V3<R> p[3];
    for (int i = 0; i < 3; i++) p[i] = mul_point(in.transform, s.vertex_position(in.mesh_index, h.primitive, i), R(1));
    uint32_t base = 3 * h.primitive + in.index_offset;
    TangentFrame<R> f[3];
    for (int i = 0; i < 3; i++) {
        uint32_t vi = s.desc->indices[base + i] + in.vertex_offset;
        f[i] = decode_tangent_frame<R>(s.desc->vertices[vi].tangent_frame);
    }
    V3<R> bary(R(0), R(h.b1), R(h.b2));
    bary.x = R(1) - bary.y - bary.z;
    hit.tangent_space.normal = normalize(mul_dir(in.normal_transform, interpolate(bary, f[0].normal, f[1].normal, f[2].normal)));
    hit.tangent_space.tangent =
        normalize(mul_dir(in.normal_transform, interpolate(bary, f[0].tangent, f[1].tangent, f[2].tangent)));
    R sign = rsign(f[0].bitangent_sign * bary.x + f[1].bitangent_sign * bary.y + f[2].bitangent_sign * bary.z);
    hit.tangent_space.bitangent = normalize(sign * cross(hit.tangent_space.normal, hit.tangent_space.tangent));
    hit.world_position = bary.x * p[0] + bary.y * p[1] + bary.z * p[2];
    for (int i = 0; i < 3; i++) {
        uint32_t uv = s.desc->vertices[s.desc->indices[base + i] + in.vertex_offset].texture_coordinates[0] |
            uint32_t(s.desc->vertices[s.desc->indices[base + i] + in.vertex_offset].texture_coordinates[1]) << 16;
        hit.u[i] = R(ttm::half_to_float(uint16_t(uv & 0xffffu)));
        hit.v[i] = R(ttm::half_to_float(uint16_t(uv >> 16)));
        hit.p[i] = p[i];
    }
    hit.tu = bary.x * hit.u[0] + bary.y * hit.u[1] + bary.z * hit.u[2];
    hit.tv = bary.x * hit.v[0] + bary.y * hit.v[1] + bary.z * hit.v[2];
    hit.bary = bary;
    return hit;
}

// math.glsl:113-125: barycentrics of a ray's crossing of the triangle plane.
template <typename R>
inline V3<R> triangle_intersection(const V3<R> p[3], V3<R> origin, V3<R> direction) {
    V3<R> edge_to_origin = origin - p[0];
    V3<R> edge_2 = p[2] - p[0];
    V3<R> edge_1 = p[1] - p[0];
    V3<R> r = cross(direction, edge_2);
    V3<R> q = cross(edge_to_origin, edge_1);
    R inverse_det = R(1) / dot(r, edge_1);
    R v1 = dot(r, edge_to_origin);
    R v2 = dot(q, direction);
    R b = v1 * inverse_det;
    R c = v2 * inverse_det;
    return V3<R>(R(1) - b - c, b, c);
}

// integrate.glsl:195-201
template <typename R>
inline Basis<R> create_surface_basis(const Basis<R>& ts, V3<R> n) {
    Basis<R> b;
    b.normal = transform_to_basis(ts, n);
    b.tangent = gram_schmidt(b.normal, ts.tangent);
    b.bitangent = normalize(cross(b.normal, b.tangent));
    return b;
}

struct TraceCounts {
    uint64_t extension = 0, shadow = 0;
};

// scene.glsl:107-136 + integrate.glsl:156-193
template <typename R>
inline V3<R> direct_light_contribution(const Scene<R>& s, V3<R> position, V3<R> wi, const Basis<R>& basis,
                                       const Material<R>& m, const Lobes<R>& lobes, Generator& g, TraceCounts& counts) {
    const tt_scene_desc* d = s.desc;
    const tt_emissive_triangle& tri = d->emissive_triangles[random_uint(g, d->emissive_triangle_count)];
    V3<R> bary = sample_triangle(random_vec2<R>(g));
    const Instance<R>& li = s.instances[tri.instance_index];
    V3<R> lp[3];
    for (int i = 0; i < 3; i++)
        lp[i] = mul_point(li.transform,
                          V3<R>(dequantize_snorm16<R>(tri.positions[i][0]), dequantize_snorm16<R>(tri.positions[i][1]),
                                dequantize_snorm16<R>(tri.positions[i][2])),
                          R(1));
    R area = rmax(R(0.00001), R(0.5) * length(cross(lp[1] - lp[0], lp[2] - lp[0])));
    V3<R> normal = normalize(cross(lp[1] - lp[0], lp[2] - lp[0]));
    V3<R> light_position = interpolate(bary, lp[0], lp[1], lp[2]);
    R light_probability = R(1) / R(float(d->emissive_triangle_count));

    V3<R> to_light = light_position - position;
    V3<R> wo = transform_from_basis(basis, to_light);
    R distance_squared = length_squared(to_light);
    V3<R> dir = to_light / Fn<R>::sqrt(distance_squared);
    R density = (R(1) / area) * distance_squared / saturate(dot(-dir, normal)) * light_probability;
    Hit h;
    counts.shadow++;
    if (!s.closest_hit(position, dir, false, R(1.0e-3), R(10000.0), h)) return V3<R>();
    if (h.instance != tri.instance_index) return V3<R>();
    uint32_t triangle_index = h.primitive + d->instances[h.instance].index_offset / 3;
    if (triangle_index % 0xffffu != uint32_t(tri.hash)) return V3<R>();
    // The sample's texture coordinate (scene.glsl:130-133) and
    // material_emissive there at level 4 (integrate.glsl:189).
    R tc[2];
    for (int k = 0; k < 2; k++)
        tc[k] = bary.x * R(ttm::half_to_float(tri.texture_coordinates[0][k])) +
            bary.y * R(ttm::half_to_float(tri.texture_coordinates[1][k])) +
            bary.z * R(ttm::half_to_float(tri.texture_coordinates[2][k]));
    V3<R> emissive = s.emissive_at(li.material_index, tc[0], tc[1], R(4));
    R bsdf_density;
    V3<R> bsdf = bsdf_evaluate(m, lobes, wi, wo, bsdf_density);
    return bsdf * emissive * rabs(cos_theta(wo)) / rmax(density, K<R>::PDF_EPSILON);
}

// integrate.glsl:203-217 + 270-321, for one sample of one pixel.
// calculate_mip_level (integrate.glsl:134-139) of the first hit from ray
// differentials one pixel over in x and in y (integrate.glsl:89-107).
template <typename R>
inline R first_hit_lod(const Camera<R>& c, const RayHit<R>& hit, R ndc_x, R ndc_y) {
    R w = R(float(c.width)), hgt = R(float(c.height));
    R tsx = R(2) / w, tsy = R(2) / hgt;
    V3<R> hx = triangle_intersection(hit.p, c.position, camera_ray_direction(c, ndc_x + tsx, ndc_y));
    V3<R> hy = triangle_intersection(hit.p, c.position, camera_ray_direction(c, ndc_x, ndc_y + tsy));
    V3<R> dx = hit.bary - hx, dy = hit.bary - hy;
    R dxu = (dx.x * hit.u[0] + dx.y * hit.u[1] + dx.z * hit.u[2]) * w;
    R dxv = (dx.x * hit.v[0] + dx.y * hit.v[1] + dx.z * hit.v[2]) * hgt;
    R dyu = (dy.x * hit.u[0] + dy.y * hit.u[1] + dy.z * hit.u[2]) * w;
    R dyv = (dy.x * hit.v[0] + dy.y * hit.v[1] + dy.z * hit.v[2]) * hgt;
    R gx = dxu * dxu + dxv * dxv, gy = dyu * dyu + dyv * dyv;
    return R(0.5) * Fn<R>::log2(gx < gy ? gy : gx);
}

template <typename R>
inline V3<R> trace_path(const Scene<R>& s, const Camera<R>& c, uint32_t light_sampling, uint32_t bounce_count,
                        V3<R> dir, Generator g, TraceCounts& counts, R ndc_x = R(0), R ndc_y = R(0)) {
    V3<R> accumulated;
    V3<R> throughput(R(1));
    V3<R> ray_dir = dir, ray_org = c.position;
    R lod = R(1);  // PathState.mip_level (create_path_state)
    for (uint32_t bounce = 0; bounce < bounce_count; bounce++) {
        Hit h;
        counts.extension++;
        if (!s.closest_hit(ray_org, ray_dir, true, R(1.0e-3), R(1000.0), h)) {
            accumulated = accumulated + V3<R>(R(1)) * throughput;
            break;
        }
        RayHit<R> hit = get_ray_hit(s, h);
        if (bounce == 0 && s.textured()) lod = first_hit_lod(c, hit, ndc_x, ndc_y);
        const Material<R> m = s.material_at(hit.material_index, hit.tu, hit.tv, lod);
        Basis<R> basis = create_surface_basis(hit.tangent_space, m.geometry_normal);
        V3<R> wi = transform_from_basis(basis, -ray_dir);
        Lobes<R> lobes = bsdf_prepare(m, wi, g);
        bool nee = light_sampling == TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION && s.desc->emissive_triangle_count > 0 &&
            m.specular_roughness > R(0.15);
        if (nee) {
            V3<R> direct = direct_light_contribution(s, hit.world_position, wi, basis, m, lobes, g, counts);
            accumulated = accumulated + throughput * direct;
        }
        if (bounce == 0 || !nee) {
            V3<R> emissive = s.emissive_at(hit.material_index, hit.tu, hit.tv, lod);
            accumulated = accumulated + emissive * throughput;
        }
        if (bounce + 1 == bounce_count) break;  // the continuation ray would never be traced
        V3<R> wo, value;
        R density;
        int lobe;
        if (!bsdf_sample(m, lobes, wi, g, wo, density, lobe, value)) break;
        ray_dir = transform_to_basis(basis, wo);
        ray_org = hit.world_position;
        throughput = throughput * (value / rmax(density, K<R>::PDF_EPSILON) * rabs(cos_theta(wo)));
    }
    return accumulated;
}

// integrate.glsl:323-340 for one pixel: the sample average (before the
// progressive accumulation of :341-346, applied by the caller).
template <typename R>
inline V3<R> render_pixel(const Scene<R>& s, const Camera<R>& c, const tt_constants& k, const Settings& st,
                          uint32_t x, uint32_t y, TraceCounts& counts) {
    Generator g = init_generator_from_pixel(x, y, k.screen_size[0], k.frame_index, st.rng, st.seed);
    V3<R> accumulated;
    for (uint32_t sample = 0; sample < k.sample_count; sample++) {
        R ox = random_float<R>(g);
        R oy = random_float<R>(g);
        ox = ox - R(0.5);
        oy = oy - R(0.5);
        R ndc_x = ((R(float(x)) + ox) / R(float(k.screen_size[0]))) * R(2) - R(1);
        R ndc_y = ((R(float(y)) + oy) / R(float(k.screen_size[1]))) * R(2) - R(1);
        V3<R> dir = camera_ray_direction(c, ndc_x, ndc_y);
        accumulated = accumulated + trace_path(s, c, k.light_sampling, k.bounce_count, dir, g, counts, ndc_x, ndc_y);
    }
    return accumulated / R(float(k.sample_count));
}

}  // namespace oracle
