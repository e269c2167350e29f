// Device view of a resident scene and the closest-hit traversal.
//
// Traversal replaces the rayQueryEXT loops of the reference
// (src/integrate/integrate.glsl:112-127 for extension rays with back-face
// culling, :168-185 for shadow rays without).  Per ray it walks the
// two-level BVH of bvh.h with a per-lane stack whose top TRACE_LDS_STACK
// entries live in LDS (entry e of lane t at lds[e * blockDim + t], so a
// wave's pushes hit 64 consecutive banks) and whose rare deeper entries spill
// to a per-lane global region.  Entering an instance pushes a sentinel and
// switches to the object-space ray (the instance's f32 inverse_transform
// applied to origin and direction); popping the sentinel switches back.
//
// Triangle test: the watertight test of Woop, Benthin & Wald (JCGT 2013)
// with its double-precision fallback, in object space on the dequantized
// snorm16 vertices; culling keeps U, V, W >= 0 (the CCW normal faces the
// origin); hits need tmin <= t <= tmax; the closest hit is the minimum of
// (t, instance, primitive).  Box tests are conservative: boxes are padded at
// build time, the slab origin is shifted by e = 64*eps*max|o| (+1e-30) and
// slab distances widen by a relative 1e-6, so no box rejects a triangle the
// exact test accepts and the result equals brute force over all triangles.
#pragma once

#include "device_math.h"

namespace ttd {

constexpr uint32_t LEAF = 0x80000000u;
constexpr uint32_t SENTINEL = 0xffffffffu;
#ifndef TT_TRACE_BLOCK
#define TT_TRACE_BLOCK 128
#endif
#ifndef TT_TRACE_LDS_STACK
#define TT_TRACE_LDS_STACK 32
#endif
constexpr int TRACE_BLOCK = TT_TRACE_BLOCK;
constexpr int TRACE_LDS_STACK = TT_TRACE_LDS_STACK;
// LDS entries + spill entries cover the deepest two-level path the builder
// allows: 3 pending children per wide level plus the instance sentinel
// (bvh.h BVH_WIDE_STACK_LIMIT = 148).
constexpr int TRACE_STACK_LIMIT = 148;
constexpr int TRACE_SPILL_STACK = TRACE_STACK_LIMIT - TRACE_LDS_STACK;

struct SceneView {
    const float4* nodes;      // BvhWide, 8 x float4
    const float4* tris;       // BvhTri, 3 x float4
    const float4* bvh_insts;  // BvhInstance, 4 x float4
    const float4* inst_boxes; // padded world box per instance, 2 x float4
    uint32_t tlas_root;
    uint32_t emissive_count;
    const float* instances;   // tt_instance, 52 floats (208 B)
    const float4* materials;  // tt_material, 19 x float4 (304 B)
    const int16_t* positions; // 3 per vertex
    const uint2* vertices;    // tt_vertex: (f16x2 uv, tangent frame)
    const uint32_t* indices;
    const uint32_t* emissive; // tt_emissive_triangle, 9 x u32 (36 B)
    const float4* lights;     // light records, 5 x float4 per emissive triangle
    const uint4* tex_records; // tt_texture_record, 5 x uint4
    const uint8_t* tex_data;
    const float* srgb_lut;
    uint32_t texture_count;
};

struct HitRecord {
    float t;
    uint32_t instance, primitive;
    float b1, b2;
};

struct TraceRay {
    f3 inv;           // 1 / direction (inf where a component is 0)
    f3 olo_inv, ohi_inv;  // slab origins shifted by +e / -e, times inv
    // Woop setup
    int kx, ky, kz;
    float sx, sy, sz;
    f3 o, d;
};

TT_DEV float component(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// A direction component smaller in magnitude than tiny = e * 1e-6 (zero
// included, which the instance transform produces for rays nearly
// perpendicular to an object axis) is replaced by +-tiny for the box tests.
// With 1 / 0 = inf the fused slab distance l * inv - (o + e) * inv would be
// inf - inf = NaN for every box on the ray's side of the axis' origin, and
// that axis, which should cull every box not straddling the ray's constant
// coordinate, would cull nothing (one such ray walked half of an 81 920-
// triangle BLAS).  With +-tiny the axis accepts exactly the boxes within e of
// that coordinate, over a t range of +-e / tiny = +-1e6 (beyond every tmax),
// and (o + e) / tiny <= 1.3e5 / 1e-6 stays finite; the shift e still covers
// the rounding of (o + e) * inv (at most 8e3 there).  Triangle tests use the
// true direction.
TT_DEV void slab_setup(TraceRay& r, f3 o, f3 d) {
    float mo = fmaxf(fmaxf(abs_(o.x), abs_(o.y)), abs_(o.z));
    float e = mo * (64.0f * 1.1920929e-7f) + 1e-30f;
    float tiny = e * 1e-6f;
    float dx = abs_(d.x) < tiny ? copysignf(tiny, d.x) : d.x;
    float dy = abs_(d.y) < tiny ? copysignf(tiny, d.y) : d.y;
    float dz = abs_(d.z) < tiny ? copysignf(tiny, d.z) : d.z;
    r.inv = f3(1.0f / dx, 1.0f / dy, 1.0f / dz);
    r.olo_inv = f3((o.x + e) * r.inv.x, (o.y + e) * r.inv.y, (o.z + e) * r.inv.z);
    r.ohi_inv = f3((o.x - e) * r.inv.x, (o.y - e) * r.inv.y, (o.z - e) * r.inv.z);
}

TT_DEV void woop_setup(TraceRay& r, f3 o, f3 d) {
    r.o = o;
    r.d = d;
    float ax = abs_(d.x), ay = abs_(d.y), az = abs_(d.z);
    int kz = (ax >= ay) ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dz = component(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.sx = component(d, kx) / dz;
    r.sy = component(d, ky) / dz;
    r.sz = 1.0f / dz;
}

// Conservative slab interval of a box: plane distances as one fused
// multiply-add from the shifted origins (the shift e covers the rounding of
// origin * inv many times over), widened by a relative 1e-6.  Directions
// are never zero here (slab_setup); a NaN distance (a NaN ray) is ignored by
// min/max, so that axis does not cull and the test stays conservative.
TT_DEV void box_interval(const TraceRay& r, float lx, float hx, float ly, float hy, float lz, float hz, float& n,
                         float& f) {
    float t0x = __builtin_fmaf(lx, r.inv.x, -r.olo_inv.x), t1x = __builtin_fmaf(hx, r.inv.x, -r.ohi_inv.x);
    float t0y = __builtin_fmaf(ly, r.inv.y, -r.olo_inv.y), t1y = __builtin_fmaf(hy, r.inv.y, -r.ohi_inv.y);
    float t0z = __builtin_fmaf(lz, r.inv.z, -r.olo_inv.z), t1z = __builtin_fmaf(hz, r.inv.z, -r.ohi_inv.z);
    n = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    f = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
}

// Box test against [tmin, tmax]; `tn` = raw entry distance for ordering.
TT_DEV bool box_hit(const TraceRay& r, float lx, float hx, float ly, float hy, float lz, float hz, float tmin,
                    float tmax, float& tn) {
    float n, f;
    box_interval(r, lx, hx, ly, hy, lz, hz, n, f);
    tn = n;
    float nc = __builtin_fmaf(-abs_(n), 1e-6f, n);
    float fc = __builtin_fmaf(abs_(f), 1e-6f, f);
    return fmaxf(nc, tmin) <= fminf(fc, tmax);
}

constexpr float INF_F = __builtin_huge_valf();

// Compare-exchange of (key, reference) pairs: the smaller key first.
TT_DEV void sort_pair(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
    bool swap = kb < ka;
    float k = swap ? kb : ka;
    kb = swap ? ka : kb;
    ka = k;
    uint32_t v = swap ? rb : ra;
    rb = swap ? ra : rb;
    ra = v;
}

TT_DEV bool closer(float t, uint32_t inst, uint32_t prim, const HitRecord& b) {
    if (t != b.t) return t < b.t;
    if (inst != b.instance) return inst < b.instance;
    return prim < b.primitive;
}

// Woop et al. 2013; U, V, W and det of an accepted hit are kept so the
// barycentrics are divided once at the end.
template <bool CULL>
TT_DEV bool triangle_hit(const TraceRay& r, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float& t_out, float& V_out,
                         float& W_out, float& det_out) {
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = component(A, r.kz), Bkz = component(B, r.kz), Ckz = component(C, r.kz);
    float Ax = component(A, r.kx) - r.sx * Akz;
    float Ay = component(A, r.ky) - r.sy * Akz;
    float Bx = component(B, r.kx) - r.sx * Bkz;
    float By = component(B, r.ky) - r.sy * Bkz;
    float Cx = component(C, r.kx) - r.sx * Ckz;
    float Cy = component(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    // Double-precision recomputation on an exact zero.  Without culling a
    // wave-uniform branch keeps it off the common path (as a per-lane branch
    // the compiler evaluates it for every triangle and selects); with
    // culling the per-lane form measured faster.
    bool degenerate = U == 0.0f || V == 0.0f || W == 0.0f;
    if (CULL || __builtin_amdgcn_ballot_w64(degenerate) != 0ull) {
        if (degenerate) {
            U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
            V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
            W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
        }
    }
    if (CULL) {
        if (U < 0.0f || V < 0.0f || W < 0.0f) return false;
    } else {
        if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    }
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    V_out = V;
    W_out = W;
    det_out = det;
    return true;
}

// Top TRACE_LDS_STACK entries of every lane's stack: entry e of lane t at
// trace_stack[e * TRACE_BLOCK + t].  A file-scope LDS array, so pushes and
// pops compile to ds_write / ds_read rather than generic flat accesses.
static __shared__ uint32_t trace_stack[TRACE_LDS_STACK * TRACE_BLOCK];

struct Stack {
    uint32_t lane;    // threadIdx.x
    uint32_t* spill;  // entry e at spill[(e - TRACE_LDS_STACK) * spill_stride]
    uint32_t spill_stride;
    uint32_t sp;
    TT_DEV void push(uint32_t v) {
        if (sp < uint32_t(TRACE_LDS_STACK))
            trace_stack[sp * TRACE_BLOCK + lane] = v;
        else
            static_cast<volatile uint32_t*>(spill)[(sp - TRACE_LDS_STACK) * spill_stride] = v;
        sp++;
    }
    TT_DEV uint32_t pop() {
        sp--;
        // The LDS read is unconditional (clamped slot) so it stays a ds_read;
        // only the rare spilled entry comes from global memory.
        uint32_t slot = sp < uint32_t(TRACE_LDS_STACK) ? sp : uint32_t(TRACE_LDS_STACK - 1);
        uint32_t v = trace_stack[slot * TRACE_BLOCK + lane];
        // volatile keeps the spill read a separate global load (otherwise it
        // is folded with the LDS read into one generic flat load).
        if (sp >= uint32_t(TRACE_LDS_STACK)) v = static_cast<volatile uint32_t*>(spill)[(sp - TRACE_LDS_STACK) * spill_stride];
        return v;
    }
};

// A shadow query: the ray was aimed at a point sampled on an emissive
// triangle of instance light_instance whose emissive hash is `hash`.  The
// reference accepts the light iff the closest hit in [1e-3, 1e4] lies on that
// instance and (primitive + index_offset / 3) % 0xffff == hash
// (integrate.glsl:168-185).  Every hit on the instance lies in the
// conservative slab interval [t_entry, t_far] of the ray against the
// instance's padded world box, so the query is capped at t_far and ends at
// the first hit that fails the check with t < t_entry: a strictly closer hit
// that is not the light decides "occluded" whatever lies behind it.
struct ShadowQuery {
    uint32_t light_instance;
    uint32_t hash;
    float t_exit;   // t_entry of the light instance's box (NaN: no early exit)
    bool occluded;  // set when a non-light hit closer than t_exit ended the query
};

// Closest hit of a world-space ray.  Returns false on a miss.
template <bool CULL, bool SHADOW = false>
TT_DEV bool trace_closest(const SceneView& s, f3 org, f3 dir, float tmin, float tmax, Stack& stack, HitRecord& best,
                          ShadowQuery* q = nullptr) {
    best.t = tmax;
    best.instance = 0xffffffffu;
    best.primitive = 0xffffffffu;
    best.b1 = 0.0f;
    best.b2 = 0.0f;
    if (s.tlas_root == SENTINEL) return false;
    bool found = false;
    float bV = 0.0f, bW = 0.0f, bdet = 1.0f;
    TraceRay r;
    slab_setup(r, org, dir);
    bool in_blas = false;
    uint32_t instance = 0;
    uint32_t triangle_base = 0;  // index_offset / 3 of the current instance
    uint32_t cur = s.tlas_root;
    stack.sp = 0;
    for (;;) {
        if (!(cur & LEAF)) {
            // Wide node: SoA boxes of four children, then their references.
            const float4* n = s.nodes + 8 * size_t(cur);
            float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5];
            uint4 ch = *reinterpret_cast<const uint4*>(n + 6);
            float limit = found ? best.t : tmax;
            float k0, k1, k2, k3;
            // Sort keys: entry distance of a hit child (made finite), +inf
            // for a miss or an unused slot (reference SENTINEL, only in
            // slots 2 and 3; its box is not trusted to miss, since the
            // conservative widening can accept a box at infinity).
            uint32_t r0 = ch.x, r1 = ch.y, r2 = ch.z, r3 = ch.w;
            {
                float tn;
                bool h0 = box_hit(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, tmin, limit, tn);
                k0 = h0 ? fminf(tn, 3.0e38f) : INF_F;
                bool h1 = box_hit(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, tmin, limit, tn);
                k1 = h1 ? fminf(tn, 3.0e38f) : INF_F;
                bool h2 = box_hit(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, tmin, limit, tn) && r2 != SENTINEL;
                k2 = h2 ? fminf(tn, 3.0e38f) : INF_F;
                bool h3 = box_hit(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, tmin, limit, tn) && r3 != SENTINEL;
                k3 = h3 ? fminf(tn, 3.0e38f) : INF_F;
            }
            // Five compare-exchanges sort the four (key, reference) pairs.
            sort_pair(k0, r0, k1, r1);
            sort_pair(k2, r2, k3, r3);
            sort_pair(k0, r0, k2, r2);
            sort_pair(k1, r1, k3, r3);
            sort_pair(k1, r1, k2, r2);
            if (k3 != INF_F) stack.push(r3);
            if (k2 != INF_F) stack.push(r2);
            if (k1 != INF_F) stack.push(r1);
            if (k0 != INF_F) {
                cur = r0;
                continue;
            }
        } else if (!in_blas) {
            instance = cur & 0x7fffffffu;
            const float4* ip = s.bvh_insts + 4 * size_t(instance);
            float4 c0 = ip[0], c1 = ip[1], c2 = ip[2], c3 = ip[3];
            uint32_t root = __float_as_uint(c0.w);
            if (root != SENTINEL) {
                f3 o(c0.x * org.x + c1.x * org.y + c2.x * org.z + c3.x, c0.y * org.x + c1.y * org.y + c2.y * org.z + c3.y,
                     c0.z * org.x + c1.z * org.y + c2.z * org.z + c3.z);
                f3 d(c0.x * dir.x + c1.x * dir.y + c2.x * dir.z, c0.y * dir.x + c1.y * dir.y + c2.y * dir.z,
                     c0.z * dir.x + c1.z * dir.y + c2.z * dir.z);
                slab_setup(r, o, d);
                woop_setup(r, o, d);
                stack.push(SENTINEL);
                in_blas = true;
                if (SHADOW) triangle_base = __float_as_uint(c1.w);
                cur = root;
                continue;
            }
        } else {
            uint32_t first = cur & 0x7ffffffu;
            uint32_t count = ((cur >> 27) & 0xfu) + 1u;
            for (uint32_t k = 0; k < count; k++) {
                const float4* tp = s.tris + 3 * size_t(first + k);
                float4 a = tp[0], b = tp[1], c = tp[2];
                float t, V, W, det;
                float limit = found ? best.t : tmax;
                if (triangle_hit<CULL>(r, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(c.x, c.y, c.z), tmin, limit, t, V, W,
                                       det)) {
                    uint32_t prim = __float_as_uint(a.w);
                    if (SHADOW) {
                        bool light = instance == q->light_instance && (prim + triangle_base) % 0xffffu == q->hash;
                        if (!light && t < q->t_exit) {
                            q->occluded = true;
                            best.t = t;
                            best.instance = instance;
                            best.primitive = prim;
                            return true;
                        }
                    }
                    if (!found || closer(t, instance, prim, best)) {
                        found = true;
                        best.t = t;
                        best.instance = instance;
                        best.primitive = prim;
                        bV = V;
                        bW = W;
                        bdet = det;
                    }
                }
            }
        }
        // Pop the next reference; a sentinel returns to world space.
        for (;;) {
            if (stack.sp == 0) {
                if (found) {
                    best.b1 = bV / bdet;
                    best.b2 = bW / bdet;
                }
                return found;
            }
            cur = stack.pop();
            if (cur == SENTINEL) {
                in_blas = false;
                slab_setup(r, org, dir);
                continue;
            }
            break;
        }
    }
}

// The reference's shadow-ray rule for a light sample (see ShadowQuery).
TT_DEV bool light_visible(const SceneView& s, f3 org, f3 dir, uint32_t light_instance, uint32_t hash, Stack& stack) {
    const float4* box = s.inst_boxes + 2 * size_t(light_instance);
    float4 lo = box[0], hi = box[1];
    TraceRay r;
    slab_setup(r, org, dir);
    float n, f;
    box_interval(r, lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, n, f);
    float t_entry = __builtin_fmaf(-abs_(n), 1e-6f, n);
    float t_far = __builtin_fmaf(abs_(f), 1e-6f, f);
    const float tmin = 1.0e-3f;
    float tmax = fminf(t_far, 10000.0f);
    if (!(fmaxf(t_entry, tmin) <= tmax)) return false;  // the light cannot be hit
    ShadowQuery q;
    q.light_instance = light_instance;
    q.hash = hash;
    q.t_exit = t_entry;
    q.occluded = false;
    HitRecord h;
    bool found = trace_closest<false, true>(s, org, dir, tmin, tmax, stack, h, &q);
    if (!found || q.occluded || h.instance != light_instance) return false;
    uint32_t index_offset = __float_as_uint(s.instances[52 * size_t(h.instance) + 49]);
    return (h.primitive + index_offset / 3u) % 0xffffu == hash;
}

}  // namespace ttd
