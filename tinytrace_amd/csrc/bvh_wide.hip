// Collapse of the binary BVH (both builders' output) into the 4-wide nodes
// the traversal reads.
//
// Every binary node at even depth below its tree's root (a BLAS root, or
// the TLAS root) becomes one wide node; its children are its grandchildren
// where its child is an inner node, or the child itself where that is a
// leaf.  So a wide node has 2 to 4 children, one fetch of 128 B tests all
// their boxes, and a ray takes half the steps.  Boxes are copied bit for
// bit, so the padded-box guarantees of the binary trees (no box rejects a
// triangle the exact test accepts) hold unchanged.
//
//   1. parent of every inner node (a node without one is a tree root);
//   2. depth parity by walking up to the root (bounded walk);
//   3. exclusive scan of "even depth" flags: the wide index of each such node;
//   4. emission of the wide nodes, inner references renumbered;
//   5. BLAS roots in the instance records renumbered; the TLAS root
//      renumbered on the host.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "bvh.h"
#include "runtime_internal.h"

namespace {

constexpr uint32_t NONE = 0xffffffffu;
// Longest parent chain walked; deeper chains than any built tree has
// (BVH_STACK_LIMIT) mean a malformed array and are reported.
constexpr uint32_t WALK_LIMIT = 4 * tt::BVH_STACK_LIMIT;

__global__ void wide_parent_kernel(const tt::BvhNode* nodes, uint32_t count, uint32_t* parent, uint32_t* error) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    for (int c = 0; c < 2; c++) {
        uint32_t ref = nodes[i].n3[c];
        if (ref & tt::BVH_LEAF) continue;
        if (ref < count) parent[ref] = i;
        else atomicOr(error, 2u);
    }
}

__global__ void wide_fill_kernel(uint32_t* p, uint32_t count, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void wide_flag_kernel(const uint32_t* parent, uint32_t count, uint32_t* flag, uint32_t* error) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t odd = 0, steps = 0;
    uint32_t p = parent[i];
    while (p != NONE && steps < WALK_LIMIT) {
        odd ^= 1u;
        steps++;
        p = parent[p];
    }
    if (p != NONE) atomicOr(error, 1u);
    flag[i] = odd ? 0u : 1u;
}

struct SlotBox {
    float lx, hx, ly, hy, lz, hz;
};

__device__ __forceinline__ SlotBox slot_box(const tt::BvhNode& n, int c) {
    SlotBox b;
    b.lx = c ? n.n1[0] : n.n0[0];
    b.hx = c ? n.n1[1] : n.n0[1];
    b.ly = c ? n.n1[2] : n.n0[2];
    b.hy = c ? n.n1[3] : n.n0[3];
    b.lz = n.n2[2 * c];
    b.hz = n.n2[2 * c + 1];
    return b;
}

__global__ void wide_emit_kernel(const tt::BvhNode* nodes, uint32_t count, const uint32_t* flag,
                                 const uint32_t* index, tt::BvhWide* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || !flag[i]) return;
    const tt::BvhNode n = nodes[i];
    tt::BvhWide w;
    int k = 0;
    auto put = [&](const SlotBox& b, uint32_t ref) {
        w.lo_x[k] = b.lx;
        w.hi_x[k] = b.hx;
        w.lo_y[k] = b.ly;
        w.hi_y[k] = b.hy;
        w.lo_z[k] = b.lz;
        w.hi_z[k] = b.hz;
        w.child[k] = ref;
        k++;
    };
    for (int c = 0; c < 2; c++) {
        uint32_t ref = n.n3[c];
        if (ref & tt::BVH_LEAF) {
            put(slot_box(n, c), ref);
        } else {
            const tt::BvhNode m = nodes[ref];
            for (int g = 0; g < 2; g++) {
                uint32_t r = m.n3[g];
                put(slot_box(m, g), (r & tt::BVH_LEAF) ? r : index[r]);
            }
        }
    }
    while (k < 4) {
        SlotBox none{INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
        put(none, NONE);
    }
    for (int q = 0; q < 4; q++) w.pad[q] = 0u;
    out[index[i]] = w;
}

// Instance records: c0.w is the BLAS root reference.
__global__ void wide_roots_kernel(float4* instances, uint32_t count, const uint32_t* index, uint32_t nodes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 c0 = instances[4 * size_t(i)];
    uint32_t root = __float_as_uint(c0.w);
    if (root != NONE && !(root & tt::BVH_LEAF) && root < nodes) {
        c0.w = __uint_as_float(index[root]);
        instances[4 * size_t(i)] = c0;
    }
}

uint32_t blocks(uint64_t n) { return uint32_t((n + 255) / 256); }

}  // namespace

namespace tt {

tt_status collapse_wide(tt_context* c, const BvhNode* nodes, uint32_t count, void* instances,
                        uint32_t instance_count, uint32_t& tlas_root, BvhWide** out, uint32_t& wide_count) {
    *out = nullptr;
    wide_count = 0;
    hipStream_t st = tt::api_stream(c);
    if (count == 0) {
        // No inner nodes: every root is a leaf (or the scene is empty).
        return TT_OK;
    }
    size_t scan_bytes = 0;
    {
        uint32_t* np = nullptr;
        hipError_t e = rocprim::exclusive_scan(nullptr, scan_bytes, np, np, 0u, size_t(count), rocprim::plus<uint32_t>(),
                                               st);
        if (e != hipSuccess) return hip_fail(c, e, "scan size");
    }
    size_t words = 3 * size_t(count) + 64;
    char* scratch = nullptr;
    TT_HIP(c, hipMalloc(&scratch, words * sizeof(uint32_t) + scan_bytes + 256));
    uint32_t* parent = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* flag = parent + count;
    uint32_t* index = flag + count;
    uint32_t* error = index + count;
    void* scan_temp = scratch + ((words * sizeof(uint32_t) + 255) & ~size_t(255));
    BvhWide* wide = nullptr;
    auto done = [&](tt_status s) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(scratch);
        if (s != TT_OK && wide) (void)hipFree(wide);
        return s;
    };
#define TT_WIDE(expr)                                                  \
    do {                                                               \
        hipError_t e_ = (expr);                                        \
        if (e_ != hipSuccess) return done(hip_fail(c, e_, #expr));     \
    } while (0)
    wide_fill_kernel<<<blocks(count), 256, 0, st>>>(parent, count, NONE);
    wide_fill_kernel<<<1, 64, 0, st>>>(error, 1, 0u);
    TT_WIDE(hipGetLastError());
    wide_parent_kernel<<<blocks(count), 256, 0, st>>>(nodes, count, parent, error);
    TT_WIDE(hipGetLastError());
    wide_flag_kernel<<<blocks(count), 256, 0, st>>>(parent, count, flag, error);
    TT_WIDE(hipGetLastError());
    TT_WIDE(rocprim::exclusive_scan(scan_temp, scan_bytes, flag, index, 0u, size_t(count), rocprim::plus<uint32_t>(),
                                    st));
    uint32_t tail[2] = {0, 0}, err = 0;
    TT_WIDE(hipMemcpyAsync(&tail[0], index + count - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TT_WIDE(hipMemcpyAsync(&tail[1], flag + count - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TT_WIDE(hipMemcpyAsync(&err, error, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TT_WIDE(hipStreamSynchronize(st));
    if (err) return done(fail(c, TT_INVALID_ARGUMENT, "BVH collapse: malformed binary tree"));
    uint32_t n = tail[0] + tail[1];
    TT_WIDE(hipMalloc(&wide, std::max<size_t>(n, 1) * sizeof(BvhWide)));
    wide_emit_kernel<<<blocks(count), 256, 0, st>>>(nodes, count, flag, index, wide);
    TT_WIDE(hipGetLastError());
    if (instance_count > 0) {
        wide_roots_kernel<<<blocks(instance_count), 256, 0, st>>>(static_cast<float4*>(instances), instance_count,
                                                                   index, count);
        TT_WIDE(hipGetLastError());
    }
    if (tlas_root != NONE && !(tlas_root & BVH_LEAF)) {
        if (tlas_root >= count) return done(fail(c, TT_INVALID_ARGUMENT, "BVH collapse: TLAS root out of range"));
        uint32_t r = 0;
        TT_WIDE(hipMemcpyAsync(&r, index + tlas_root, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        TT_WIDE(hipStreamSynchronize(st));
        tlas_root = r;
    }
#undef TT_WIDE
    *out = wide;
    wide_count = n;
    return done(TT_OK);
}

}  // namespace tt
