// Device LBVH build of every BLAS (Karras 2012, "Maximizing parallelism in
// the construction of BVHs, octrees, and k-d trees"), the fast-build
// alternative to the host binned-SAH builder.  All meshes are built in one
// pass:
//   1. per triangle: dequantised object-space vertices (v / 32767, as the
//      host builder), box, centroid; centroid bounds per mesh (atomics on
//      order-preserving float encodings)
//   2. 64-bit keys (mesh << 32 | 30-bit Morton code of the centroid in its
//      mesh's centroid bounds), sorted with rocPRIM's radix sort, so every
//      mesh is one contiguous run of Morton-ordered triangles
//   3. per internal node: Karras' range / split search inside its mesh's run
//      (key ties broken by position), children and parents
//   4. bottom-up boxes: the second thread to reach a node unions its
//      children and writes their padded boxes into the 64-B node
// Leaves hold one triangle.  Boxes use the host builder's padding, so the
// traversal's guarantees (closest hit == brute force) are unchanged.  The
// TLAS (over instances, small) is then built on the host from the per-mesh
// summaries this pass returns.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "bvh.h"
#include "runtime_internal.h"

namespace {

constexpr uint32_t NONE = 0xffffffffu;

// Order-preserving float <-> uint for atomic min / max.
__device__ __forceinline__ uint32_t ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Host builder's outward padding: 2^-20 of the coordinate magnitude.
__device__ __forceinline__ float pad_lo(float v) { return v - (fabsf(v) * 9.5367431640625e-07f + 1e-30f); }
__device__ __forceinline__ float pad_hi(float v) { return v + (fabsf(v) * 9.5367431640625e-07f + 1e-30f); }

struct MeshRun {
    uint32_t tri_offset;     // first triangle (global, = sorted position)
    uint32_t tri_count;
    uint32_t vertex_offset;
    uint32_t index_offset;
    uint32_t node_base;      // first internal node
    uint32_t pad[3];
};

__device__ __forceinline__ uint32_t find_run(const MeshRun* runs, uint32_t count, uint32_t g, bool by_node) {
    // Last run whose start <= g.
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        uint32_t start = by_node ? runs[mid].node_base : runs[mid].tri_offset;
        if (start <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void lbvh_setup_kernel(const MeshRun* runs, uint32_t mesh_count, const int16_t* positions,
                                  const uint32_t* indices, uint32_t total, tt::BvhTri* raw, float* boxes,
                                  uint32_t* centroid_bounds, uint32_t* tri_mesh) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    // The last run starting at or before g (empty runs share their start
    // with the next one, which the search prefers).
    uint32_t m = find_run(runs, mesh_count, g, false);
    const MeshRun r = runs[m];
    uint32_t t = g - r.tri_offset;
    float v[3][3];
    for (int c = 0; c < 3; c++) {
        uint32_t vi = r.vertex_offset + indices[r.index_offset + 3 * t + c];
        for (int a = 0; a < 3; a++) v[c][a] = float(positions[3 * size_t(vi) + a]) / 32767.0f;
    }
    tt::BvhTri bt;
    for (int a = 0; a < 3; a++) {
        bt.v0[a] = v[0][a];
        bt.v1[a] = v[1][a];
        bt.v2[a] = v[2][a];
    }
    bt.v0[3] = __uint_as_float(t);
    bt.v1[3] = 0.0f;
    bt.v2[3] = 0.0f;
    raw[g] = bt;
    tri_mesh[g] = m;
    for (int a = 0; a < 3; a++) {
        float lo = fminf(fminf(v[0][a], v[1][a]), v[2][a]);
        float hi = fmaxf(fmaxf(v[0][a], v[1][a]), v[2][a]);
        boxes[6 * size_t(g) + a] = lo;
        boxes[6 * size_t(g) + 3 + a] = hi;
        float cen = (lo + hi) * 0.5f;
        atomicMin(centroid_bounds + 6 * m + a, ordered(cen));
        atomicMax(centroid_bounds + 6 * m + 3 + a, ordered(cen));
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void lbvh_morton_kernel(const float* boxes, const uint32_t* tri_mesh, const uint32_t* centroid_bounds,
                                   uint32_t total, uint64_t* keys, uint32_t* values) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    uint32_t m = tri_mesh[g];
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float lo = unordered(centroid_bounds[6 * m + a]), hi = unordered(centroid_bounds[6 * m + 3 + a]);
        float cen = (boxes[6 * size_t(g) + a] + boxes[6 * size_t(g) + 3 + a]) * 0.5f;
        float ext = hi - lo;
        float u = ext > 0.0f ? (cen - lo) / ext : 0.5f;
        q[a] = uint32_t(fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f));
    }
    uint32_t code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    keys[g] = (uint64_t(m) << 32) | code;
    values[g] = g;
}

__global__ void lbvh_reorder_kernel(const uint32_t* values, const tt::BvhTri* raw, const float* boxes, uint32_t total,
                                    tt::BvhTri* tris, float* leaf_boxes) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= total) return;
    uint32_t g = values[j];
    tris[j] = raw[g];
    for (int k = 0; k < 6; k++) leaf_boxes[6 * size_t(j) + k] = boxes[6 * size_t(g) + k];
}

// Common-prefix length of positions i and j of a run (-1 outside it); equal
// keys compare by position.
__device__ __forceinline__ int delta(const uint64_t* keys, uint32_t first, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = uint32_t(keys[first + i]), b = uint32_t(keys[first + j]);
    if (a == b) return 32 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clz(a ^ b);
}

__global__ void lbvh_hierarchy_kernel(const MeshRun* runs, uint32_t mesh_count, const uint64_t* keys,
                                      uint32_t internal_total, tt::BvhNode* nodes, uint32_t* parent_internal,
                                      uint32_t* parent_leaf) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= internal_total) return;
    uint32_t m = find_run(runs, mesh_count, k, true);
    const MeshRun r = runs[m];
    const int n = int(r.tri_count);
    const int i = int(k - r.node_base);
    const uint32_t first = r.tri_offset;
    int d = (delta(keys, first, n, i, i + 1) - delta(keys, first, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, first, n, i, i - d);
    int lmax = 2;
    while (delta(keys, first, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, first, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, first, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) / 2;
        if (delta(keys, first, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t left, right;
    if (lo == gamma) {
        left = tt::BVH_LEAF | (first + uint32_t(gamma));
        parent_leaf[first + gamma] = k;
    } else {
        left = r.node_base + uint32_t(gamma);
        parent_internal[left] = k;
    }
    if (hi == gamma + 1) {
        right = tt::BVH_LEAF | (first + uint32_t(gamma + 1));
        parent_leaf[first + gamma + 1] = k;
    } else {
        right = r.node_base + uint32_t(gamma + 1);
        parent_internal[right] = k;
    }
    tt::BvhNode& node = nodes[k];
    node.n3[0] = left;
    node.n3[1] = right;
    node.n3[2] = 0;
    node.n3[3] = 0;
    if (i == 0) parent_internal[k] = NONE;
}

__device__ __forceinline__ double box_area(const float* b) {
    double dx = double(b[3]) - b[0], dy = double(b[4]) - b[1], dz = double(b[5]) - b[2];
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

// Bottom-up boxes from every leaf; per mesh: depth (longest leaf-to-root
// chain) and the area sums of the SAH statistics.
__global__ void lbvh_boxes_kernel(const uint32_t* tri_mesh_sorted, const MeshRun* runs, uint32_t total,
                                  const float* leaf_boxes, float* internal_boxes, tt::BvhNode* nodes,
                                  const uint32_t* parent_internal, const uint32_t* parent_leaf, uint32_t* visits,
                                  uint32_t* heights, uint32_t* depth, double* area_internal, double* area_leaf) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= total) return;
    uint32_t m = tri_mesh_sorted[j];
    atomicAdd(area_leaf + m, box_area(leaf_boxes + 6 * size_t(j)));
    if (runs[m].tri_count < 2) return;
    uint32_t node = parent_leaf[j];
    while (node != NONE) {
        __threadfence();
        if (atomicAdd(visits + node, 1u) == 0u) return;  // the sibling's thread finishes this node
        __threadfence();
        tt::BvhNode& nd = nodes[node];
        float cb[2][6];
        uint32_t h = 0;
        for (int c = 0; c < 2; c++) {
            uint32_t ref = nd.n3[c];
            const volatile float* src = (ref & tt::BVH_LEAF) ? leaf_boxes + 6 * size_t(ref & 0x7fffffffu)
                                                             : internal_boxes + 6 * size_t(ref);
            for (int k = 0; k < 6; k++) cb[c][k] = src[k];
            if (!(ref & tt::BVH_LEAF)) h = max(h, static_cast<const volatile uint32_t*>(heights)[ref]);
        }
        // Height of this node above its deepest leaf.
        heights[node] = h + 1;
        float u[6];
        for (int a = 0; a < 3; a++) {
            u[a] = fminf(cb[0][a], cb[1][a]);
            u[3 + a] = fmaxf(cb[0][3 + a], cb[1][3 + a]);
        }
        for (int k = 0; k < 6; k++) internal_boxes[6 * size_t(node) + k] = u[k];
        nd.n0[0] = pad_lo(cb[0][0]);
        nd.n0[1] = pad_hi(cb[0][3]);
        nd.n0[2] = pad_lo(cb[0][1]);
        nd.n0[3] = pad_hi(cb[0][4]);
        nd.n1[0] = pad_lo(cb[1][0]);
        nd.n1[1] = pad_hi(cb[1][3]);
        nd.n1[2] = pad_lo(cb[1][1]);
        nd.n1[3] = pad_hi(cb[1][4]);
        nd.n2[0] = pad_lo(cb[0][2]);
        nd.n2[1] = pad_hi(cb[0][5]);
        nd.n2[2] = pad_lo(cb[1][2]);
        nd.n2[3] = pad_hi(cb[1][5]);
        atomicAdd(area_internal + m, box_area(u));
        atomicMax(depth + m, h + 1);
        node = parent_internal[node];
    }
}

uint32_t grid_of(uint64_t n, uint32_t b) { return uint32_t((n + b - 1) / b); }

}  // namespace

namespace tt {

// Builds every BLAS of `d` on the device: nodes into nodes[0 .. internal
// count), triangles (Morton order) into tris; fills the per-mesh summaries.
tt_status build_blas_device(tt_context* c, const tt_scene_desc& d, const int16_t* positions, const uint32_t* indices,
                            BvhNode* nodes, BvhTri* tris, std::vector<BlasSummary>& summary, uint32_t& node_count) {
    hipStream_t st = tt::api_stream(c);
    const uint32_t meshes = d.mesh_count;
    std::vector<MeshRun> runs(std::max(meshes, 1u));
    uint64_t total = 0, internal = 0;
    for (uint32_t m = 0; m < meshes; m++) {
        MeshRun& r = runs[m];
        r.tri_offset = uint32_t(total);
        r.tri_count = d.meshes[m].index_count / 3;
        r.vertex_offset = d.meshes[m].vertex_offset;
        r.index_offset = d.meshes[m].index_offset;
        r.node_base = uint32_t(internal);
        total += r.tri_count;
        internal += r.tri_count > 1 ? r.tri_count - 1 : 0;
    }
    node_count = uint32_t(internal);
    summary.assign(meshes, BlasSummary());
    if (total == 0) return TT_OK;

    // Scratch: runs, raw tris, boxes, keys (2), values (2), leaf/internal
    // boxes, parents, visits, bounds, per-mesh stats.
    size_t n = total;
    size_t bytes_runs = runs.size() * sizeof(MeshRun);
    size_t sizes[] = {bytes_runs,           n * sizeof(BvhTri),      n * 6 * sizeof(float), n * sizeof(uint32_t),
                      2 * n * sizeof(uint64_t), 2 * n * sizeof(uint32_t), n * 6 * sizeof(float),
                      n * 6 * sizeof(float), 2 * n * sizeof(uint32_t), 2 * n * sizeof(uint32_t),
                      meshes * 6 * sizeof(uint32_t), meshes * sizeof(uint32_t), 2 * meshes * sizeof(double)};
    size_t offs[13], total_bytes = 0;
    for (int k = 0; k < 13; k++) {
        offs[k] = total_bytes;
        total_bytes += (sizes[k] + 255) & ~size_t(255);
    }
    uint32_t mesh_bits = 1;
    while ((1ull << mesh_bits) < meshes) mesh_bits++;
    size_t sort_bytes = 0;
    {
        uint64_t* nk = nullptr;
        uint32_t* nv = nullptr;
        hipError_t e = rocprim::radix_sort_pairs(nullptr, sort_bytes, nk, nk, nv, nv, n, 0, 32 + mesh_bits, st);
        if (e != hipSuccess) return tt::hip_fail(c, e, "radix sort size");
    }
    char* scratch = nullptr;
    TT_HIP(c, hipMalloc(&scratch, total_bytes + sort_bytes + 256));
    auto at = [&](int k) { return scratch + offs[k]; };
    MeshRun* d_runs = reinterpret_cast<MeshRun*>(at(0));
    BvhTri* raw = reinterpret_cast<BvhTri*>(at(1));
    float* boxes = reinterpret_cast<float*>(at(2));
    uint32_t* tri_mesh = reinterpret_cast<uint32_t*>(at(3));
    uint64_t* keys = reinterpret_cast<uint64_t*>(at(4));
    uint32_t* values = reinterpret_cast<uint32_t*>(at(5));
    float* leaf_boxes = reinterpret_cast<float*>(at(6));
    float* internal_boxes = reinterpret_cast<float*>(at(7));
    uint32_t* parents = reinterpret_cast<uint32_t*>(at(8));
    uint32_t* visits = reinterpret_cast<uint32_t*>(at(9));
    uint32_t* heights = visits + n;
    uint32_t* cbounds = reinterpret_cast<uint32_t*>(at(10));
    uint32_t* depth = reinterpret_cast<uint32_t*>(at(11));
    double* areas = reinterpret_cast<double*>(at(12));
    void* sort_temp = scratch + total_bytes;
    auto done = [&](tt_status s) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(scratch);
        return s;
    };
#define TT_LBVH(expr)                                                    \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) return done(tt::hip_fail(c, e_, #expr));   \
    } while (0)
    // Sorted mesh ids per position: the same runs, so reuse tri_mesh after
    // the sort by recomputing from the runs (keys carry the mesh id).
    std::vector<uint32_t> bounds_init(size_t(meshes) * 6);
    for (uint32_t m = 0; m < meshes; m++)
        for (int a = 0; a < 3; a++) {
            bounds_init[6 * m + a] = 0xffffffffu;  // ordered(+inf) < this: min starts high
            bounds_init[6 * m + 3 + a] = 0u;
        }
    TT_LBVH(hipMemcpyAsync(d_runs, runs.data(), bytes_runs, hipMemcpyHostToDevice, st));
    TT_LBVH(hipMemcpyAsync(cbounds, bounds_init.data(), bounds_init.size() * 4, hipMemcpyHostToDevice, st));
    TT_LBVH(hipMemsetAsync(visits, 0, n * sizeof(uint32_t), st));
    TT_LBVH(hipMemsetAsync(depth, 0, meshes * sizeof(uint32_t), st));
    TT_LBVH(hipMemsetAsync(areas, 0, 2 * meshes * sizeof(double), st));
    lbvh_setup_kernel<<<grid_of(n, 256), 256, 0, st>>>(d_runs, meshes, positions, indices, uint32_t(n), raw, boxes,
                                                       cbounds, tri_mesh);
    TT_LBVH(hipGetLastError());
    lbvh_morton_kernel<<<grid_of(n, 256), 256, 0, st>>>(boxes, tri_mesh, cbounds, uint32_t(n), keys, values);
    TT_LBVH(hipGetLastError());
    TT_LBVH(rocprim::radix_sort_pairs(sort_temp, sort_bytes, keys, keys + n, values, values + n, n, 0,
                                      32 + mesh_bits, st));
    const uint64_t* skeys = keys + n;
    const uint32_t* svalues = values + n;
    lbvh_reorder_kernel<<<grid_of(n, 256), 256, 0, st>>>(svalues, raw, boxes, uint32_t(n), tris, leaf_boxes);
    TT_LBVH(hipGetLastError());
    // The sort keeps every mesh on its run of positions, so tri_mesh (built
    // by triangle index over the same runs) also maps sorted positions.
    if (internal > 0) {
        lbvh_hierarchy_kernel<<<grid_of(internal, 256), 256, 0, st>>>(d_runs, meshes, skeys, uint32_t(internal), nodes,
                                                                      parents, parents + n);
        TT_LBVH(hipGetLastError());
    }
    lbvh_boxes_kernel<<<grid_of(n, 256), 256, 0, st>>>(tri_mesh, d_runs, uint32_t(n), leaf_boxes, internal_boxes,
                                                       nodes, parents, parents + n, visits, heights, depth, areas,
                                                       areas + meshes);
    TT_LBVH(hipGetLastError());

    // Per-mesh summaries.
    std::vector<uint32_t> h_depth(meshes);
    std::vector<double> h_area(2 * size_t(meshes));
    TT_LBVH(hipMemcpyAsync(h_depth.data(), depth, meshes * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TT_LBVH(hipMemcpyAsync(h_area.data(), areas, 2 * meshes * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<float> root_boxes(size_t(meshes) * 6);
    for (uint32_t m = 0; m < meshes; m++) {
        const MeshRun& r = runs[m];
        if (r.tri_count == 0) continue;
        const float* src = r.tri_count > 1 ? internal_boxes + 6 * size_t(r.node_base) : leaf_boxes + 6 * size_t(r.tri_offset);
        TT_LBVH(hipMemcpyAsync(&root_boxes[6 * size_t(m)], src, 6 * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    TT_LBVH(hipStreamSynchronize(st));
#undef TT_LBVH
    for (uint32_t m = 0; m < meshes; m++) {
        const MeshRun& r = runs[m];
        BlasSummary& b = summary[m];
        if (r.tri_count == 0) continue;
        b.empty = false;
        b.root = r.tri_count > 1 ? r.node_base : (BVH_LEAF | r.tri_offset);
        for (int a = 0; a < 3; a++) {
            b.lo[a] = root_boxes[6 * size_t(m) + a];
            b.hi[a] = root_boxes[6 * size_t(m) + 3 + a];
        }
        b.depth = std::max(h_depth[m], 1u);
        double dx = double(b.hi[0]) - b.lo[0], dy = double(b.hi[1]) - b.lo[1], dz = double(b.hi[2]) - b.lo[2];
        double root_area = 2.0 * (dx * dy + dy * dz + dz * dx);
        double ai = h_area[m], al = h_area[meshes + m];
        if (root_area > 0.0) {
            b.cost = (0.5 * ai + al) / root_area;
            // Wide nodes are the even-depth half of the binary ones; their
            // area is taken as half of the binary nodes' total.
            b.bytes = (0.5 * ai * double(sizeof(BvhWide)) + al * double(sizeof(BvhTri))) / root_area;
        }
    }
    return done(TT_OK);
}

}  // namespace tt
