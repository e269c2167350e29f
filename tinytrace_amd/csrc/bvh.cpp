// Host binned-SAH builder for the two-level BVH described in bvh.h.
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <cmath>
#include <exception>
#include <mutex>
#include <thread>

// SAH cost of one traversal step relative to one triangle test, and the
// largest BLAS leaf.
#ifndef TT_SAH_TRAVERSAL_COST
#define TT_SAH_TRAVERSAL_COST 1.0
#endif
#ifndef TT_SAH_BINS
#define TT_SAH_BINS 16
#endif
#ifndef TT_BLAS_MAX_LEAF
#define TT_BLAS_MAX_LEAF 4
#endif

namespace tt {

namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* p) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool valid() const { return lo[0] <= hi[0]; }
    double area() const {
        if (!valid()) return 0.0;
        double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

// Outward padding: 2^-20 of the coordinate magnitude (covers the rounding of
// the watertight triangle test's vertex translation, a few ulp of |v|).
Box padded(const Box& b) {
    Box r = b;
    for (int a = 0; a < 3; a++) {
        r.lo[a] = b.lo[a] - (std::fabs(b.lo[a]) * 9.5367431640625e-07f + 1e-30f);
        r.hi[a] = b.hi[a] + (std::fabs(b.hi[a]) * 9.5367431640625e-07f + 1e-30f);
    }
    return r;
}

struct Prim {
    Box box;
    float centroid[3];
    uint32_t id;
};

struct Builder {
    std::vector<Prim>& prims;
    std::vector<BvhNode>& nodes;  // local node array (indices local)
    uint32_t max_leaf;
    uint32_t max_depth = 0;
    // Leaves in creation order: (first prim position, count) -> ref patch.
    std::vector<std::pair<uint32_t, uint32_t>> leaves;

    Builder(std::vector<Prim>& p, std::vector<BvhNode>& n, uint32_t ml) : prims(p), nodes(n), max_leaf(ml) {}

    // Returns a reference; leaves are encoded as BVH_LEAF | leaf ordinal for
    // now and patched by the caller.
    uint32_t build(uint32_t first, uint32_t count, uint32_t depth, Box& out_box) {
        max_depth = std::max(max_depth, depth);
        Box box, cbox;
        for (uint32_t i = first; i < first + count; i++) {
            box.grow(prims[i].box);
            cbox.grow(prims[i].centroid);
        }
        out_box = box;
        if (count == 1 || (count <= max_leaf && count <= 2)) return make_leaf(first, count);

        int axis = -1;
        uint32_t mid = 0;
        double leaf_cost = double(count);
        if (depth < BVH_MEDIAN_DEPTH) {
            const int BINS = TT_SAH_BINS;
            double best = INFINITY;
            int best_axis = -1, best_bin = 0;
            for (int a = 0; a < 3; a++) {
                float ext = cbox.hi[a] - cbox.lo[a];
                if (!(ext > 0.0f)) continue;
                float scale = float(BINS) / ext;
                Box bb[BINS];
                uint32_t bc[BINS] = {};
                for (uint32_t i = first; i < first + count; i++) {
                    int b = int((prims[i].centroid[a] - cbox.lo[a]) * scale);
                    b = std::min(std::max(b, 0), BINS - 1);
                    bb[b].grow(prims[i].box);
                    bc[b]++;
                }
                double ra[BINS];
                uint32_t rc[BINS];
                Box acc;
                uint32_t n = 0;
                for (int b = BINS - 1; b > 0; b--) {
                    acc.grow(bb[b]);
                    n += bc[b];
                    ra[b] = acc.area();
                    rc[b] = n;
                }
                acc = Box();
                n = 0;
                for (int b = 0; b < BINS - 1; b++) {
                    acc.grow(bb[b]);
                    n += bc[b];
                    if (n == 0 || rc[b + 1] == 0) continue;
                    double c = acc.area() * n + ra[b + 1] * rc[b + 1];
                    if (c < best) {
                        best = c;
                        best_axis = a;
                        best_bin = b;
                    }
                }
            }
            double parent_area = box.area();
            double split_cost = parent_area > 0.0 ? TT_SAH_TRAVERSAL_COST + best / parent_area : INFINITY;
            if (best_axis >= 0 && !(count <= max_leaf && leaf_cost <= split_cost)) {
                float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
                float scale = float(BINS) / ext;
                auto it = std::partition(prims.begin() + first, prims.begin() + first + count, [&](const Prim& p) {
                    int b = int((p.centroid[best_axis] - cbox.lo[best_axis]) * scale);
                    b = std::min(std::max(b, 0), BINS - 1);
                    return b <= best_bin;
                });
                mid = uint32_t(it - prims.begin());
                if (mid > first && mid < first + count) axis = best_axis;
            } else if (count <= max_leaf) {
                return make_leaf(first, count);
            }
        } else if (count <= max_leaf) {
            return make_leaf(first, count);
        }
        if (axis < 0) {
            // Object median on the widest centroid axis.
            int a = 0;
            float ext = -1.0f;
            for (int k = 0; k < 3; k++)
                if (cbox.hi[k] - cbox.lo[k] > ext) {
                    ext = cbox.hi[k] - cbox.lo[k];
                    a = k;
                }
            mid = first + count / 2;
            std::nth_element(prims.begin() + first, prims.begin() + mid, prims.begin() + first + count,
                             [&](const Prim& x, const Prim& y) {
                                 if (x.centroid[a] != y.centroid[a]) return x.centroid[a] < y.centroid[a];
                                 return x.id < y.id;
                             });
        }
        uint32_t index = uint32_t(nodes.size());
        nodes.emplace_back();
        Box lb, rb;
        uint32_t l = build(first, mid - first, depth + 1, lb);
        uint32_t r = build(mid, first + count - mid, depth + 1, rb);
        Box pl = padded(lb), pr = padded(rb);
        BvhNode& node = nodes[index];
        node.n0[0] = pl.lo[0];
        node.n0[1] = pl.hi[0];
        node.n0[2] = pl.lo[1];
        node.n0[3] = pl.hi[1];
        node.n1[0] = pr.lo[0];
        node.n1[1] = pr.hi[0];
        node.n1[2] = pr.lo[1];
        node.n1[3] = pr.hi[1];
        node.n2[0] = pl.lo[2];
        node.n2[1] = pl.hi[2];
        node.n2[2] = pr.lo[2];
        node.n2[3] = pr.hi[2];
        node.n3[0] = l;
        node.n3[1] = r;
        node.n3[2] = 0;
        node.n3[3] = 0;
        return index;
    }

    uint32_t make_leaf(uint32_t first, uint32_t count) {
        uint32_t ordinal = uint32_t(leaves.size());
        leaves.emplace_back(first, count);
        return BVH_LEAF | ordinal;
    }
};

// SAH statistics of a subtree: expected visits weighted by surface area.
struct Stats {
    double cost = 0.0;
    double bytes = 0.0;
};

template <typename LeafFn>
void subtree_stats(const std::vector<BvhNode>& nodes, uint32_t ref, uint32_t node_base, double p, LeafFn&& leaf,
                   Stats& s, uint32_t depth = 0) {
    // nodes[k] holds node node_base + k.
    // p = probability that this reference's box is hit (relative to root).
    // Binary nodes at even depth are the wide nodes traversal visits (one
    // 128-B fetch, four box tests); odd-depth ones are folded into them.
    if (ref & BVH_LEAF) {
        leaf(ref, p, s);
        return;
    }
    const BvhNode& n = nodes[ref - node_base];
    if (depth % 2 == 0) {
        s.cost += p * 1.0;
        s.bytes += p * double(sizeof(BvhWide));
    }
    auto area = [](float lx, float hx, float ly, float hy, float lz, float hz) {
        double dx = double(hx) - lx, dy = double(hy) - ly, dz = double(hz) - lz;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    };
    double a0 = area(n.n0[0], n.n0[1], n.n0[2], n.n0[3], n.n2[0], n.n2[1]);
    double a1 = area(n.n1[0], n.n1[1], n.n1[2], n.n1[3], n.n2[2], n.n2[3]);
    double parent = std::max(a0, a1);
    // Parent area from the union of the children.
    double ux = double(std::max(n.n0[1], n.n1[1])) - std::min(n.n0[0], n.n1[0]);
    double uy = double(std::max(n.n0[3], n.n1[3])) - std::min(n.n0[2], n.n1[2]);
    double uz = double(std::max(n.n2[1], n.n2[3])) - std::min(n.n2[0], n.n2[2]);
    parent = std::max(parent, 2.0 * (ux * uy + uy * uz + uz * ux));
    double q0 = parent > 0 ? a0 / parent : 1.0, q1 = parent > 0 ? a1 / parent : 1.0;
    subtree_stats(nodes, n.n3[0], node_base, p * q0, leaf, s, depth + 1);
    subtree_stats(nodes, n.n3[1], node_base, p * q1, leaf, s, depth + 1);
}

struct Blas {
    std::vector<BvhNode> nodes;
    std::vector<BvhTri> tris;
    uint32_t root = BVH_LEAF;  // local reference
    Box box;                   // object-space bounds (unpadded)
    uint32_t depth = 0;
    Stats stats;
    bool empty = true;
};

void build_blas(const tt_scene_desc& d, uint32_t mesh, Blas& out) {
    const tt_mesh_record& m = d.meshes[mesh];
    uint32_t tri_count = m.index_count / 3;
    if (tri_count == 0) return;
    std::vector<Prim> prims(tri_count);
    std::vector<BvhTri> raw(tri_count);
    for (uint32_t t = 0; t < tri_count; t++) {
        float v[3][3];
        for (int c = 0; c < 3; c++) {
            uint32_t vi = m.vertex_offset + d.indices[m.index_offset + 3 * t + c];
            for (int a = 0; a < 3; a++) v[c][a] = float(d.positions[3 * size_t(vi) + a]) / 32767.0f;
        }
        Prim& p = prims[t];
        for (int c = 0; c < 3; c++) p.box.grow(v[c]);
        for (int a = 0; a < 3; a++) p.centroid[a] = (p.box.lo[a] + p.box.hi[a]) * 0.5f;
        p.id = t;
        BvhTri& bt = raw[t];
        for (int a = 0; a < 3; a++) {
            bt.v0[a] = v[0][a];
            bt.v1[a] = v[1][a];
            bt.v2[a] = v[2][a];
        }
        uint32_t prim_bits = t;
        std::memcpy(&bt.v0[3], &prim_bits, 4);
        bt.v1[3] = 0.0f;
        bt.v2[3] = 0.0f;
    }
    Builder b(prims, out.nodes, TT_BLAS_MAX_LEAF);
    uint32_t root = b.build(0, tri_count, 0, out.box);
    // Lay out triangles in leaf order and patch leaf references.
    std::vector<uint32_t> leaf_ref(b.leaves.size());
    for (size_t l = 0; l < b.leaves.size(); l++) {
        uint32_t first = uint32_t(out.tris.size());
        uint32_t count = b.leaves[l].second;
        for (uint32_t k = 0; k < count; k++) out.tris.push_back(raw[prims[b.leaves[l].first + k].id]);
        leaf_ref[l] = BVH_LEAF | ((count - 1) << 27) | first;
    }
    auto patch = [&](uint32_t r) { return (r & BVH_LEAF) ? leaf_ref[r & 0x7fffffffu] : r; };
    for (auto& n : out.nodes) {
        n.n3[0] = patch(n.n3[0]);
        n.n3[1] = patch(n.n3[1]);
    }
    out.root = patch(root);
    out.depth = b.max_depth;
    out.empty = false;
    // Expected bytes/cost relative to the BLAS root box.
    subtree_stats(
        out.nodes, out.root, 0, 1.0,
        [&](uint32_t ref, double p, Stats& s) {
            double n = double(((ref >> 27) & 0xfu) + 1);
            s.cost += p * n;
            s.bytes += p * n * double(sizeof(BvhTri));
        },
        out.stats);
}

// Runs fn(i) for every i in [0, count) on up to max_threads host threads.
// An exception thrown by fn stops the work not yet handed out and is
// rethrown here once every thread has joined (one escaping a thread's
// function would end the process).
template <typename Fn>
void parallel_for(uint32_t count, unsigned max_threads, Fn&& fn) {
    std::atomic<uint32_t> next{0};
    std::exception_ptr error;
    std::mutex error_mutex;
    auto worker = [&] {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= count) break;
            try {
                fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> lock(error_mutex);
                if (!error) error = std::current_exception();
                next.store(count);
            }
        }
    };
    const unsigned nt = std::max(1u, std::min(max_threads, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    try {
        pool.reserve(nt);
        for (unsigned t = 0; t < nt; t++) pool.emplace_back(worker);
    } catch (...) {
        for (auto& t : pool) t.join();
        throw;
    }
    for (auto& t : pool) t.join();
    if (error) std::rethrow_exception(error);
}

// Object-space boxes that together bound mesh m: its triangles binned by
// centroid into a PART_GRID^3 grid (16^3) over the mesh's box, one box per
// non-empty cell around the vertices of the cell's triangles.  Every triangle
// lies inside its own cell's box, so the union of the boxes' world images
// bounds the instance; for a rotated mesh that union is much tighter than the
// world box of the whole mesh's box (the axis-aligned box of a rotated cube
// is up to sqrt(3) times wider on each axis).  Vertices are dequantized as the
// BLAS does (v / 32767 in float).
#ifndef TT_PART_GRID
#define TT_PART_GRID 16
#endif
constexpr uint32_t PART_GRID = TT_PART_GRID;

std::vector<Box> part_boxes(const tt_scene_desc& d, uint32_t mesh) {
    const tt_mesh_record& m = d.meshes[mesh];
    const uint32_t tri_count = m.index_count / 3;
    auto vertex = [&](uint32_t t, int c, float* v) {
        uint32_t vi = m.vertex_offset + d.indices[m.index_offset + 3 * t + c];
        for (int a = 0; a < 3; a++) v[a] = float(d.positions[3 * size_t(vi) + a]) / 32767.0f;
    };
    Box whole;
    for (uint32_t t = 0; t < tri_count; t++)
        for (int c = 0; c < 3; c++) {
            float v[3];
            vertex(t, c, v);
            whole.grow(v);
        }
    std::vector<Box> cells(PART_GRID * PART_GRID * PART_GRID);
    for (uint32_t t = 0; t < tri_count; t++) {
        float v[3][3];
        uint32_t cell = 0;
        for (int c = 0; c < 3; c++) vertex(t, c, v[c]);
        for (int a = 0; a < 3; a++) {
            float lo = std::min(v[0][a], std::min(v[1][a], v[2][a]));
            float hi = std::max(v[0][a], std::max(v[1][a], v[2][a]));
            float ext = whole.hi[a] - whole.lo[a];
            float u = ext > 0.0f ? (0.5f * (lo + hi) - whole.lo[a]) / ext : 0.0f;
            uint32_t k = uint32_t(std::min(std::max(u * float(PART_GRID), 0.0f), float(PART_GRID - 1)));
            cell = cell * PART_GRID + k;
        }
        for (int c = 0; c < 3; c++) cells[cell].grow(v[c]);
    }
    std::vector<Box> out;
    for (const Box& b : cells)
        if (b.valid()) out.push_back(b);
    return out;
}

// Bounds of the image of object-space box b under the instance transform
// (column-major 4x4), in double: centre M c + t, half-extent |M| h.
void grow_world(const float* T, const Box& b, double lo[3], double hi[3]) {
    for (int r = 0; r < 3; r++) {
        double c = double(T[12 + r]), h = 0.0;
        for (int k = 0; k < 3; k++) {
            double m = double(T[k * 4 + r]);
            c += m * 0.5 * (double(b.lo[k]) + double(b.hi[k]));
            h += std::fabs(m) * 0.5 * (double(b.hi[k]) - double(b.lo[k]));
        }
        lo[r] = std::min(lo[r], c - h);
        hi[r] = std::max(hi[r], c + h);
    }
}

}  // namespace

// TLAS over the instances, instance records, world boxes and the two-level
// SAH statistics, given every mesh's BLAS (roots as global references).  The
// TLAS nodes are appended to out.nodes and numbered from node_base on.
bool build_tlas(const tt_scene_desc& d, const std::vector<BlasSummary>& blas, uint32_t node_base,
                BvhBuildResult& out) {
    // Instance world boxes: the union of the world images of the mesh's part
    // boxes (in double), padded by 1e-5 of (|lo| + |hi| + extent) of the
    // image of the whole mesh's box, which contains the union.
    std::vector<std::vector<Box>> parts(d.mesh_count);
    {
        std::vector<uint8_t> used(d.mesh_count, 0);
        for (uint32_t i = 0; i < d.instance_count; i++)
            if (d.instances[i].mesh_index < d.mesh_count) used[d.instances[i].mesh_index] = 1;
        parallel_for(d.mesh_count, 16u, [&](uint32_t m) {
            if (used[m] && !blas[m].empty) parts[m] = part_boxes(d, m);
        });
    }
    std::vector<Prim> prims;
    std::vector<Box> world(d.instance_count);
    for (uint32_t i = 0; i < d.instance_count; i++) {
        const tt_instance& in = d.instances[i];
        if (in.mesh_index >= d.mesh_count) {
            out.error = "instance refers to a missing mesh";
            return false;
        }
        const BlasSummary& b = blas[in.mesh_index];
        if (b.empty) continue;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        Box whole;
        for (int a = 0; a < 3; a++) {
            whole.lo[a] = b.lo[a];
            whole.hi[a] = b.hi[a];
        }
        grow_world(in.transform, whole, lo, hi);
        double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        double pad[3];
        for (int r = 0; r < 3; r++) pad[r] = (std::fabs(lo[r]) + std::fabs(hi[r]) + ext) * 1e-5;
        const std::vector<Box>& pb = parts[in.mesh_index];
        if (!pb.empty()) {
            double plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (const Box& q : pb) grow_world(in.transform, q, plo, phi);
            for (int r = 0; r < 3; r++) {
                lo[r] = std::max(lo[r], plo[r]);
                hi[r] = std::min(hi[r], phi[r]);
            }
        }
        Prim p;
        for (int r = 0; r < 3; r++) {
            p.box.lo[r] = float(lo[r] - pad[r]);
            p.box.hi[r] = float(hi[r] + pad[r]);
            p.centroid[r] = float(0.5 * (lo[r] + hi[r]));
        }
        p.id = i;
        world[i] = p.box;
        prims.push_back(p);
    }

    // TLAS (one instance per leaf), after the BLAS nodes.
    std::vector<BvhNode> tlas_nodes;
    uint32_t tlas_root = BVH_LEAF | 0x7fffffffu;
    uint32_t tlas_depth = 0;
    Box tlas_box;
    if (!prims.empty()) {
        Builder b(prims, tlas_nodes, 1);
        uint32_t root = b.build(0, uint32_t(prims.size()), 0, tlas_box);
        std::vector<uint32_t> leaf_ref(b.leaves.size());
        for (size_t l = 0; l < b.leaves.size(); l++) leaf_ref[l] = BVH_LEAF | prims[b.leaves[l].first].id;
        auto patch = [&](uint32_t r) { return (r & BVH_LEAF) ? leaf_ref[r & 0x7fffffffu] : r + node_base; };
        for (auto& n : tlas_nodes) {
            n.n3[0] = patch(n.n3[0]);
            n.n3[1] = patch(n.n3[1]);
        }
        tlas_root = patch(root);
        tlas_depth = b.max_depth;
        out.empty = false;
    }
    std::vector<BvhNode> tlas_copy = tlas_nodes;
    out.nodes.insert(out.nodes.end(), tlas_copy.begin(), tlas_copy.end());
    uint32_t deepest = 0;
    for (const BlasSummary& b : blas)
        if (!b.empty) deepest = std::max(deepest, b.depth);
    out.tlas_root = tlas_root;
    out.max_depth = tlas_depth + deepest;
    if (out.max_depth + 2 > BVH_STACK_LIMIT) {
        out.error = "BVH too deep for the traversal stack";
        return false;
    }

    out.instances.resize(d.instance_count);
    for (uint32_t i = 0; i < d.instance_count; i++) {
        const tt_instance& in = d.instances[i];
        BvhInstance& bi = out.instances[i];
        for (int r = 0; r < 3; r++) {
            bi.c0[r] = in.inverse_transform[0 + r];
            bi.c1[r] = in.inverse_transform[4 + r];
            bi.c2[r] = in.inverse_transform[8 + r];
            bi.c3[r] = in.inverse_transform[12 + r];
        }
        uint32_t root = blas[in.mesh_index].empty ? (BVH_LEAF | 0x7fffffffu) : blas[in.mesh_index].root;
        std::memcpy(&bi.c0[3], &root, 4);
        // c1.w: first global triangle index of the mesh (index_offset / 3).
        uint32_t base = in.index_offset / 3;
        std::memcpy(&bi.c1[3], &base, 4);
        bi.c2[3] = 0.0f;
        bi.c3[3] = 0.0f;
    }
    // Padded world boxes (empty meshes keep lo = +inf > hi = -inf: no ray
    // enters them).
    out.instance_boxes.assign(size_t(d.instance_count) * 8, 0.0f);
    for (uint32_t i = 0; i < d.instance_count; i++) {
        float* b = &out.instance_boxes[size_t(i) * 8];
        for (int r = 0; r < 3; r++) {
            b[r] = world[i].lo[r];
            b[4 + r] = world[i].hi[r];
        }
    }

    // Two-level SAH expectation: TLAS nodes, then per instance leaf the
    // instance record plus the BLAS expectation.
    Stats s;
    if (!out.empty) {
        subtree_stats(
            tlas_copy, tlas_root, node_base, 1.0,
            [&](uint32_t ref, double p, Stats& st) {
                uint32_t i = ref & 0x7fffffffu;
                const BlasSummary& b = blas[d.instances[i].mesh_index];
                st.cost += p * (1.0 + b.cost);
                st.bytes += p * (double(sizeof(BvhInstance)) + b.bytes);
            },
            s);
    }
    out.sah_cost = s.cost;
    out.expected_bytes = s.bytes;
    return true;
}

bool build_bvh(const tt_scene_desc& d, BvhBuildResult& out) {
    out = BvhBuildResult();
    uint64_t total_tris = 0;
    for (uint32_t m = 0; m < d.mesh_count; m++) total_tris += d.meshes[m].index_count / 3;
    if (total_tris >= (1ull << 27)) {
        out.error = "too many triangles for the 27-bit leaf encoding";
        return false;
    }
    std::vector<Blas> blas(d.mesh_count);
    parallel_for(d.mesh_count, 32u, [&](uint32_t m) { build_blas(d, m, blas[m]); });
    // Concatenate every BLAS (references rebased), then the TLAS.
    std::vector<BlasSummary> summary(d.mesh_count);
    for (uint32_t m = 0; m < d.mesh_count; m++) {
        Blas& b = blas[m];
        BlasSummary& bs = summary[m];
        bs.empty = b.empty;
        if (b.empty) continue;
        uint32_t node_base = uint32_t(out.nodes.size());
        uint32_t tri_base = uint32_t(out.tris.size());
        auto rebase = [&](uint32_t r) {
            if (r & BVH_LEAF) return BVH_LEAF | (r & (0xfu << 27)) | ((r & 0x7ffffffu) + tri_base);
            return r + node_base;
        };
        for (auto n : b.nodes) {
            n.n3[0] = rebase(n.n3[0]);
            n.n3[1] = rebase(n.n3[1]);
            out.nodes.push_back(n);
        }
        out.tris.insert(out.tris.end(), b.tris.begin(), b.tris.end());
        bs.root = rebase(b.root);
        for (int a = 0; a < 3; a++) {
            bs.lo[a] = b.box.lo[a];
            bs.hi[a] = b.box.hi[a];
        }
        bs.depth = b.depth;
        bs.cost = b.stats.cost;
        bs.bytes = b.stats.bytes;
    }
    return build_tlas(d, summary, uint32_t(out.nodes.size()), out);
}

}  // namespace tt
