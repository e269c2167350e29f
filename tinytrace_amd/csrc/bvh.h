// Two-level BVH in the layout the gfx950 traversal kernel reads.
//
// Stands in for the driver's acceleration-structure build that the reference
// requests per mesh (BLAS, crates/tinytrace_backend/src/resource.rs:804-878,
// called from src/scene.rs:130-162) and over instances (TLAS,
// resource.rs:1014-1083, called from src/scene.rs:164-171).
//
// HBM layout
//   nodes : BvhNode[]  64 B, both children's boxes in the parent so one
//                      16-B-aligned fetch of 4 x dwordx4 tests two boxes
//                      (Aila & Laine 2009).  Every BLAS first, then the
//                      TLAS.
//   tris  : BvhTri[]   48 B, object-space vertices dequantized from snorm16
//                      (v / 32767), in leaf order; v0.w holds the mesh-local
//                      primitive index.
//   insts : BvhInstance[] 64 B, the instance's f32 inverse_transform columns
//                      (xyz); c0.w the BLAS root reference, c1.w the mesh's
//                      first global triangle (index_offset / 3).
//   boxes : 2 x float4 per instance, its padded world box (lo.xyz, hi.xyz);
//                      every hit on the instance lies inside, so the
//                      conservative slab interval of a ray against it bounds
//                      the hit distances of all its triangles (shadow rays).
// Node references (u32): bit 31 clear = inner node index; set = leaf.  TLAS
// leaves hold one instance (bits 0..30); BLAS leaves hold a triangle range:
// first slot in bits 0..26, count-1 in bits 27..30 (count <= 16).
//
// The builders produce these binary nodes; the scene then collapses them on
// the device into the 4-wide nodes traversal reads (bvh_wide.hip):
//   wide  : BvhWide[]  128 B, one cache line: the four children's boxes as
//                      SoA (lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4],
//                      hi.z[4]) and their four references.  Every binary
//                      node at even depth below its tree's root becomes one
//                      wide node whose children are its grandchildren (or
//                      its children where those are leaves).  Boxes are
//                      copied unchanged; unused slots hold +inf boxes, which
//                      no ray enters, and the reference 0xffffffff.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tinytrace.h"

namespace tt {

struct BvhNode {
    float n0[4];  // c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y
    float n1[4];  // c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y
    float n2[4];  // c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z
    uint32_t n3[4];  // ref0, ref1, 0, 0
};
static_assert(sizeof(BvhNode) == 64, "node size");

struct BvhWide {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(BvhWide) == 128, "wide node size");

struct BvhTri {
    float v0[4];
    float v1[4];
    float v2[4];
};
static_assert(sizeof(BvhTri) == 48, "triangle size");

struct BvhInstance {
    float c0[4];
    float c1[4];
    float c2[4];
    float c3[4];
};
static_assert(sizeof(BvhInstance) == 64, "instance size");

constexpr uint32_t BVH_LEAF = 0x80000000u;
constexpr uint32_t BVH_MAX_LEAF = 16;
// Depth at which splits switch to object medians; bounds every tree's depth
// by BVH_MEDIAN_DEPTH + 27 and so the traversal stack.
constexpr uint32_t BVH_MEDIAN_DEPTH = 20;
constexpr uint32_t BVH_STACK_LIMIT = 96;
// The 4-wide traversal pushes at most 3 references per wide level; a
// two-level path of binary depth D <= BVH_STACK_LIMIT - 2 has at most
// D / 2 + 2 wide levels, plus the instance sentinel.
constexpr uint32_t BVH_WIDE_STACK_LIMIT = 3 * ((BVH_STACK_LIMIT - 2) / 2 + 2) + 1;  // 148

struct BvhBuildResult {
    std::vector<BvhNode> nodes;
    std::vector<BvhTri> tris;
    std::vector<BvhInstance> instances;
    std::vector<float> instance_boxes;  // 8 per instance: lo.xyz, 0, hi.xyz, 0
    uint32_t tlas_root = BVH_LEAF;  // BVH_LEAF | 0x7fffffff = empty scene
    bool empty = true;
    uint32_t max_depth = 0;
    double sah_cost = 0.0;
    double expected_bytes = 0.0;
    std::string error;
};

// Host binned-SAH build of every BLAS and the TLAS.  The statistics
// (sah_cost, expected_bytes) describe the 4-wide tree the device makes of
// it: binary nodes at even depth cost one 128-B wide node visit, those at
// odd depth nothing.
bool build_bvh(const tt_scene_desc& desc, BvhBuildResult& out);

// One mesh's BLAS as the TLAS build needs it.
struct BlasSummary {
    bool empty = true;
    uint32_t root = BVH_LEAF;      // global node / leaf reference
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // object-space bounds (unpadded)
    uint32_t depth = 0;
    double cost = 0.0, bytes = 0.0;  // SAH expectation relative to the root box
};

// TLAS, instance records, world boxes and statistics over the given BLASes;
// appends the TLAS nodes to out.nodes, numbered from node_base on.
bool build_tlas(const tt_scene_desc& desc, const std::vector<BlasSummary>& blas, uint32_t node_base,
                BvhBuildResult& out);

}  // namespace tt
