// Triangle-mesh BVH (the oclBvhTrimesh program, libCL/geom/oclBvhTrimesh.h:21-82) and the particle
// versus mesh collision kernel clCollideBVH (libCL/phys/oclFluid3D.cl:365-550) on gfx950.
#pragma once

#include <vector>

#include "mfl_common.hpp"
#include "radix_sort.hpp"

namespace mfl {

// BVHNode, libCL/geom/oclBvhTrimesh.h:33-41: internal nodes 0 .. T-2 (bit = split bit), leaves
// T-1 .. 2T-2 (left = right = 0, bit = triangle index).
struct BvhNode {
    float4 bbMin;
    float4 bbMax;
    int32_t left;
    int32_t right;
    uint32_t bit;
    uint32_t trav;
};
static_assert(sizeof(BvhNode) == 48, "BVHNode is 48 bytes");
static_assert(sizeof(BvhNode) == sizeof(mfl_bvh_node), "mfl_bvh_node mirrors BVHNode");

class BvhTrimesh {
public:
    int compile();  // oclProgram::compile analogue (device check)
    // oclBvhTrimesh::compute (libCL/geom/oclBvhTrimesh.cpp:53-162): builds the tree of the
    // index_count / 3 triangles of `index` over `vertex` (vertex_count float4) on `s`, then waits for
    // the stream and reads the root back (the reference maps bfBvhRoot, cpp:156-157).
    int compute(const float4* vertex, size_t vertex_count, const uint32_t* index, size_t index_count, hipStream_t s);
    BvhNode* nodes() const { return nodes_.as<BvhNode>(); }
    uint32_t node_count() const { return count_; }
    uint32_t root() const { return root_; }
    uint32_t* morton_keys() const { return keys_.as<uint32_t>(); }
    uint32_t* morton_vals() const { return vals_.as<uint32_t>(); }

    // clCollideBVH (libCL/phys/oclFluid3D.cl:498-550) with this tree (nodes, root) over particles
    // [0, n) of caller-order position / velocity / force buffers and the mesh it was built from. A
    // triangle or vertex index outside the given counts ends that particle's walk without a change.
    // Uses this object's spill list: calls with one BVH on concurrent streams race.
    int collide(float4* force, float4* pos, float4* vel, size_t n, const mfl_params& p, const uint32_t* triangle,
                size_t index_count, const float4* vertex, size_t vertex_count, hipStream_t s);

private:
    bool compiled_ = false;
    RadixSort sort_;
    // leaf entries of the walk: tag | triangle (bit 15 while the tree fits 16-bit ring slots)
    uint32_t leaf_tag() const { return count_ <= 0xFFFFu ? 0x8000u : 0x80000000u; }

    DeviceBuffer part_, keys_, vals_, nodes_, split_, range_, info_, spill_, wide_, seg_, seg_meta_;
    std::vector<uint64_t> seg_host_;  // staging of the segment-tree level offsets and sizes
    uint32_t count_ = 0, root_ = 0;
};

}  // namespace mfl
