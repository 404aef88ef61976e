/*
 * CPU oracle for libCL's triangle-mesh BVH (oclBvhTrimesh) and the particle-vs-mesh collision kernel
 * clCollideBVH of the 3D SPH program.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libcl_amd's BVH build and collision kernels. The product
 * library never links or calls it. Every function cites the reference file:line it restates; the
 * parity pinning of each is listed in the header comment of bvh_oracle.c.
 */
#ifndef BVH_ORACLE_H
#define BVH_ORACLE_H

#include <stdint.h>

#include "sph_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BVHNode, libCL/geom/oclBvhTrimesh.h:33-41 and libCL/geom/oclBvhTrimesh.cl:105-113 (48 bytes). */
typedef struct {
    float bbMin[4];
    float bbMax[4];
    int32_t left;
    int32_t right;
    uint32_t bit;
    uint32_t trav;
} orc_bvh_node;

/* code(), libCL/geom/oclBvhTrimesh.cl:78-85 */
uint32_t orc_morton_code(uint32_t x);
/* clAABB, libCL/geom/oclBvhTrimesh.cl:31-72, over a fresh box */
void orc_bvh_aabb(const float* vertex4, uint32_t nv, float* min4, float* max4);
/* clMorton, libCL/geom/oclBvhTrimesh.cl:87-99 */
void orc_bvh_morton(const uint32_t* index, const float* vertex4, uint32_t ntri, const float* min4,
                    const float* max4, uint32_t* key, uint32_t* val);
/* oclBvhTrimesh::compute, libCL/geom/oclBvhTrimesh.cpp:53-162: nodes has 2 ntri - 1 entries, key/val
   (ntri entries each) receive the sorted Morton keys and triangle indices. Returns 0 when ntri < 2 or
   a triangle index is not below nv. */
int orc_bvh_build(const float* vertex4, uint32_t nv, const uint32_t* index, uint32_t nindex, orc_bvh_node* nodes,
                  uint32_t* root, uint32_t* key, uint32_t* val);
/* clCollideBVH, libCL/phys/oclFluid3D.cl:498-550, for particles [0, n): hit[p] (optional) receives the
   leaf node that stopped the walk or 0xFFFFFFFF, filled[p] (optional) 1 where the 256-entry ring
   filled up (which ends the walk, as in the reference). */
void orc_collide_bvh(float* force4, float* pos4, float* vel4, uint32_t n, const orc_params* p,
                     const orc_bvh_node* nodes, uint32_t root, const uint32_t* tri, const float* vertex4,
                     uint32_t* hit, uint32_t* filled);

#ifdef __cplusplus
}
#endif
#endif
