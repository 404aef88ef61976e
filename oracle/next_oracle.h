/*
 * CPU oracle for libCL's oclFluid3Dnext program (a 2D particle step on a dense cell grid).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of libcl_amd's oclFluid3Dnext kernels; the product never links
 * it. Every function cites the reference file:line it restates; the parity pinning is listed in the
 * header comment of next_oracle.c.
 */
#ifndef NEXT_ORACLE_H
#define NEXT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* System, libCL/phys/oclFluid3Dnext.cl:15-30 (oclFluid3Dnext.h:46-61 calls field 2 kernelRadius) */
typedef struct {
    float deltaTime;
    float particleRadius;
    float density;
    float stiffness0;
    float stiffness1;
    float viscosity0;
    float viscosity1;
    float radius;
    float cellSize;
    int32_t cellCountX;
    int32_t cellCountY;
} orc_next_params;

/* clInitFluid, libCL/phys/oclFluid3Dnext.cl:95-113 */
void orc_next_init_params(orc_next_params* p);
/* W, libCL/phys/oclFluid3Dnext.cl:116-133 */
float orc_next_w(float h, float q);
/* clHash, cl:36-44 */
void orc_next_hash(const float* state4, uint32_t n, const orc_next_params* p, uint32_t* cell, uint32_t* index);
/* clReorder, cl:46-51 */
void orc_next_reorder(const uint32_t* index, const float* state4, float* sorted4, uint32_t n);
/* clInitBounds + clFindBounds, cl:53-92, over ncells entries (a cell index >= ncells is not recorded) */
void orc_next_bounds(const uint32_t* cell, uint32_t n, uint32_t ncells, uint32_t* start, uint32_t* end);
/* clComputePressure, cl:135-181 (pressure: float2 per sorted particle) */
void orc_next_pressure(const float* sorted4, float* pressure2, const uint32_t* start, const uint32_t* end, uint32_t n,
                       const orc_next_params* p);
/* clComputeForces, cl:183-238 (force: float2 per sorted particle) */
void orc_next_forces(const float* sorted4, const float* pressure2, float* force2, const uint32_t* start,
                     const uint32_t* end, uint32_t n, const orc_next_params* p);
/* clIntegrateForce, cl:251-263 (state in caller order) */
void orc_next_integrate(float* state4, const float* force2, const uint32_t* index, uint32_t n,
                        const orc_next_params* p);
/* oclFluid3Dnext::compute, libCL/phys/oclFluid3Dnext.cpp:264-317, on caller arrays; cell/index/sorted4/
   pressure2/force2 are scratch of n entries, start/end hold cellCountX * cellCountY entries. */
void orc_next_step(float* state4, uint32_t n, const orc_next_params* p, uint32_t* cell, uint32_t* index,
                   float* sorted4, uint32_t* start, uint32_t* end, float* pressure2, float* force2);

#ifdef __cplusplus
}
#endif
#endif
