/*
 * CPU oracle for the libCL 3D SPH step (oclFluid3D) and its key/value radix sort (oclRadixSort).
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference kernels, used by the
 * tests as the checker and by bench.py as the timed CPU baseline ("port"). The product library
 * (libcl_amd/libmfl.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference root).
 * Parity pinning (what checks each function independently of this file) is listed in the header
 * comment of sph_oracle.c.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* libCL/phys/oclFluid3D.h:48-63 and libCL/phys/oclFluid3D.cl:15-30: 13 fp32 fields, this order. */
typedef struct {
    float deltaTime;
    float particleRadius;
    float cellSize;
    float mass;
    float viscosity;
    float pressure;
    float density;
    float spacing;
    float stiffness;
    float viscosityConstant;
    float pressureConstant;
    float kernelConstant;
    float velocitylimit;
} orc_params;

#define ORC_BUCKETS 262144u          /* libCL/phys/oclFluid3D.cl:32 */
#define ORC_EMPTY 0xFFFFFFFFu        /* libCL/phys/oclFluid3D.cpp:309 */
#define ORC_REF_CLEAR 65536u         /* libCL/phys/oclFluid3D.cpp:310-312: 262,144 bytes of uint */
#define ORC_SORT_MAX 4194304u        /* libCL/sort/oclRadixSort.cpp:24 (BLOCK^3*4/16) */

/* clInitFluid defaults, libCL/phys/oclFluid3D.cl:130-148 */
void orc_init_params(orc_params* p);
/* kernelConstant / viscosityConstant / pressureConstant from cellSize and viscosity,
   libCL/phys/oclFluid3D.cl:124-128,145-147 */
void orc_derive_constants(orc_params* p);

/* libCL/phys/oclFluid3D.cl:34-52 */
uint32_t orc_hash_key(int gx, int gy, int gz);
void orc_grid_position(const float* pos4, float cell_size, int* g3);

/* clHash, libCL/phys/oclFluid3D.cl:57-64 */
void orc_hash(const float* pos4, uint32_t n, float cell_size, uint32_t* cell, uint32_t* index);

/* oclRadixSort::compute, libCL/sort/oclRadixSort.cpp:62-144 with kernels
   libCL/sort/oclRadixSort.cl:20-222. Returns 1 on success, 0 on the reference's rejections. */
int orc_radix_sort(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit);
/* same algorithm without the reference's N < 4,194,304 limit (for large-N baselines) */
int orc_radix_sort_unbounded(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit);
/* stable merge sort of (key,val) by bits [start_bit,end_bit) of key: an independent checker */
void orc_stable_sort(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit);

/* clReorder (fluid), libCL/phys/oclFluid3D.cl:66-72 */
void orc_reorder(const uint32_t* index, const float* pos4, const float* vel4, float* spos4, float* svel4,
                 uint32_t n);
/* clEnqueueFillBuffer x2, libCL/phys/oclFluid3D.cpp:309-313 (first clear_count entries only) */
void orc_clear_buckets(uint32_t* cell_start, uint32_t* cell_end, uint32_t clear_count);
/* clFindBounds, libCL/phys/oclFluid3D.cl:85-113 */
void orc_find_bounds(const uint32_t* cell, uint32_t n, uint32_t* cell_start, uint32_t* cell_end);
/* clCalculateDensity, libCL/phys/oclFluid3D.cl:150-200; counts (may be NULL) receives each sorted
   particle's number of accepted visits */
void orc_density(float* spos4, float* svel4, const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n,
                 const orc_params* p, uint32_t* counts);
/* clCalculateForces, libCL/phys/oclFluid3D.cl:202-255 */
void orc_forces(const float* spos4, const float* svel4, float* force4, const uint32_t* index,
                const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n, const orc_params* p);
/* clClipBox, libCL/phys/oclFluid3D.cl:304-361 */
void orc_clip_box(const float* spos4, const float* svel4, float* force4, const uint32_t* index, uint32_t n,
                  const orc_params* p, const float* min4, const float* max4);
/* clIntegrateForce, libCL/phys/oclFluid3D.cl:258-271 */
void orc_integrate_force(float* vel4, const float* force4, uint32_t n, const orc_params* p);
/* clGravity, libCL/phys/oclFluid3D.cl:294-298 with the (0,0,1,0) argument of oclFluid3D.cpp:256-258 */
void orc_gravity(float* force4, uint32_t n, const float* gravity4);
/* clIntegrateVelocity, libCL/phys/oclFluid3D.cl:284-292 */
void orc_integrate_velocity(float* pos4, const float* vel4, uint32_t n, const orc_params* p);

/* One oclFluid3D::compute() (libCL/phys/oclFluid3D.cpp:293-342) on caller-owned arrays.
   cell/index/spos/svel are scratch of length n; cell_start/cell_end hold ORC_BUCKETS entries and
   carry state between steps exactly like the reference's bucket tables.
   integrate = 0 stops after the force pass (the EVT_INTEGRATE callback path, cpp:325-328). */
typedef struct {
    uint32_t n;
    orc_params params;
    float box_min[4];
    float box_max[4];
    float gravity[4];
    uint32_t clear_count;
    float* pos;
    float* vel;
    float* force;
    uint32_t* cell;
    uint32_t* index;
    float* spos;
    float* svel;
    uint32_t* cell_start;
    uint32_t* cell_end;
    uint32_t* counts; /* may be NULL: accepted visits per sorted particle of the last density pass */
} orc_fluid;

int orc_step(orc_fluid* f, int integrate);
int orc_steps(orc_fluid* f, int nsteps);

/* number of OpenMP threads the oracle uses (1 when built without OpenMP) */
int orc_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
