/*
 * mfl.h - C ABI of libmfl.so, the MI355X (gfx950) implementation of libCL's 3D SPH step
 * (oclFluid3D) and key/value radix sort (oclRadixSort).
 *
 * Every entry point names the reference interface it replaces (paths relative to the reference root).
 * Conventions, following the reference's (libCL/oclCommon.cpp:19,39-46, libCL/oclProgram.h:146-150):
 *   - int-returning calls return 1 on success and 0 on failure; the failure text is available from
 *     mfl_last_error() (thread local) instead of the global sStatusCL + Log(ERR) stream.
 *   - device pointers are plain `void*` / `uint32_t*` into device memory of the program's device;
 *     `stream` is a hipStream_t (NULL = the null stream). compute calls only enqueue work, the
 *     caller synchronises (mfl_device_finish is clFinish).
 *   - element layouts are the reference's: float4 (16 B, AoS) positions / velocities / forces,
 *     uint32 cell keys and indices, two uint32[262144] bucket tables.
 */
#ifndef MFL_H
#define MFL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFL_ABI_VERSION 1

/* libCL/phys/oclFluid3D.h:48-63 (13 x fp32, this order; 52 bytes) */
typedef struct mfl_params {
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
} mfl_params;

typedef struct mfl_fluid3d mfl_fluid3d;
typedef struct mfl_radixsort mfl_radixsort;

/* Buffer selectors for mfl_fluid3d_{get,set,read,write}_buffer.
   POSITION/VELOCITY/FORCE: libCL/phys/oclFluid3D.h:34-40 (get/set{Position,Velocity,Force}Buffer);
   SORTED_*, INDEX, PARAMS: oclFluid3D.h:42-46; CELL, CELL_START, CELL_END: the internal bfCell /
   bfCellStart / bfCellEnd of oclFluid3D.h:92-94, exposed read-only for parity checks. */
enum mfl_buffer_kind {
    MFL_POSITION = 0,
    MFL_VELOCITY = 1,
    MFL_FORCE = 2,
    MFL_SORTED_POSITION = 3,
    MFL_SORTED_VELOCITY = 4,
    MFL_INDEX = 5,
    MFL_CELL = 6,
    MFL_CELL_START = 7,
    MFL_CELL_END = 8,
    MFL_PARAMS = 9,
    /* uint32[N] by sorted slot: the number of accepted visits of each particle in the last density pass
       (the `count` of libCL/phys/oclFluid3D.cl:179-182 that decides count > 2), written only while
       mfl_fluid3d_set_record_counts is on. Read-only. */
    MFL_NEIGHBOUR_COUNT = 10
};

/* ---- runtime ------------------------------------------------------------------------------------ */

/* Last failure text of this thread ("" if none). Replaces sStatusCL/oclError()
   (libCL/oclCommon.cpp:19,39-46) and the sticky Log(ERR) (libCL/oclObject.cpp:55-58). */
const char* mfl_last_error(void);
void mfl_clear_error(void);
int mfl_abi_version(void);
/* Selects the HIP device later create calls bind to (oclContext::getDevice(i), libCL/oclContext.h). */
int mfl_set_device(int ordinal);
/* clFinish on the program's queue (libCL/TEST/main.cpp:217): hipDeviceSynchronize. */
int mfl_device_finish(void);
/* Blocking copy between any two host / device addresses (oclBuffer::read / write,
   libCL/oclBuffer.cpp:47-83, for buffers that are not program-owned). */
int mfl_memcpy(void* dst, const void* src, size_t bytes);
/* Device memory of `bytes` bytes on the current device for caller-owned buffers (oclBuffer::create,
   libCL/oclBuffer.h:61, without the host mapping), NULL on failure; mfl_device_free releases it
   (~oclBuffer). */
void* mfl_device_alloc(size_t bytes);
void mfl_device_free(void* ptr);

/* ---- params ------------------------------------------------------------------------------------- */

/* Defaults of clInitFluid (libCL/phys/oclFluid3D.cl:130-148), evaluated on the host. */
void mfl_params_default(mfl_params* p);
/* The least float T with sqrtf(T) >= h: the neighbour passes accept a visit at squared distance r2
   when r2 < T, which for IEEE (correctly rounded, non-decreasing) sqrt is exactly sqrt(r2) < h, the
   library's form of the reference's fast_length(d) < h (libCL/phys/oclFluid3D.cl:177-178). Host only;
   no counterpart in the reference. */
float mfl_sqrt_cut(float h);
/* kernelConstant, viscosityConstant, pressureConstant from cellSize and viscosity, with the
   expressions of libCL/phys/oclFluid3D.cl:124-128,145-147. Never applied implicitly. */
void mfl_params_derive(mfl_params* p);

/* ---- oclRadixSort (libCL/sort/oclRadixSort.h:26-28) ---------------------------------------------- */

/* oclRadixSort::oclRadixSort(oclContext&, oclProgram*) (libCL/sort/oclRadixSort.cpp:27-48) */
mfl_radixsort* mfl_radixsort_create(void);
void mfl_radixsort_destroy(mfl_radixsort* s);
/* oclProgram::compile (libCL/oclProgram.cpp:58-120): kernels ship precompiled for gfx950; this
   verifies the device can run them. */
int mfl_radixsort_compile(mfl_radixsort* s);
/* oclRadixSort::compute(dev, key, val, startBit, endBit) (libCL/sort/oclRadixSort.cpp:62-144):
   stable sort of n (key, value) pairs in place by bits [start_bit, end_bit) of the key. Rejects
   (end_bit - start_bit) % 4 != 0 like the reference (cpp:76-80); the reference's n < 4,194,304 cap
   (cpp:70-74) is lifted (n < 2^32). n == 0 or start_bit >= end_bit is a no-op. */
int mfl_radixsort_compute(mfl_radixsort* s, uint32_t* keys, uint32_t* vals, size_t n, int start_bit,
                          int end_bit, void* stream);

/* No counterpart in the reference. mfl_radixsort_compute that allocates nothing, for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): the sort's temporaries live in a fixed device array of the library (18 MiB + 2 KiB,
   held by every GPU the library runs on until the process ends) instead of buffers of `s`. Same result,
   checks and return convention as mfl_radixsort_compute; n above MFL_SORT_NOALLOC_MAX_PAIRS is rejected.
   Every call on one GPU shares that array: calls that overlap in time on one GPU (other streams, other
   threads) race; calls in one stream, or separated by synchronisation, do not. */
#define MFL_SORT_NOALLOC_MAX_PAIRS 1048576u
int mfl_radixsort_compute_noalloc(mfl_radixsort* s, uint32_t* keys, uint32_t* vals, size_t n, int start_bit,
                                  int end_bit, void* stream);

/* ---- oclFluid3D (libCL/phys/oclFluid3D.h:25-74) -------------------------------------------------- */

/* oclFluid3D::oclFluid3D (libCL/phys/oclFluid3D.cpp:21-83): owns position/velocity/force buffers for
   the default 256 particles, bucket tables initialised to 0xFFFFFFFF (the reference leaves them
   uninitialised), reference box [-1,1]^3 and gravity (0,0,1,0) (cpp:256-268). */
mfl_fluid3d* mfl_fluid3d_create(void);
/* oclFluid3D::~oclFluid3D (cpp:85-90): frees program-owned buffers only. */
void mfl_fluid3d_destroy(mfl_fluid3d* f);
/* oclFluid3D::compile (cpp:273-285): loads clInitFluid's default Params (host copy + device
   param buffer). */
int mfl_fluid3d_compile(mfl_fluid3d* f);
/* oclFluid3D::compute (cpp:293-342): one SPH step enqueued on `stream`: hash, sort, reorder, bucket
   clear, bounds, density, forces, then either the EVT_INTEGRATE callback or clip-box + integrate +
   gravity + integrate-velocity. Fails (like a rejected NDRange) unless the particle count is a
   non-zero multiple of 256. */
int mfl_fluid3d_compute(mfl_fluid3d* f, void* stream);
/* oclFluid3D::setParticleCount (cpp:92-124): re-creates every program-owned buffer (contents lost);
   user buffers are not resized (the reference resizes them with a WARN): a later compute fails if a
   user buffer is smaller than the new count. */
int mfl_fluid3d_set_particle_count(mfl_fluid3d* f, size_t n);
/* oclFluid3D::getParticleCount (cpp:126-129) */
size_t mfl_fluid3d_get_particle_count(mfl_fluid3d* f);
/* set{Position,Velocity,Force}Buffer (cpp:131-165): `ptr` is caller-owned device memory of `bytes`
   bytes (>= 16 * particle count); ownership stays with the caller. Only POSITION, VELOCITY, FORCE. */
int mfl_fluid3d_set_buffer(mfl_fluid3d* f, int kind, void* ptr, size_t bytes);
/* get*Buffer (cpp:171-201): device pointer and byte size of any buffer kind. */
int mfl_fluid3d_get_buffer(mfl_fluid3d* f, int kind, void** ptr, size_t* bytes);
/* oclBuffer::write / read of a program buffer (libCL/oclBuffer.cpp:47-83), blocking. */
int mfl_fluid3d_write_buffer(mfl_fluid3d* f, int kind, const void* host, size_t bytes);
int mfl_fluid3d_read_buffer(mfl_fluid3d* f, int kind, void* host, size_t bytes);
/* oclFluid3D::getParameters (cpp:204-207): a host struct, valid at any time, read by every compute
   call (the device param buffer is refreshed from it on each step). */
mfl_params* mfl_fluid3d_params(mfl_fluid3d* f);
/* Addition: the clip box (hard-coded [-1,1]^3 in bindBuffers, cpp:260-268). */
int mfl_fluid3d_set_box(mfl_fluid3d* f, const float* min4, const float* max4);
/* Addition: the clGravity direction argument (hard-coded (0,0,1,0), cpp:256-258). */
int mfl_fluid3d_set_gravity(mfl_fluid3d* f, const float* dir4);
/* Addition: how many leading bucket entries are reset to 0xFFFFFFFF each step. Default 65,536,
   which is what the reference's byte-sized fill of an element count resets (cpp:309-313);
   262,144 resets the whole table. */
int mfl_fluid3d_set_bucket_clear(mfl_fluid3d* f, uint32_t count);
/* Addition: replay compute() from a captured hipGraph (re-captured whenever the particle count,
   buffers, params, box, gravity or bucket clear change). Ignored while an integrate callback is set. */
int mfl_fluid3d_set_use_graph(mfl_fluid3d* f, int enable);
/* Addition for parity checks: while enabled, every density pass also writes each particle's number of
   accepted visits (integer, the `count` of libCL/phys/oclFluid3D.cl:179-182) to the
   MFL_NEIGHBOUR_COUNT buffer. Off by default (4 bytes per particle more per step). */
int mfl_fluid3d_set_record_counts(mfl_fluid3d* f, int enable);

/* Addition: the order in which the density and force passes visit particles. Visit sets, counts and
   acceptance do not depend on it; the LIST pass's sums do not either (one thread per particle, in the
   reference's visit order), the tile passes' sums only in FP32 summation order (the particles that
   share a staged neighbourhood set how a particle's visits split over its lanes). It is there for
   cache locality: SORTED walks the hash-sorted arrays, CALLER walks the caller's particle order, MORTON
   walks a Morton order of the particles' cells rebuilt every `refresh_steps` steps (0 = once).
   Until this is called, mfl_fluid3d_set_particle_count picks the order by size: SORTED up to 131,072
   particles (their state fits one XCD's L2, so the hash order costs no locality, and it never goes
   stale between refreshes), MORTON with a 16-step refresh above. A call also makes the next compute
   rebuild the Morton order (after a caller rewrites the state, for instance). */
enum mfl_traversal { MFL_TRAVERSAL_SORTED = 0, MFL_TRAVERSAL_CALLER = 1, MFL_TRAVERSAL_MORTON = 2 };
int mfl_fluid3d_set_traversal(mfl_fluid3d* f, int order, uint32_t refresh_steps);
/* The current order and refresh period. */
int mfl_fluid3d_get_traversal(mfl_fluid3d* f, int* order, uint32_t* refresh_steps);

/* Addition: launch shape of the neighbour passes. Results differ between shapes only in FP32
   summation order (visit sets, counts and acceptance are identical).
   LIST: the density pass records every particle's accepted visits, the force pass reads the record.
   TILE: both passes stage the 27 bucket ranges of each cell's particles in LDS (`tile_depth` entries
   per wave, `block_waves` waves per workgroup), keeping only entries that can lie within h of the
   particles; a neighbourhood that keeps more than tile_depth entries is walked from global memory.
   TILE_LIST: the TILE density pass also records the visits, and the LIST force pass reads them.
   TILE_KEPT: the TILE density pass records each cell's kept entries and every lane's acceptance bits;
   the force pass stages the kept entries in LDS (64 at a time) and sums the accepted ones. The default
   is TILE_KEPT, 2 lanes, depth 96, 2 waves per workgroup. */
enum mfl_neighbour_pass { MFL_PASS_LIST = 0, MFL_PASS_TILE = 1, MFL_PASS_TILE_LIST = 2, MFL_PASS_TILE_KEPT = 3 };
typedef struct mfl_tuning {
    int32_t neighbour_pass; /* mfl_neighbour_pass */
    int32_t density_lanes;  /* lanes per particle in the density pass: LIST 1 or 8; the tile passes 2, 4 or 8 */
    uint32_t tile_depth;    /* tile passes: LDS entries per staged cell, multiple of 16 in 64..64 x lanes */
    uint32_t block_waves;   /* tile passes: waves per workgroup, 1, 2 or 4 */
    uint32_t sort_items;    /* keys per thread of each radix sort pass: 4, 8 or 16; 0 = by key count */
    uint32_t force_chunks;  /* LIST, TILE_LIST force pass: visits per trip / 4, 1 or 2 */
} mfl_tuning;
int mfl_fluid3d_get_tuning(mfl_fluid3d* f, mfl_tuning* out);
int mfl_fluid3d_set_tuning(mfl_fluid3d* f, const mfl_tuning* t);
/* Times `steps` eager steps of each candidate shape from the program's current state, restoring the
   state (positions, velocities, forces, bucket tables) before each trial and at the end, and keeps the
   fastest; `best` and `ms_per_step` (either may be NULL) receive it. Needs an allocated state
   (compile + particle count); enqueues on `stream` and synchronises it. */
int mfl_fluid3d_autotune(mfl_fluid3d* f, void* stream, uint32_t steps, mfl_tuning* best, double* ms_per_step);

/* Per-stage kernel timing with HIP events on the compute stream: the analogue of oclKernel::profile(true)
   + getStartTime/getEndTime (libCL/oclKernel.cpp:84-102,167-203). `stage_mask` has bit s set for each
   MFL_STAGE_* to time (0 disables). While a mask is set, compute() launches eagerly (no graph replay). */
enum mfl_stage {
    MFL_STAGE_HASH = 0,
    MFL_STAGE_SORT = 1,
    MFL_STAGE_REORDER = 2,
    MFL_STAGE_BOUNDS = 3,
    MFL_STAGE_DENSITY = 4,
    MFL_STAGE_FORCES = 5,
    MFL_STAGE_ORDER = 6,
    MFL_STAGE_COUNT = 7
};
int mfl_fluid3d_set_profiling(mfl_fluid3d* f, uint32_t stage_mask);
/* Waits for the recorded events and returns the accumulated time (ms) and launch count of a stage
   since the last reset; reset != 0 clears the accumulators afterwards. */
int mfl_fluid3d_stage_time(mfl_fluid3d* f, int stage, double* total_ms, uint64_t* launches, int reset);

/* addEventHandler(srtEvent named EVT_INTEGRATE = "OnIntegrate") (cpp:19,287-291,325-328): when set,
   compute() calls fn after the force pass instead of clip-box + integration. NULL removes it. The
   callback's return value is ignored, as in the reference. */
typedef int (*mfl_integrate_fn)(mfl_fluid3d* f, void* stream, void* user);
int mfl_fluid3d_set_integrate_callback(mfl_fluid3d* f, mfl_integrate_fn fn, void* user);
/* getKernel(name) + clEnqueueNDRangeKernel with the program's bound arguments (libCL/oclProgram.h:78,
   libCL/TEST/main.cpp:282-302). Names: clHash, clReorder, clFindBounds, clCalculateDensity,
   clCalculateForces, clClipBox, clIntegrateForce, clGravity, clIntegrateVelocity. */
int mfl_fluid3d_enqueue_kernel(mfl_fluid3d* f, const char* name, void* stream);

/* ---- triangle-mesh BVH and particle collision (oclBvhTrimesh, clCollideBVH) ---------------------- */

/* BVHNode (libCL/geom/oclBvhTrimesh.h:33-41, libCL/geom/oclBvhTrimesh.cl:105-113), 48 bytes: internal
   nodes 0 .. T-2 (bit = highest differing bit of the neighbouring sorted Morton keys, trav = 2 after
   a build), leaves T-1 .. 2T-2 (left = right = 0, bit = triangle index, trav = 0). */
typedef struct mfl_bvh_node {
    float bbMin[4];
    float bbMax[4];
    int32_t left;
    int32_t right;
    uint32_t bit;
    uint32_t trav;
} mfl_bvh_node;
typedef struct mfl_bvh_trimesh mfl_bvh_trimesh;
/* oclBvhTrimesh::oclBvhTrimesh / ~oclBvhTrimesh (libCL/geom/oclBvhTrimesh.cpp:19-51) */
mfl_bvh_trimesh* mfl_bvh_trimesh_create(void);
void mfl_bvh_trimesh_destroy(mfl_bvh_trimesh* b);
/* oclProgram::compile: device check */
int mfl_bvh_trimesh_compile(mfl_bvh_trimesh* b);
/* oclBvhTrimesh::compute(device, bfVertex, bfIndex) (cpp:53-162): T = index_count / 3 triangles of
   `index` (uint32 vertex indices, device) over `vertex` (vertex_count float4, device): global box of
   all vertices, 30-bit Morton key of each triangle's centroid, stable 32-bit key sort, 2T-1 nodes,
   root. Enqueues on `stream`, then waits for it and reads the root back (the reference maps
   bfBvhRoot, cpp:156-157). Fails for T < 2 or a triangle index not below vertex_count. The global box
   starts empty on every build (the reference's bfAABB is never initialised and accumulates). Where
   Morton keys repeat (the reference's linking races there) ties link to the left neighbour. */
int mfl_bvh_trimesh_compute(mfl_bvh_trimesh* b, const void* vertex, size_t vertex_count, const uint32_t* index,
                            size_t index_count, void* stream);
/* getNodeBuffer / getNodeCount / getRootNode (cpp:164-177): device array of mfl_bvh_node */
int mfl_bvh_trimesh_get_node_buffer(mfl_bvh_trimesh* b, void** ptr, size_t* bytes);
uint32_t mfl_bvh_trimesh_get_node_count(mfl_bvh_trimesh* b);
uint32_t mfl_bvh_trimesh_get_root_node(mfl_bvh_trimesh* b);
/* bfMortonKey / bfMortonVal after the sort (device, T entries each) */
int mfl_bvh_trimesh_get_morton(mfl_bvh_trimesh* b, void** keys, void** vals, size_t* count);

/* clCollideBVH (libCL/phys/oclFluid3D.cl:498-550) with its kernel arguments: forceOut, position,
   velocity (n caller-order float4 each, device), param (host copy), the tree of `bvh` (its node
   buffer and root, getNodeBuffer / getRootNode) and the mesh it was built from: triangle (index
   buffer, index_count entries) and vertex (vertex_count float4). A triangle or vertex index outside
   the counts ends that particle's walk without a change (the reference reads past the buffers).
   Breadth-first walk with the reference's 256-entry ring; the first triangle that collides moves the
   particle, zeroes its velocity and adds the contact force. The walk keeps 16 ring slots per particle
   in LDS and redoes the rare walks that need more with all 256; the list of those lives in `bvh`, so
   calls with one BVH on concurrent streams race. Fails if `bvh` has not been built. */
int mfl_collide_bvh(void* force, void* position, void* velocity, size_t n, const mfl_params* p, mfl_bvh_trimesh* bvh,
                    const uint32_t* triangle, size_t index_count, const void* vertex, size_t vertex_count,
                    void* stream);
/* The same kernel bound to a program's position / velocity / force buffers and Params, as an
   EVT_INTEGRATE handler enqueues it (getKernel("clCollideBVH") + the program's buffers). */
int mfl_fluid3d_collide_bvh(mfl_fluid3d* f, mfl_bvh_trimesh* bvh, const uint32_t* triangle, size_t index_count,
                            const void* vertex, size_t vertex_count, void* stream);

/* ---- oclFluid3Dnext (libCL/phys/oclFluid3Dnext.{h,cpp,cl}) ---------------------------------------- */

/* System of libCL/phys/oclFluid3Dnext.cl:15-30 (Params, oclFluid3Dnext.h:46-61, names field 2
   kernelRadius): 9 fp32 fields, then the grid size as two int32. */
typedef struct mfl_next_params {
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
} mfl_next_params;
/* clInitFluid (oclFluid3Dnext.cl:95-113): dt 0.0333333333, density 10000, stiffness 0.004 / 0.01,
   viscosity 1 / 1, radius 12, cellSize 2, grid 512 x 384, particleRadius 0.87 (1/density)^(1/3) */
void mfl_next_params_default(mfl_next_params* p);
/* buffers: the state (x, y, vx, vy as float4, caller order) is the program's "position" buffer
   (setPositionBuffer, cpp:113-123); pressure and force are float2 per sorted particle */
enum mfl_next_buffer {
    MFL_NEXT_POSITION = 0,
    MFL_NEXT_SORTED_STATE = 1,
    MFL_NEXT_INDEX = 2,
    MFL_NEXT_CELL = 3,
    MFL_NEXT_CELL_START = 4,
    MFL_NEXT_CELL_END = 5,
    MFL_NEXT_PRESSURE = 6,
    MFL_NEXT_FORCE = 7,
    MFL_NEXT_PARAMS = 8
};
typedef struct mfl_fluid3dnext mfl_fluid3dnext;
/* oclFluid3Dnext::oclFluid3Dnext / ~oclFluid3Dnext (cpp:19-82); 16,384 particles initially (cpp:53) */
mfl_fluid3dnext* mfl_fluid3dnext_create(void);
void mfl_fluid3dnext_destroy(mfl_fluid3dnext* f);
/* compile (cpp:231-255): clInitFluid defaults, cell tables of cellCountX * cellCountY entries */
int mfl_fluid3dnext_compile(mfl_fluid3dnext* f);
/* setParticleCount / getParticleCount (cpp:85-111) */
int mfl_fluid3dnext_set_particle_count(mfl_fluid3dnext* f, size_t n);
size_t mfl_fluid3dnext_get_particle_count(mfl_fluid3dnext* f);
/* getParameters (cpp:150-153): live host copy, applied at the next compute (the grid tables grow when
   cellCountX * cellCountY does; the reference keeps its compile-time tables) */
mfl_next_params* mfl_fluid3dnext_params(mfl_fluid3dnext* f);
int mfl_fluid3dnext_set_buffer(mfl_fluid3dnext* f, int kind, void* ptr, size_t bytes);
int mfl_fluid3dnext_get_buffer(mfl_fluid3dnext* f, int kind, void** ptr, size_t* bytes);
int mfl_fluid3dnext_write_buffer(mfl_fluid3dnext* f, int kind, const void* host, size_t bytes);
int mfl_fluid3dnext_read_buffer(mfl_fluid3dnext* f, int kind, void* host, size_t bytes);
/* oclFluid3Dnext::compute (cpp:264-317): clHash, sort of bits [0, 24), clReorder, clInitBounds,
   clFindBounds, clComputePressure, clComputeForces, clIntegrateForce. Fails for a particle count that
   is not a non-zero multiple of 256 (cLocalSize). A cell index outside the grid is not recorded in
   the bounds (the reference writes past its tables). */
int mfl_fluid3dnext_compute(mfl_fluid3dnext* f, void* stream);

/* ---- batched scenes (addition; no counterpart in the reference) --------------------------------- */

/* Many independent oclFluid3D systems advanced together: one set of launches per step for all scenes
   (optionally one hipGraph replay), scene-offset keys and one segmented sort. Scene s owns particles
   [offset_s, offset_s + count_s) of the concatenated float4 position / velocity / force buffers, its
   own Params, box and pair of 2^18-entry bucket tables (scene-local indices, as a single-scene program
   holds them), and its state after every step equals that of a single-scene program run alone.
   counts[s] must be non-zero multiples of 256. Buffer kinds as for mfl_fluid3d_*: CELL_START and
   CELL_END are num_scenes * 262144 entries, INDEX holds batch-global caller indices. */
typedef struct mfl_fluid_batch mfl_fluid_batch;
mfl_fluid_batch* mfl_batch_create(uint32_t num_scenes, const uint32_t* counts);
void mfl_batch_destroy(mfl_fluid_batch* b);
int mfl_batch_scene_range(mfl_fluid_batch* b, uint32_t scene, uint32_t* offset, uint32_t* count);
/* scene < 0 applies to every scene */
int mfl_batch_set_params(mfl_fluid_batch* b, int scene, const mfl_params* p);
int mfl_batch_get_params(mfl_fluid_batch* b, uint32_t scene, mfl_params* p);
int mfl_batch_set_box(mfl_fluid_batch* b, int scene, const float* min4, const float* max4);
int mfl_batch_set_gravity(mfl_fluid_batch* b, const float* dir4);
int mfl_batch_set_bucket_clear(mfl_fluid_batch* b, uint32_t count);
int mfl_batch_set_use_graph(mfl_fluid_batch* b, int enable);
int mfl_batch_get_buffer(mfl_fluid_batch* b, int kind, void** ptr, size_t* bytes);
int mfl_batch_write_buffer(mfl_fluid_batch* b, int kind, const void* host, size_t bytes);
int mfl_batch_read_buffer(mfl_fluid_batch* b, int kind, void* host, size_t bytes);
int mfl_batch_compute(mfl_fluid_batch* b, void* stream);
/* Launch shape of the batched step, as mfl_fluid3d_*_tuning: MFL_PASS_LIST (group traversal in sorted
   order, density_lanes 8, recorded-visit forces), MFL_PASS_TILE_LIST or MFL_PASS_TILE_KEPT (the tile
   density pass in a Morton order of each scene's cells, refreshed every 16 steps; density_lanes 4 or 8;
   recorded-visit or kept-entry forces). The autotuner times the candidates from the batch's current
   state and leaves that state unchanged. */
int mfl_batch_get_tuning(mfl_fluid_batch* b, mfl_tuning* out);
int mfl_batch_set_tuning(mfl_fluid_batch* b, const mfl_tuning* t);
int mfl_batch_autotune(mfl_fluid_batch* b, void* stream, uint32_t steps, mfl_tuning* best, double* ms_per_step);

/* ---- adjoint of a rollout (addition; no counterpart in the reference) ---------------------------- */

/* Reverse-mode derivative of `steps` compute() steps of a program, with checkpoints every
   `checkpoint_every` steps and recomputation in between. The sort permutation and the neighbour visits
   of each step (duplicates and stale-range visits included) are replayed from the forward as fixed
   data. list_cap (rounded up to a multiple of 4) bounds the recorded visits per particle (backward fails with an error beyond it).
   Gradients: dL/dx_0 and dL/dv_0 (caller-order float4 device buffers), and dL/d viscosity,
   stiffness, rest density (FP64). The program must keep its particle count and buffers meanwhile. */
enum mfl_adjoint_buffer { MFL_ADJ_POSITION_GRAD = 0, MFL_ADJ_VELOCITY_GRAD = 1 };
typedef struct mfl_adjoint mfl_adjoint;
mfl_adjoint* mfl_adjoint_create(mfl_fluid3d* f, uint32_t steps, uint32_t checkpoint_every, uint32_t list_cap);
void mfl_adjoint_destroy(mfl_adjoint* a);
/* runs the rollout from the program's current state (the program ends in the final state) */
int mfl_adjoint_forward(mfl_adjoint* a, void* stream);
/* pos_bar / vel_bar: dL/dx_T, dL/dv_T as caller-order float4 device arrays (NULL = zero) */
int mfl_adjoint_backward(mfl_adjoint* a, const void* pos_bar, const void* vel_bar, void* stream);
int mfl_adjoint_get_buffer(mfl_adjoint* a, int which, void** ptr, size_t* bytes);
/* out[0..3] = dL/d viscosity, dL/d stiffness, dL/d rest density, dL/d viscosityConstant */
int mfl_adjoint_param_grads(mfl_adjoint* a, double* out4);

#ifdef __cplusplus
}
#endif
#endif /* MFL_H */
