/*
 * CPU oracle for the libCL 3D SPH step and key/value radix sort.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load this file's library, and only as the checker / CPU baseline. The product path is the HIP
 * library in libcl_amd/ and never routes through here.
 *
 * Restated from (paths relative to the reference root):
 *   libCL/phys/oclFluid3D.cl:15-361  (kernels; clEmitter :273-282 is never enqueued and is skipped)
 *   libCL/phys/oclFluid3D.cpp:213-342 (argument binding, box and gravity values, enqueue order, fill size)
 *   libCL/sort/oclRadixSort.cl:20-222, libCL/sort/oclRadixSort.cpp:62-144 (4-bit LSD radix sort)
 *
 * Arithmetic follows the kernels' expression order in IEEE fp32 (built with -ffp-contract=off).
 * OpenCL's fast_length is taken as sqrt(dot) and pow(x, 3.0f) as powf; the reference pins no
 * runtime for these builtins, so their last-ulp behaviour is parity-unpinned and absorbed by the
 * 1e-5 field tolerance.
 *
 * How each function is pinned without the reference running (tests/test_oracle_*.py):
 *   orc_hash_key / orc_grid_position : hand-derived known answers (tests/golden/hash_kat.json) computed
 *                                      with Python integers, plus negative / cell-boundary cases.
 *   orc_radix_sort                   : numpy stable argsort (an independent library routine) on
 *                                      random, duplicate-heavy, partial-block and sub-range inputs;
 *                                      the reference's rejections (size cap, bit range % 4).
 *   orc_find_bounds + clear          : first / last+1 occurrence via numpy searchsorted, stale-bucket
 *                                      rule checked on a two-step hand-built case.
 *   orc_density / orc_forces         : O(N^2) fp64 brute force in numpy (tests/golden/ fixtures written by
 *                                      tests/golden/make_golden.py), a 4-particle tetrahedron closed
 *                                      form, the count<=2 fallback, momentum conservation.
 *   orc_clip_box / integrate / gravity: per-wall closed forms, the clamp, the exact gravity bit pattern.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define GROUP 256 /* work-group size of every reference NDRange (oclFluid3D.cpp:16) */

int orc_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* ---------------------------------------------------------------- params */

/* libCL/phys/oclFluid3D.cl:119-128: M_PI_F is the float nearest pi; the double literals make the
   outer arithmetic double on fp64-capable devices, the pow() calls take and return float. */
static const float kPiF = 3.14159274101257324219f;

void orc_derive_constants(orc_params* p)
{
    /* DensityConstant(h) = 315.0/(64.0*M_PI_F*pow(h,9.0f))   (cl:124, used at cl:145) */
    double h9 = (double)powf(p->cellSize, 9.0f);
    p->kernelConstant = (float)(315.0 / (64.0 * (double)kPiF * h9));
    /* ViscosityConstant(h) = (-45.0/(M_PI_F * pow(h, 6.0f)))   (cl:128): M_PI_F*pow() is a float product */
    float pih6 = kPiF * powf(p->cellSize, 6.0f);
    double visc_c = -45.0 / (double)pih6;
    p->viscosityConstant = (float)(-(double)p->viscosity * visc_c); /* cl:146 */
    p->pressureConstant = (float)(-visc_c);                          /* cl:126,147 (same expression) */
}

void orc_init_params(orc_params* p)
{
    /* libCL/phys/oclFluid3D.cl:132-147 */
    memset(p, 0, sizeof(*p));
    p->deltaTime = 0.005f;
    p->mass = 0.006f;
    p->density = 1000.0f;
    p->viscosity = 6.0f;
    p->pressure = 0.0f;
    p->velocitylimit = 500.0f;
    p->stiffness = 1.0f;
    p->spacing = (float)(0.87 * (double)powf(p->mass / p->density, 1.0f / 3.0f));
    p->particleRadius = (float)(0.5 * (double)p->spacing);
    p->cellSize = (float)(2.0 * (double)p->spacing);
    orc_derive_constants(p);
}

/* ---------------------------------------------------------------- hash */

void orc_grid_position(const float* pos4, float cell_size, int* g3)
{
    /* libCL/phys/oclFluid3D.cl:34-42: (int)floor(p / cellSize) per axis, IEEE divide */
    g3[0] = (int)floorf(pos4[0] / cell_size);
    g3[1] = (int)floorf(pos4[1] / cell_size);
    g3[2] = (int)floorf(pos4[2] / cell_size);
}

uint32_t orc_hash_key(int gx, int gy, int gz)
{
    /* libCL/phys/oclFluid3D.cl:44-52: uint products wrap mod 2^32, then % BUCKETS */
    uint32_t n = (73856093u * (uint32_t)gx) ^ (19349663u * (uint32_t)gy) ^ (83492791u * (uint32_t)gz);
    return n % ORC_BUCKETS;
}

void orc_hash(const float* pos4, uint32_t n, float cell_size, uint32_t* cell, uint32_t* index)
{
    /* libCL/phys/oclFluid3D.cl:57-64 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        int g[3];
        orc_grid_position(pos4 + 4 * i, cell_size, g);
        cell[i] = orc_hash_key(g[0], g[1], g[2]);
        index[i] = (uint32_t)i;
    }
}

/* ---------------------------------------------------------------- radix sort */

#define SORT_BITS 4  /* libCL/sort/oclRadixSort.cl:15, oclRadixSort.cpp:18 */
#define SORT_RADIX 16

/* One 4-bit pass of libCL/sort/oclRadixSort.cpp:95-142. */
static void radix_pass(uint32_t* key, uint32_t* val, uint32_t* tkey, uint32_t* tval, uint32_t n, int shift,
                       uint32_t* block_scan, uint32_t* block_offset)
{
    const int64_t nb = ((int64_t)n + GROUP - 1) / GROUP;

    /* clBlockSort (cl:20-112): per 256-block, four stable 1-bit splits from the low bit (a stable
       partition by the 4-bit digit); padding lanes hold 0xFFFFFFFF (cl:34-40) and are counted in the
       block histogram as digit 0xF. Digit start (blockOffset, 0 for absent digits) and count
       (blockScan, digit-major) are recorded. */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; b++) {
        uint32_t lk[GROUP], lv[GROUP];
        uint32_t cnt[SORT_RADIX] = {0}, first[SORT_RADIX], fill[SORT_RADIX];
        for (int t = 0; t < GROUP; t++) {
            int64_t g = b * GROUP + t;
            uint32_t k = 0xFFFFFFFFu, v = 0xFFFFFFFFu;
            if (g < (int64_t)n) {
                k = key[g];
                v = val[g];
            }
            lk[t] = k;
            lv[t] = v;
            cnt[(k >> shift) & 0xF]++;
        }
        uint32_t run = 0;
        for (int d = 0; d < SORT_RADIX; d++) {
            first[d] = run;
            fill[d] = run;
            run += cnt[d];
        }
        uint32_t sk[GROUP], sv[GROUP];
        for (int t = 0; t < GROUP; t++) {
            uint32_t d = (lk[t] >> shift) & 0xF;
            uint32_t a = fill[d]++;
            sk[a] = lk[t];
            sv[a] = lv[t];
        }
        for (int t = 0; t < GROUP; t++) {
            int64_t g = b * GROUP + t;
            if (g < (int64_t)n) {
                tkey[g] = sk[t];
                tval[g] = sv[t];
            }
        }
        for (int d = 0; d < SORT_RADIX; d++) {
            block_offset[b * SORT_RADIX + d] = cnt[d] ? first[d] : 0;
            block_scan[(int64_t)d * nb + b] = cnt[d];
        }
    }

    /* clBlockScan + clBlockPrefix (cl:114-187): exclusive scan of the digit-major count table. */
    uint32_t run = 0;
    for (int64_t i = 0; i < nb * SORT_RADIX; i++) {
        uint32_t c = block_scan[i];
        block_scan[i] = run;
        run += c;
    }

    /* clReorder (cl:189-222): dst = scan[d*nB+b] + tid - start[b*16+d] */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; b++) {
        for (int t = 0; t < GROUP; t++) {
            int64_t g = b * GROUP + t;
            if (g >= (int64_t)n) break;
            uint32_t k = tkey[g];
            uint32_t d = (k >> shift) & 0xF;
            uint32_t dst = block_scan[(int64_t)d * nb + b] + (uint32_t)t - block_offset[b * SORT_RADIX + d];
            key[dst] = k;
            val[dst] = tval[g];
        }
    }
}

static int radix_sort_impl(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit, int bounded)
{
    /* checks of libCL/sort/oclRadixSort.cpp:64-80 (key/val sizes agree by construction here) */
    if (bounded && n >= ORC_SORT_MAX) return 0;
    if ((end_bit - start_bit) % SORT_BITS != 0) return 0;
    if (n == 0 || start_bit >= end_bit) return 1;
    const size_t nb = ((size_t)n + GROUP - 1) / GROUP;
    uint32_t* tkey = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* tval = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* scan = (uint32_t*)malloc(sizeof(uint32_t) * nb * SORT_RADIX);
    uint32_t* offs = (uint32_t*)malloc(sizeof(uint32_t) * nb * SORT_RADIX);
    if (!tkey || !tval || !scan || !offs) {
        free(tkey); free(tval); free(scan); free(offs);
        return 0;
    }
    for (int j = start_bit; j < end_bit; j += SORT_BITS) /* cpp:95 */
        radix_pass(key, val, tkey, tval, n, j, scan, offs);
    free(tkey); free(tval); free(scan); free(offs);
    return 1;
}

int orc_radix_sort(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit)
{
    return radix_sort_impl(key, val, n, start_bit, end_bit, 1);
}

int orc_radix_sort_unbounded(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit)
{
    return radix_sort_impl(key, val, n, start_bit, end_bit, 0);
}

/* Independent checker: top-down stable merge sort on the masked key. */
static uint32_t sort_mask(int start_bit, int end_bit)
{
    int w = end_bit - start_bit;
    return w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1u);
}

void orc_stable_sort(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit)
{
    if (n < 2 || start_bit >= end_bit) return;
    const uint32_t mask = sort_mask(start_bit, end_bit);
    uint32_t* ak = key;
    uint32_t* av = val;
    uint32_t* bk = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* bv = (uint32_t*)malloc(sizeof(uint32_t) * n);
    /* bottom-up: runs of width w merged into runs of 2w, stable (left run wins ties) */
    for (uint64_t w = 1; w < n; w *= 2) {
        for (uint64_t lo = 0; lo < n; lo += 2 * w) {
            uint64_t mid = lo + w < n ? lo + w : n;
            uint64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            uint64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                uint32_t ki = (ak[i] >> start_bit) & mask, kj = (ak[j] >> start_bit) & mask;
                if (kj < ki) { bk[o] = ak[j]; bv[o++] = av[j++]; }
                else         { bk[o] = ak[i]; bv[o++] = av[i++]; }
            }
            while (i < mid) { bk[o] = ak[i]; bv[o++] = av[i++]; }
            while (j < hi)  { bk[o] = ak[j]; bv[o++] = av[j++]; }
        }
        uint32_t* t;
        t = ak; ak = bk; bk = t;
        t = av; av = bv; bv = t;
    }
    if (ak != key) {
        memcpy(key, ak, sizeof(uint32_t) * n);
        memcpy(val, av, sizeof(uint32_t) * n);
        free(ak); free(av);
    } else {
        free(bk); free(bv);
    }
}

/* ---------------------------------------------------------------- fluid kernels */

/* The squared length inside fast_length(vector) (cl:177, cl:232). The reference leaves the builtin to
   its OpenCL runtime; this restatement fixes it as fma(z, z, fma(y, y, x * x)), the form an OpenCL
   compiler contracting `x*x + y*y + z*z` produces (FP_CONTRACT is on by default in OpenCL C) and the
   one the HIP kernels use, so both sides take the same accept / reject decision for every visit. */
static inline float dot3(float x, float y, float z) { return fmaf(z, z, fmaf(y, y, x * x)); }

void orc_reorder(const uint32_t* index, const float* pos4, const float* vel4, float* spos4, float* svel4,
                 uint32_t n)
{
    /* libCL/phys/oclFluid3D.cl:66-72 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t s = index[i];
        memcpy(spos4 + 4 * i, pos4 + 4 * (size_t)s, 16);
        memcpy(svel4 + 4 * i, vel4 + 4 * (size_t)s, 16);
    }
}

void orc_clear_buckets(uint32_t* cell_start, uint32_t* cell_end, uint32_t clear_count)
{
    /* libCL/phys/oclFluid3D.cpp:309-313: the fill size argument is bytes, so the reference resets
       only the first 262,144 / 4 = 65,536 entries; clear_count generalises it. */
    if (clear_count > ORC_BUCKETS) clear_count = ORC_BUCKETS;
    for (uint32_t b = 0; b < clear_count; b++) {
        cell_start[b] = ORC_EMPTY;
        cell_end[b] = ORC_EMPTY;
    }
}

void orc_find_bounds(const uint32_t* cell, uint32_t n, uint32_t* cell_start, uint32_t* cell_end)
{
    /* libCL/phys/oclFluid3D.cl:85-113 (numParticles = global size = n) */
    for (uint32_t i = 0; i < n; i++) {
        uint32_t h = cell[i];
        if (i == 0) {
            cell_start[h] = 0;
            continue; /* cl:100-104 returns before the last-particle check */
        }
        uint32_t prev = cell[i - 1];
        if (h != prev) {
            cell_end[prev] = i;
            cell_start[h] = i;
        }
        if (i == n - 1) cell_end[h] = n;
    }
}

void orc_density(float* spos4, float* svel4, const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n,
                 const orc_params* p, uint32_t* counts)
{
    const float h = p->cellSize;
    /* libCL/phys/oclFluid3D.cl:150-200. Reads of position[j].w race with other items' writes in the
       reference too; only xyz of j are used. Two passes keep this deterministic: compute, then store. */
    float* rho = (float*)malloc(sizeof(float) * (n ? n : 1));
    float* prs = (float*)malloc(sizeof(float) * (n ? n : 1));
#pragma omp parallel for schedule(dynamic, GROUP)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float* pi = spos4 + 4 * i;
        int g[3];
        orc_grid_position(pi, h, g);
        float density = 0.0f;
        int count = 0;
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    uint32_t c = orc_hash_key(g[0] + x, g[1] + y, g[2] + z);
                    uint32_t start = cell_start[c];
                    if (start == ORC_EMPTY) continue;
                    /* ranges lie in [0, n] for any n the reference accepts (a multiple of 256);
                       the clamp only matters for n == 1, where cl:100-104 never writes the end */
                    uint32_t end = cell_end[c] < n ? cell_end[c] : n;
                    for (uint32_t j = start; j < end; j++) {
                        if (j == (uint32_t)i) continue;
                        const float* pj = spos4 + 4 * (size_t)j;
                        float vx = pi[0] - pj[0], vy = pi[1] - pj[1], vz = pi[2] - pj[2];
                        float vlen = sqrtf(dot3(vx, vy, vz)); /* fast_length, w = 0 */
                        if (vlen < h) {
                            density += powf(h * h - vlen * vlen, 3.0f); /* cl:123 */
                            count++;
                        }
                    }
                }
        if (counts) counts[i] = (uint32_t)count; /* accepted visits of particle i (cl:179-182) */
        if (count > 2) {
            float d = fmaxf(1.0f, p->mass * p->kernelConstant * density);
            rho[i] = d;
            prs[i] = p->pressure + p->stiffness * (d - p->density);
        } else {
            rho[i] = p->density;
            prs[i] = p->pressure;
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        spos4[4 * i + 3] = rho[i];
        svel4[4 * i + 3] = prs[i];
    }
    free(rho);
    free(prs);
}

void orc_forces(const float* spos4, const float* svel4, float* force4, const uint32_t* index,
                const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n, const orc_params* p)
{
    const float h = p->cellSize;
    /* libCL/phys/oclFluid3D.cl:202-255 */
#pragma omp parallel for schedule(dynamic, GROUP)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float* pi = spos4 + 4 * i;
        const float* vi = svel4 + 4 * i;
        float P[3] = {0, 0, 0}, V[3] = {0, 0, 0};
        int g[3];
        orc_grid_position(pi, h, g);
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    uint32_t c = orc_hash_key(g[0] + x, g[1] + y, g[2] + z);
                    uint32_t start = cell_start[c];
                    if (start == ORC_EMPTY) continue;
                    /* ranges lie in [0, n] for any n the reference accepts (a multiple of 256);
                       the clamp only matters for n == 1, where cl:100-104 never writes the end */
                    uint32_t end = cell_end[c] < n ? cell_end[c] : n;
                    for (uint32_t j = start; j < end; j++) {
                        if (j == (uint32_t)i) continue;
                        const float* pj = spos4 + 4 * (size_t)j;
                        float v[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                        float vlen = sqrtf(dot3(v[0], v[1], v[2])); /* fast_length, w = 0 */
                        if (vlen < h) {
                            const float* vj = svel4 + 4 * (size_t)j;
                            float density2 = pi[3] * pj[3];
                            float pressure2 = vi[3] + vj[3];
                            float hr = h - vlen;
                            float q = pressure2 / density2;
                            for (int c3 = 0; c3 < 3; c3++) {
                                /* pressure2/density2*PressureVariable(h, vector/vlen, vlen), cl:125,240 */
                                P[c3] += q * (((v[c3] / vlen) * hr) * hr);
                                /* (velJ-velI)/density2*ViscosityVariable(h, vlen), cl:127,242 */
                                V[c3] += ((vj[c3] - vi[c3]) / density2) * hr;
                            }
                        }
                    }
                }
        float* f = force4 + 4 * (size_t)index[i];
        for (int c3 = 0; c3 < 3; c3++) f[c3] += (p->pressureConstant * P[c3] + p->viscosityConstant * V[c3]) * p->mass;
        f[3] = 0.0f; /* cl:253 */
    }
}

#define BOUNDARY_EPSILON 0.00001f /* libCL/phys/oclFluid3D.cl:304 */
#define BOUNDARY_STIFFNESS 2000   /* cl:306 */
#define BOUNDARY_DAMPENING 256    /* cl:307 */

/* repulsionForce, cl:309-312, accumulated into f as the float4 sum of cl:327 etc. */
static void add_repulsion(float* f, const float* nrm, const float* vel, float dist)
{
    float dot = nrm[0] * vel[0] + nrm[1] * vel[1] + nrm[2] * vel[2] + nrm[3] * vel[3];
    float s = BOUNDARY_STIFFNESS * dist - BOUNDARY_DAMPENING * dot;
    for (int c = 0; c < 4; c++) f[c] += nrm[c] * s;
}

void orc_clip_box(const float* spos4, const float* svel4, float* force4, const uint32_t* index, uint32_t n,
                  const orc_params* p, const float* mn, const float* mx)
{
    static const float ny0[4] = {0, 1, 0, 0}, ny1[4] = {0, -1, 0, 0};
    static const float nz0[4] = {0, 0, 1, 0}, nz1[4] = {0, 0, -1, 0};
    static const float nx0[4] = {1, 0, 0, 0}, nx1[4] = {-1, 0, 0, 0};
    const float r = p->particleRadius;
    /* libCL/phys/oclFluid3D.cl:315-361: walls in the order y-, y+, z-, z+, x-, x+ */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float* pos = spos4 + 4 * i;
        const float* vel = svel4 + 4 * i;
        float* fo = force4 + 4 * (size_t)index[i];
        float f[4] = {fo[0], fo[1], fo[2], fo[3]};
        float diff;
        diff = r - (pos[1] - mn[1]);
        if (diff > BOUNDARY_EPSILON) add_repulsion(f, ny0, vel, diff);
        diff = r - (mx[1] - pos[1]);
        if (diff > BOUNDARY_EPSILON) add_repulsion(f, ny1, vel, diff);
        diff = r - (pos[2] - mn[2]);
        if (diff > BOUNDARY_EPSILON) add_repulsion(f, nz0, vel, diff);
        diff = r - (mx[2] - pos[2]);
        if (diff > BOUNDARY_EPSILON) add_repulsion(f, nz1, vel, diff);
        diff = r - (pos[0] - mn[0]);
        if (diff > BOUNDARY_EPSILON) add_repulsion(f, nx0, vel, diff);
        diff = r - (mx[0] - pos[0]);
        if (diff > BOUNDARY_EPSILON) add_repulsion(f, nx1, vel, diff);
        memcpy(fo, f, 16);
    }
}

void orc_integrate_force(float* vel4, const float* force4, uint32_t n, const orc_params* p)
{
    /* libCL/phys/oclFluid3D.cl:258-271: the clamp rescales a private copy; force is not written */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        float f[4];
        memcpy(f, force4 + 4 * i, 16);
        float speed = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3]);
        if (speed > p->velocitylimit) {
            float s = p->velocitylimit / speed;
            for (int c = 0; c < 4; c++) f[c] *= s;
        }
        for (int c = 0; c < 4; c++) vel4[4 * i + c] += f[c] * p->deltaTime;
    }
}

void orc_gravity(float* force4, uint32_t n, const float* gravity4)
{
    /* libCL/phys/oclFluid3D.cl:294-298: force = -9.8f * gravity */
    float g[4];
    for (int c = 0; c < 4; c++) g[c] = -9.8f * gravity4[c];
    for (uint32_t i = 0; i < n; i++) memcpy(force4 + 4 * (size_t)i, g, 16);
}

void orc_integrate_velocity(float* pos4, const float* vel4, uint32_t n, const orc_params* p)
{
    /* libCL/phys/oclFluid3D.cl:284-292 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++)
        for (int c = 0; c < 4; c++) pos4[4 * i + c] += vel4[4 * i + c] * p->deltaTime;
}

int orc_step(orc_fluid* f, int integrate)
{
    /* libCL/phys/oclFluid3D.cpp:293-342 */
    const uint32_t n = f->n;
    orc_hash(f->pos, n, f->params.cellSize, f->cell, f->index);              /* :295 */
    if (!orc_radix_sort_unbounded(f->cell, f->index, n, 0, 24)) return 0;    /* :299 */
    orc_reorder(f->index, f->pos, f->vel, f->spos, f->svel, n);              /* :304 */
    orc_clear_buckets(f->cell_start, f->cell_end, f->clear_count);           /* :309-313 */
    orc_find_bounds(f->cell, n, f->cell_start, f->cell_end);                 /* :315 */
    orc_density(f->spos, f->svel, f->cell_start, f->cell_end, n, &f->params, f->counts); /* :318 */
    orc_forces(f->spos, f->svel, f->force, f->index, f->cell_start, f->cell_end, n, &f->params); /* :321 */
    if (!integrate) return 1;                                                /* :325-328 */
    orc_clip_box(f->spos, f->svel, f->force, f->index, n, &f->params, f->box_min, f->box_max); /* :331 */
    orc_integrate_force(f->vel, f->force, n, &f->params);                    /* :333 */
    orc_gravity(f->force, n, f->gravity);                                    /* :335 */
    orc_integrate_velocity(f->pos, f->vel, n, &f->params);                   /* :337 */
    return 1;
}

int orc_steps(orc_fluid* f, int nsteps)
{
    for (int s = 0; s < nsteps; s++)
        if (!orc_step(f, 1)) return 0;
    return 1;
}
