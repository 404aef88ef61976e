/*
 * Plain-C restatement of libCL's oclFluid3Dnext step (libCL/phys/oclFluid3Dnext.cpp:264-317 with the
 * kernels of libCL/phys/oclFluid3Dnext.cl:36-263): a 2D particle state (x, y, vx, vy) hashed into a
 * dense cellCountX x cellCountY grid, a density sum with a cubic spline over the 3 x 3 neighbouring
 * cells, a Tait pressure, the force of the last neighbour visited (the reference assigns, it does not
 * accumulate: cl:221-223), gravity, and a clamped symplectic Euler update.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of libcl_amd's oclFluid3Dnext kernels; the product never
 * links it.
 *
 * Definitions the OpenCL sources leave open, fixed here and in libcl_amd/csrc/fluid3dnext.hip:
 * length(v) = sqrt(v.x v.x + v.y v.y); no contraction of multiply-adds; the double constants of the
 * kernels (0.87, 0.25, 9.81, clInitFluid's literals) are applied as C applies them (0.87 * (float)
 * in double, then rounded to float; 0.25 and 9.81 rounded to float first, as OpenCL converts a scalar
 * to a vector's element type); pow(x, 7.0f) is the C library's powf (the GPU's powf differs by at most
 * an ulp or two, which the tests' tolerance covers); clamp(x, lo, hi) = fmin(fmax(x, lo), hi); a cell
 * index outside [0, cellCountX * cellCountY) is not recorded in the bounds (the reference writes past
 * its tables there).
 *
 * Parity pinning (tests/test_next_oracle.py): W against the closed-form cubic spline and its
 * normalisation (integral of W over the line = 1); the hash against numpy floor division; the bounds
 * against numpy first/last occurrence; the density against an fp64 all-pairs sum restricted to the
 * 3 x 3 cell block; the pressure against the Tait formula in fp64; the force against an independent
 * search for the last neighbour in visit order; the update against its closed form.
 */
#include "next_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

#define NEXT_EMPTY 0xFFFFFFFFu

/* clInitFluid, libCL/phys/oclFluid3Dnext.cl:95-113 (double literals rounded into the float fields;
   particleRadius evaluates 0.87 * pow(1.0 / density, 1 / 3.0f) in double) */
void orc_next_init_params(orc_next_params* p)
{
    memset(p, 0, sizeof(*p));
    p->deltaTime = (float)0.0333333333;
    p->density = 10000.0f;
    p->stiffness0 = (float)0.004;
    p->stiffness1 = (float)0.01;
    p->viscosity0 = 1.0f;
    p->viscosity1 = 1.0f;
    p->radius = 12.0f;
    p->cellSize = 2.0f;
    p->cellCountX = (int32_t)(1024 / p->cellSize);
    p->cellCountY = (int32_t)(768 / p->cellSize);
    p->particleRadius = (float)(0.87 * pow(1.0 / (double)p->density, (double)(1 / 3.0f)));
}

/* W, libCL/phys/oclFluid3Dnext.cl:116-133 */
float orc_next_w(float h, float q)
{
    float value = 0.0f;
    q /= h;
    if (q < 1.0f) {
        value = 1.0f - 1.5f * q * q + 0.75f * q * q * q;
    } else if (q < 2.0f) {
        q = 2.0f - q;
        value = 0.25f * q * q * q;
    }
    return 2.0f / (3.0f * h) * value;
}

/* clHash, cl:36-44: cell = gy * cellCountX + gx in int arithmetic (two's complement wrap) */
void orc_next_hash(const float* state4, uint32_t n, const orc_next_params* p, uint32_t* cell, uint32_t* index)
{
    for (uint32_t i = 0; i < n; i++) {
        const int32_t gx = (int32_t)floorf(state4[4 * (size_t)i] / p->cellSize);
        const int32_t gy = (int32_t)floorf(state4[4 * (size_t)i + 1] / p->cellSize);
        cell[i] = (uint32_t)gy * (uint32_t)p->cellCountX + (uint32_t)gx;
        index[i] = i;
    }
}

/* clReorder, cl:46-51 */
void orc_next_reorder(const uint32_t* index, const float* state4, float* sorted4, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) memcpy(sorted4 + 4 * (size_t)i, state4 + 4 * (size_t)index[i], 16);
}

/* clInitBounds + clFindBounds, cl:53-92 */
void orc_next_bounds(const uint32_t* cell, uint32_t n, uint32_t ncells, uint32_t* start, uint32_t* end)
{
    for (uint32_t c = 0; c < ncells; c++) start[c] = end[c] = NEXT_EMPTY;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t h = cell[i];
        if (i == 0) {
            if (h < ncells) start[h] = 0;
            continue;
        }
        const uint32_t prev = cell[i - 1];
        if (h != prev) {
            if (prev < ncells) end[prev] = i;
            if (h < ncells) start[h] = i;
        }
        if (i == n - 1 && h < ncells) end[h] = n;
    }
}

static void cell_window(float x, float y, const orc_next_params* p, int* x0, int* x1, int* y0, int* y1)
{
    const int gx = (int)floorf(x / p->cellSize), gy = (int)floorf(y / p->cellSize);
    *x0 = gx - 1 > 0 ? gx - 1 : 0;
    *x1 = gx + 1 < p->cellCountX - 1 ? gx + 1 : p->cellCountX - 1;
    *y0 = gy - 1 > 0 ? gy - 1 : 0;
    *y1 = gy + 1 < p->cellCountY - 1 ? gy + 1 : p->cellCountY - 1;
}

/* clComputePressure, cl:135-181 */
void orc_next_pressure(const float* sorted4, float* pressure2, const uint32_t* start, const uint32_t* end, uint32_t n,
                       const orc_next_params* p)
{
#pragma omp parallel for schedule(static, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const float xi = sorted4[4 * (size_t)i], yi = sorted4[4 * (size_t)i + 1];
        int x0, x1, y0, y1;
        cell_window(xi, yi, p, &x0, &x1, &y0, &y1);
        float density = 0.0f;
        for (int x = x0; x <= x1; x++)
            for (int y = y0; y <= y1; y++) {
                const uint32_t c = (uint32_t)(y * p->cellCountX + x);
                const uint32_t s = start[c];
                if (s == NEXT_EMPTY) continue;
                const uint32_t e = end[c];
                for (uint32_t j = s; j < e; j++) {
                    if (j == i) continue;
                    const float dx = sorted4[4 * (size_t)j] - xi, dy = sorted4[4 * (size_t)j + 1] - yi;
                    density += orc_next_w(p->cellSize, sqrtf(dx * dx + dy * dy));
                }
            }
        float pr = 0.0f;
        if (density != 0.0f) {
            const float tait = (float)(0.87 * (double)(powf(density / 100.0f, 7.0f) - 1.0f));
            pr = tait / (density * density);
        }
        pressure2[2 * (size_t)i] = pr;
        pressure2[2 * (size_t)i + 1] = pr;
    }
}

/* clComputeForces, cl:183-238: the pressure and surface terms of the last neighbour visited (x outer,
   y inner, j ascending), plus 0.25 x surface minus gravity */
void orc_next_forces(const float* sorted4, const float* pressure2, float* force2, const uint32_t* start,
                     const uint32_t* end, uint32_t n, const orc_next_params* p)
{
#pragma omp parallel for schedule(static, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const float xi = sorted4[4 * (size_t)i], yi = sorted4[4 * (size_t)i + 1];
        const float pi = pressure2[2 * (size_t)i];
        int x0, x1, y0, y1;
        cell_window(xi, yi, p, &x0, &x1, &y0, &y1);
        float fpx = 0.0f, fpy = 0.0f, fsx = 0.0f, fsy = 0.0f;
        for (int x = x0; x <= x1; x++)
            for (int y = y0; y <= y1; y++) {
                const uint32_t c = (uint32_t)(y * p->cellCountX + x);
                const uint32_t s = start[c];
                if (s == NEXT_EMPTY) continue;
                const uint32_t e = end[c];
                for (uint32_t j = s; j < e; j++) {
                    if (j == i) continue;
                    const float rx = sorted4[4 * (size_t)j] - xi, ry = sorted4[4 * (size_t)j + 1] - yi;
                    const float r = sqrtf(rx * rx + ry * ry);
                    const float w = orc_next_w(p->cellSize, r);
                    const float gx = rx / r * w, gy = ry / r * w;
                    const float pj = pressure2[2 * (size_t)j];
                    const float k = -(pi + pj);
                    fpx = k * gx;
                    fpy = k * gy;
                    fsx = rx * w;
                    fsy = ry * w;
                }
            }
        force2[2 * (size_t)i] = ((fpx + 0.0f) + 0.25f * fsx) - 0.0f;
        force2[2 * (size_t)i + 1] = ((fpy + 0.0f) + 0.25f * fsy) - (float)9.81;
    }
}

/* clIntegrateForce, cl:251-263 */
void orc_next_integrate(float* state4, const float* force2, const uint32_t* index, uint32_t n,
                        const orc_next_params* p)
{
    for (uint32_t i = 0; i < n; i++) {
        float* s = state4 + 4 * (size_t)index[i];
        const float vx = s[2] + force2[2 * (size_t)i] * p->deltaTime;
        const float vy = s[3] + force2[2 * (size_t)i + 1] * p->deltaTime;
        const float x = s[0] + vx * p->deltaTime, y = s[1] + vy * p->deltaTime;
        s[0] = fminf(fmaxf(x, 1.0f), 1024.0f);
        s[1] = fminf(fmaxf(y, 1.0f), 768.0f);
        s[2] = vx;
        s[3] = vy;
    }
}

/* oclFluid3Dnext::compute, libCL/phys/oclFluid3Dnext.cpp:264-317 */
void orc_next_step(float* state4, uint32_t n, const orc_next_params* p, uint32_t* cell, uint32_t* index,
                   float* sorted4, uint32_t* start, uint32_t* end, float* pressure2, float* force2)
{
    const uint32_t ncells = (uint32_t)p->cellCountX * (uint32_t)p->cellCountY;
    orc_next_hash(state4, n, p, cell, index);
    orc_stable_sort(cell, index, n, 0, 24);  /* mRadixSort.compute(bfCell, bfIndex, 0, 24), cpp:272 */
    orc_next_reorder(index, state4, sorted4, n);
    orc_next_bounds(cell, n, ncells, start, end);
    orc_next_pressure(sorted4, pressure2, start, end, n, p);
    orc_next_forces(sorted4, pressure2, force2, start, end, n, p);
    orc_next_integrate(state4, force2, index, n, p);
}
