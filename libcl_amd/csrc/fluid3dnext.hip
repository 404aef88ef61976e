// oclFluid3Dnext (libCL/phys/oclFluid3Dnext.cpp:264-317, kernels libCL/phys/oclFluid3Dnext.cl:36-263) for
// gfx950: a 2D particle state (x, y, vx, vy) hashed into a dense cellCountX x cellCountY grid.
//
// Step: hash (+ clear of both cell tables, which no kernel between the hash and the bounds reads) ->
// 24-bit stable radix sort -> reorder + bounds in one gather -> density / Tait pressure over the 3 x 3
// neighbouring cells -> force + integration in one pass. The reference's force kernel assigns the terms
// of each neighbour instead of adding them (cl:221-223), so only the last neighbour visited (x outer, y
// inner, j ascending) contributes: the force pass walks the window backwards and stops at the first
// neighbour, one pair per particle instead of the window. Integration (clIntegrateForce, cl:251-263)
// touches only the caller slot index[i] of sorted particle i, so it runs in the same thread.
//
// Definitions the OpenCL sources leave open (the oracle uses the same): length(v) = sqrt(x x + y y);
// no contraction; 0.87 * (powf(...) - 1) in double then rounded; 0.25 and 9.81 rounded to float;
// clamp = fmin(fmax(x, lo), hi); cell indices outside the grid are not recorded in the bounds.
#include "fluid3dnext.hpp"

#include <cmath>
#include <cstring>
#include <string>

namespace mfl {
namespace {

constexpr uint32_t kBlock = 256;  // cLocalSize, cpp:16
constexpr uint32_t kNoCell = 0xFFFFFFFFu;

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// W, cl:116-133. POW2: h is a power of two, so q * (1 / h) equals q / h exactly and the pair loop
// multiplies instead of dividing.
template <bool POW2 = false>
__device__ inline float next_w(float h, float q, float inv_h = 0.f)
{
    float value = 0.f;
    q = POW2 ? q * inv_h : q / h;
    if (q < 1.f) {
        value = 1.0f - 1.5f * q * q + 0.75f * q * q * q;
    } else if (q < 2.f) {
        q = 2.0f - q;
        value = 0.25f * q * q * q;
    }
    return 2.0f / (3.0f * h) * value;
}

// h a power of two whose reciprocal is a normal float: q * (1 / h) == q / h for every q
inline bool reciprocal_exact(float h)
{
    int e = 0;
    return h > 0.f && std::isfinite(h) && std::frexp(h, &e) == 0.5f && e > -125 && e < 126;
}

struct Window {
    int x0, x1, y0, y1;
};

// the clamped 3 x 3 cell window of cl:141-148
__device__ inline Window window_of(float x, float y, const mfl_next_params& p)
{
    const int gx = (int)floorf(x / p.cellSize), gy = (int)floorf(y / p.cellSize);
    return Window{max(gx - 1, 0), min(gx + 1, p.cellCountX - 1), max(gy - 1, 0), min(gy + 1, p.cellCountY - 1)};
}

// clHash (cl:36-44) + clInitBounds (cl:53-58) over all cells, and the sort's first-pass digit counts
// (RadixSort::first_pass_plan): block b hashes the keys of sort tile xcd_tile(), ITEMS per thread at
// tile offsets k * 256 + thread, and histograms their low `bits` bits in LDS. The sort generates the
// identity index itself (cl:43). The grid is one block per tile.
template <int ITEMS>
__global__ __launch_bounds__(kBlock) void k_next_hash_counts(const float4* __restrict__ state, uint32_t n,
                                                             mfl_next_params p, uint32_t* __restrict__ cell,
                                                             uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                             uint32_t ncells, uint32_t* __restrict__ counts,
                                                             uint32_t ntiles, uint32_t bits)
{
    __shared__ uint32_t hist[512];
    const uint32_t radix = 1u << bits, mask = radix - 1u;
    for (uint32_t d = threadIdx.x; d < radix; d += kBlock) hist[d] = 0;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < ncells; c += gridDim.x * kBlock) {
        start[c] = kNoCell;
        end[c] = kNoCell;
    }
    __syncthreads();
    const uint32_t tile = xcd_tile();
    const size_t base = (size_t)tile * (kBlock * ITEMS);
    // every load issued before the first count (a clamped index keeps the loads unconditional)
    float4 sv[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t g = base + (size_t)k * kBlock + threadIdx.x;
        sv[k] = state[g < n ? g : n - 1];
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t g = base + (size_t)k * kBlock + threadIdx.x;
        if (g < n) {
            const float4 st = sv[k];
            const int gx = (int)floorf(st.x / p.cellSize), gy = (int)floorf(st.y / p.cellSize);
            const uint32_t key = (uint32_t)gy * (uint32_t)p.cellCountX + (uint32_t)gx;
            cell[g] = key;
            atomicAdd(&hist[key & mask], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < radix; d += kBlock) counts[(size_t)d * ntiles + tile] = hist[d];
}

// clReorder (cl:46-51) + clFindBounds (cl:60-92)
__global__ __launch_bounds__(kBlock) void k_next_reorder_bounds(const uint32_t* __restrict__ index,
                                                               const float4* __restrict__ state,
                                                               float4* __restrict__ sorted,
                                                               const uint32_t* __restrict__ cell, uint32_t n,
                                                               uint32_t* __restrict__ start,
                                                               uint32_t* __restrict__ end, uint32_t ncells)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    sorted[i] = state[index[i]];
    const uint32_t h = cell[i];
    if (i == 0) {
        if (h < ncells) start[h] = 0;
        return;
    }
    const uint32_t prev = cell[i - 1];
    if (h != prev) {
        if (prev < ncells) end[prev] = i;
        if (h < ncells) start[h] = i;
    }
    if (i == n - 1 && h < ncells) end[h] = n;
}

// clComputePressure, cl:135-181
template <bool POW2>
__global__ __launch_bounds__(kBlock) void k_next_pressure(const float4* __restrict__ sorted,
                                                         float2* __restrict__ pressure,
                                                         const uint32_t* __restrict__ start,
                                                         const uint32_t* __restrict__ end, uint32_t n,
                                                         mfl_next_params p)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 si = sorted[i];
    const Window w = window_of(si.x, si.y, p);
    const float inv_h = 1.0f / p.cellSize;
    float density = 0.f;
    for (int x = w.x0; x <= w.x1; x++)
        for (int y = w.y0; y <= w.y1; y++) {
            const uint32_t c = (uint32_t)(y * p.cellCountX + x);
            const uint32_t s = start[c];
            if (s == kNoCell) continue;
            const uint32_t e = end[c];
            for (uint32_t j = s; j < e; j++) {
                if (j == i) continue;
                const float4 sj = sorted[j];
                const float dx = sj.x - si.x, dy = sj.y - si.y;
                density += next_w<POW2>(p.cellSize, sqrtf(dx * dx + dy * dy), inv_h);
            }
        }
    float pr = 0.f;
    if (density != 0.f) {
        const float tait = (float)(0.87 * (double)(powf(density / 100.0f, 7.0f) - 1.0f));
        pr = tait / (density * density);
    }
    pressure[i] = make_float2(pr, pr);
}

// clComputeForces (cl:183-238) + clIntegrateForce (cl:251-263)
__global__ __launch_bounds__(kBlock) void k_next_forces_integrate(const float4* __restrict__ sorted,
                                                                 const float2* __restrict__ pressure,
                                                                 float2* __restrict__ force,
                                                                 const uint32_t* __restrict__ index,
                                                                 float4* __restrict__ state,
                                                                 const uint32_t* __restrict__ start,
                                                                 const uint32_t* __restrict__ end, uint32_t n,
                                                                 mfl_next_params p)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 si = sorted[i];
    const float pi = pressure[i].x;
    const Window w = window_of(si.x, si.y, p);
    float fpx = 0.f, fpy = 0.f, fsx = 0.f, fsy = 0.f;
    bool found = false;
    // the last neighbour of the forward visit order is the first one walking backwards
    for (int x = w.x1; x >= w.x0 && !found; x--)
        for (int y = w.y1; y >= w.y0 && !found; y--) {
            const uint32_t c = (uint32_t)(y * p.cellCountX + x);
            const uint32_t s = start[c];
            if (s == kNoCell) continue;
            const uint32_t e = end[c];
            for (uint32_t j = e; j-- > s;) {
                if (j == i) continue;
                const float4 sj = sorted[j];
                const float rx = sj.x - si.x, ry = sj.y - si.y;
                const float r = sqrtf(rx * rx + ry * ry);
                const float wv = next_w(p.cellSize, r);
                const float gx = rx / r * wv, gy = ry / r * wv;
                const float k = -(pi + pressure[j].x);
                fpx = k * gx;
                fpy = k * gy;
                fsx = rx * wv;
                fsy = ry * wv;
                found = true;
                break;
            }
        }
    const float2 f = make_float2(((fpx + 0.0f) + 0.25f * fsx) - 0.0f, ((fpy + 0.0f) + 0.25f * fsy) - 9.81f);
    force[i] = f;
    const uint32_t o = index[i];
    float4 s = state[o];
    const float vx = s.z + f.x * p.deltaTime, vy = s.w + f.y * p.deltaTime;
    s.x = fminf(fmaxf(s.x + vx * p.deltaTime, 1.0f), 1024.0f);
    s.y = fminf(fmaxf(s.y + vy * p.deltaTime, 1.0f), 768.0f);
    s.z = vx;
    s.w = vy;
    state[o] = s;
}

}  // namespace

Fluid3DNext::Fluid3DNext() { mfl_next_params_default(&params); }

bool Fluid3DNext::fit_tables()
{
    if (params.cellCountX <= 0 || params.cellCountY <= 0 ||
        (uint64_t)params.cellCountX * (uint64_t)params.cellCountY > (1u << 30)) {
        set_error("cellCountX x cellCountY must be positive and at most 2^30");
        return false;
    }
    ncells_ = (uint32_t)params.cellCountX * (uint32_t)params.cellCountY;
    return start_.fit(ncells_ * 4ull) && end_.fit(ncells_ * 4ull);
}

bool Fluid3DNext::allocate_internal()
{
    const size_t n = n_;
    if (state_.owned && !state_.fit(n * 16)) return false;
    return sorted_.fit(n * 16) && cell_.fit(n * 4) && index_.fit(n * 4) && pressure_.fit(n * 8) &&
           force_.fit(n * 8) && params_buf_.fit(sizeof(mfl_next_params)) && sort_.reserve(n, 24);
}

int Fluid3DNext::compile()
{
    if (!sort_.compile()) return 0;
    mfl_next_params_default(&params);  // clInitFluid, cl:95-113
    if (!fit_tables() || !allocate_internal()) return 0;
    MFL_HIP(hipMemcpy(params_buf_.ptr, &params, sizeof(params), hipMemcpyHostToDevice));
    uploaded_ = params;
    params_synced_ = true;
    compiled_ = true;
    return 1;
}

int Fluid3DNext::set_particle_count(size_t n)
{
    if (n > 0xFFFFFFFFull / 16) {
        set_error("particle count too large");
        return 0;
    }
    if (!state_.owned && state_.bytes < n * 16) {
        set_error("the bound position buffer holds fewer than " + std::to_string(n) + " particles");
        return 0;
    }
    n_ = n;
    return allocate_internal() ? 1 : 0;
}

DeviceBuffer* Fluid3DNext::select(int kind)
{
    switch (kind) {
    case MFL_NEXT_POSITION: return &state_;
    case MFL_NEXT_SORTED_STATE: return &sorted_;
    case MFL_NEXT_INDEX: return &index_;
    case MFL_NEXT_CELL: return &cell_;
    case MFL_NEXT_CELL_START: return &start_;
    case MFL_NEXT_CELL_END: return &end_;
    case MFL_NEXT_PRESSURE: return &pressure_;
    case MFL_NEXT_FORCE: return &force_;
    case MFL_NEXT_PARAMS: return &params_buf_;
    default: set_error("unknown buffer kind " + std::to_string(kind)); return nullptr;
    }
}

int Fluid3DNext::set_buffer(int kind, void* ptr, size_t bytes)
{
    if (kind != MFL_NEXT_POSITION) {
        set_error("only the position (state) buffer can be replaced");
        return 0;
    }
    if (!ptr) {
        set_error("null buffer");
        return 0;
    }
    if (bytes < n_ * 16) {
        set_error("buffer of " + std::to_string(bytes) + " bytes is smaller than " + std::to_string(n_) +
                  " float4 particles");
        return 0;
    }
    state_.release();  // frees a program-owned buffer only (deleteBuffer, oclFluid3Dnext.h:113-119)
    state_.ptr = ptr;
    state_.bytes = bytes;
    state_.owned = false;
    return 1;
}

int Fluid3DNext::get_buffer(int kind, void** ptr, size_t* bytes)
{
    DeviceBuffer* b = select(kind);
    if (!b) return 0;
    if (ptr) *ptr = b->ptr;
    if (bytes) *bytes = kind == MFL_NEXT_PARAMS ? sizeof(mfl_next_params) : b->bytes;
    return 1;
}

int Fluid3DNext::write_buffer(int kind, const void* host, size_t bytes)
{
    DeviceBuffer* b = select(kind);
    if (!b) return 0;
    if (bytes > b->bytes) {
        set_error("write of " + std::to_string(bytes) + " bytes exceeds the buffer");
        return 0;
    }
    MFL_HIP(hipMemcpy(b->ptr, host, bytes, hipMemcpyHostToDevice));
    return 1;
}

int Fluid3DNext::read_buffer(int kind, void* host, size_t bytes)
{
    DeviceBuffer* b = select(kind);
    if (!b) return 0;
    if (bytes > b->bytes) {
        set_error("read of " + std::to_string(bytes) + " bytes exceeds the buffer");
        return 0;
    }
    MFL_HIP(hipDeviceSynchronize());
    MFL_HIP(hipMemcpy(host, b->ptr, bytes, hipMemcpyDeviceToHost));
    return 1;
}

int Fluid3DNext::compute(hipStream_t s)
{
    if (!compiled_) {
        set_error("oclFluid3Dnext: compile() first");
        return 0;
    }
    if (n_ == 0 || n_ % kBlock != 0) {
        set_error("particle count " + std::to_string(n_) + " is not a non-zero multiple of 256 (cLocalSize)");
        return 0;
    }
    if (!(params.cellSize > 0.f)) {
        set_error("cellSize must be positive");
        return 0;
    }
    if (!fit_tables()) return 0;  // Params may have changed the grid since compile()
    const uint32_t n = (uint32_t)n_;
    auto* state = state_.as<float4>();
    auto* sorted = sorted_.as<float4>();
    auto* cell = cell_.as<uint32_t>();
    auto* index = index_.as<uint32_t>();
    auto* start = start_.as<uint32_t>();
    auto* end = end_.as<uint32_t>();
    if (!params_synced_ || std::memcmp(&params, &uploaded_, sizeof(params)) != 0) {  // bfParams, on change only
        uploaded_ = params;
        MFL_HIP(hipMemcpyAsync(params_buf_.ptr, &uploaded_, sizeof(params), hipMemcpyHostToDevice, s));
        params_synced_ = true;
    }
    // the hash kernel also counts the sort's first digit per tile (one count launch fewer)
    SortCountPlan plan;
    if (!sort_.first_pass_plan(n, 0, 24, &plan)) return 0;
#define MFL_NEXT_HASH(IT)                                                                                         \
    hipLaunchKernelGGL(k_next_hash_counts<IT>, dim3(plan.ntiles), dim3(kBlock), 0, s, state, n, params, cell, start, \
                       end, ncells_, plan.counts, plan.ntiles, plan.bits)
    switch (plan.items) {
    case 16: MFL_NEXT_HASH(16); break;
    case 8: MFL_NEXT_HASH(8); break;
    default: MFL_NEXT_HASH(4); break;
    }
#undef MFL_NEXT_HASH
    MFL_HIP(hipGetLastError());
    if (!sort_.sort_bits(cell, index, n, 0, 24, s, true, true)) return 0;  // mRadixSort.compute, cpp:272
    hipLaunchKernelGGL(k_next_reorder_bounds, dim3(blocks_for(n)), dim3(kBlock), 0, s, index, state, sorted, cell, n,
                       start, end, ncells_);
    if (reciprocal_exact(params.cellSize))
        hipLaunchKernelGGL(k_next_pressure<true>, dim3(blocks_for(n)), dim3(kBlock), 0, s, sorted, pressure_.as<float2>(),
                           start, end, n, params);
    else
        hipLaunchKernelGGL(k_next_pressure<false>, dim3(blocks_for(n)), dim3(kBlock), 0, s, sorted, pressure_.as<float2>(),
                           start, end, n, params);
    hipLaunchKernelGGL(k_next_forces_integrate, dim3(blocks_for(n)), dim3(kBlock), 0, s, sorted,
                       pressure_.as<float2>(), force_.as<float2>(), index, state, start, end, n, params);
    MFL_HIP(hipGetLastError());
    return 1;
}

}  // namespace mfl
